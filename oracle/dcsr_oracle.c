/*
 * dcsr_oracle.c -- CPU restatement of the reference's host (cusp 0.5.1 /
 * thrust 1.8.1 "sequential" backend) path for the Dynamic-CSR operations.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * never by the product library (dcsr_amd/csrc), which has no CPU fallback.
 *
 * Every function restates one reference routine, cited file:line (paths are
 * relative to the reference checkout).  Semantics that matter for bit-exact
 * structure are reproduced exactly: FP32 accumulation order, exact-zero drop,
 * stable ordering, and the output order inside each row.  Build with
 * -ffp-contract=off (no fused multiply-add), as the reference host path
 * compiled for x86-64 has no FMA.
 *
 * Pinning (how each function is checked, see tests/test_oracle_*.py):
 *   orc_sort_by_row        vs oracle/_ref (cusp::sort_by_row) + brute force
 *   orc_indices_to_offsets vs oracle/_ref + closed form (bincount/cumsum)
 *   orc_csr_add            vs oracle/_ref (cusp::add) bit-exact incl. order,
 *                          vs a dictionary-of-keys brute force on small inputs
 *   orc_csr_spmv           vs oracle/_ref bit-exact, vs scipy FP64 within tol
 *   orc_spgemm_*           vs oracle/_ref bit-exact incl. order, vs scipy
 *                          structure on data without cancellation
 *   orc_csr_to_hyb         vs oracle/_ref (generic csr->hyb conversion)
 *   orc_update             composition of the above (sort, offsets, add)
 *   orc_spmm_dense         no reference counterpart: defined as k column-wise
 *                          orc_csr_spmv calls; pinned against orc_csr_spmv and
 *                          scipy FP64.
 *
 *   orc_hyb_spmv           vs oracle/_ref (cusp::multiply on a host
 *                          hyb_matrix) bit-exact, vs scipy FP64 within tol
 *   orc_hyb_update         restates the reference's device-only append
 *                          (no cusp host counterpart); pinned by a per-row
 *                          list brute force and by the linear-operator
 *                          identity HYB(A) + B == A + B (duplicates summed)
 *
 * Checker utilities (no reference counterpart; they compare, they do not
 * restate):
 *   orc_sort_rows_by_col   stable per-row sort by column (the comparison
 *                          form); pinned against numpy's stable lexsort
 *   orc_spgemm_check_f64   |C - A*B| <= rtol (|A||B|) per stored entry of C,
 *                          with A*B and |A||B| accumulated in FP64; pinned
 *                          against scipy FP64 products and planted errors
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int32_t idx_t;

/* cusp::sort_by_row -> cusp::counting_sort_by_key (stable counting sort on the
 * row key carrying a permutation), then gather of (col, val) through the
 * permutation.  cusp/detail/sort.inl:39-70, sequential/sort.h:80-124.
 * min_row is forced to 0 (sort.h:100) and max_row = max element. */
int orc_sort_by_row(int64_t n, idx_t *rows, idx_t *cols, float *vals)
{
    if (n <= 0) return 0;
    idx_t maxr = rows[0];
    for (int64_t i = 0; i < n; ++i) {
        if (rows[i] < 0) return -1;
        if (rows[i] > maxr) maxr = rows[i];
    }
    size_t nb = (size_t)maxr + 2;
    size_t *counts = (size_t *)calloc(nb, sizeof(size_t));
    idx_t *tk = (idx_t *)malloc(n * sizeof(idx_t));
    idx_t *tc = (idx_t *)malloc(n * sizeof(idx_t));
    float *tv = (float *)malloc(n * sizeof(float));
    if (!counts || !tk || !tc || !tv) { free(counts); free(tk); free(tc); free(tv); return -2; }
    memcpy(tk, rows, n * sizeof(idx_t));
    memcpy(tc, cols, n * sizeof(idx_t));
    memcpy(tv, vals, n * sizeof(float));
    for (int64_t i = 0; i < n; ++i) counts[tk[i] + 1]++;
    for (size_t i = 0; i + 1 < nb; ++i) counts[i + 1] += counts[i];
    for (int64_t i = 0; i < n; ++i) {
        size_t p = counts[tk[i]]++;
        rows[p] = tk[i];
        cols[p] = tc[i];
        vals[p] = tv[i];
    }
    free(counts); free(tk); free(tc); free(tv);
    return 0;
}

/* cusp::indices_to_offsets = thrust::lower_bound of 0..R over the (sorted)
 * index array.  cusp/detail/format_utils.inl:188-200. */
void orc_indices_to_offsets(int64_t n, const idx_t *rows, int64_t R, idx_t *offsets)
{
    for (int64_t r = 0; r <= R; ++r) {
        int64_t lo = 0, hi = n;
        while (lo < hi) {
            int64_t mid = lo + (hi - lo) / 2;
            if (rows[mid] < r) lo = mid + 1; else hi = mid;
        }
        offsets[r] = (idx_t)lo;
    }
}

/* cusp::add on host CSR -> sequential elementwise with thrust::plus.
 * cusp/detail/elementwise.inl:107-116, sequential/elementwise.h:44-131.
 * Per row: A entries accumulate into A_row[j] (from 0, stored order), B
 * entries into B_row[j]; each touched column is visited through the linked
 * list (most recently first-touched column first) and emitted when
 * A_row[j] + B_row[j] != 0.  Returns nnz of C. */
int64_t orc_csr_add(int64_t R, int64_t C,
                    const idx_t *Aoff, const idx_t *Acol, const float *Aval,
                    const idx_t *Boff, const idx_t *Bcol, const float *Bval,
                    idx_t *Coff, idx_t *Ccol, float *Cval)
{
    idx_t *next = (idx_t *)malloc((C > 0 ? C : 1) * sizeof(idx_t));
    float *arow = (float *)calloc(C > 0 ? C : 1, sizeof(float));
    float *brow = (float *)calloc(C > 0 ? C : 1, sizeof(float));
    if (!next || !arow || !brow) { free(next); free(arow); free(brow); return -1; }
    for (int64_t j = 0; j < C; ++j) next[j] = -1;
    int64_t nnz = 0;
    Coff[0] = 0;
    for (int64_t i = 0; i < R; ++i) {
        idx_t head = -2, length = 0;
        for (idx_t jj = Aoff[i]; jj < Aoff[i + 1]; ++jj) {
            idx_t j = Acol[jj];
            arow[j] += Aval[jj];
            if (next[j] == -1) { next[j] = head; head = j; length++; }
        }
        for (idx_t jj = Boff[i]; jj < Boff[i + 1]; ++jj) {
            idx_t j = Bcol[jj];
            brow[j] += Bval[jj];
            if (next[j] == -1) { next[j] = head; head = j; length++; }
        }
        for (idx_t jj = 0; jj < length; ++jj) {
            float result = arow[head] + brow[head];
            if (result != 0.0f) {
                Ccol[nnz] = head;
                Cval[nnz] = result;
                nnz++;
            }
            idx_t prev = head;
            head = next[head];
            next[prev] = -1;
            arow[prev] = 0.0f;
            brow[prev] = 0.0f;
        }
        Coff[i + 1] = (idx_t)nnz;
    }
    free(next); free(arow); free(brow);
    return nnz;
}

/* cusp::multiply(csr, x, y) on host -> sequential csr_spmv with
 * initialize = zero, combine = multiplies, reduce = plus.
 * sequential/multiply/csr_spmv.h:41-73, generic/multiply.inl:70-72. */
void orc_csr_spmv(int64_t R, const idx_t *off, const idx_t *col, const float *val,
                  const float *x, float *y)
{
    for (int64_t i = 0; i < R; ++i) {
        float acc = 0.0f;
        for (idx_t jj = off[i]; jj < off[i + 1]; ++jj) {
            float p = val[jj] * x[col[jj]];
            acc = acc + p;
        }
        y[i] = acc;
    }
}

/* Dense multi-RHS product Y = A X (X row-major C x k with leading dim ldx,
 * Y row-major R x k with ldy).  No reference counterpart (the reference has
 * SpMV only); defined as k independent orc_csr_spmv column products. */
void orc_spmm_dense(int64_t R, const idx_t *off, const idx_t *col, const float *val,
                    int64_t k, const float *X, int64_t ldx, float *Y, int64_t ldy)
{
    for (int64_t i = 0; i < R; ++i) {
        for (int64_t c = 0; c < k; ++c) {
            float acc = 0.0f;
            for (idx_t jj = off[i]; jj < off[i + 1]; ++jj) {
                float p = val[jj] * X[(int64_t)col[jj] * ldx + c];
                acc = acc + p;
            }
            Y[i * ldy + c] = acc;
        }
    }
}

/* Symbolic pass of host CSR SpGEMM: number of distinct (i,k) with a partial
 * product, explicit zeros included.  sequential/multiply/csr_spgemm.h:38-73. */
int64_t orc_spgemm_count(int64_t R, int64_t Cb,
                         const idx_t *Aoff, const idx_t *Acol,
                         const idx_t *Boff, const idx_t *Bcol)
{
    int64_t *mask = (int64_t *)malloc((Cb > 0 ? Cb : 1) * sizeof(int64_t));
    if (!mask) return -1;
    for (int64_t k = 0; k < Cb; ++k) mask[k] = -1;
    int64_t nnz = 0;
    for (int64_t i = 0; i < R; ++i)
        for (idx_t jj = Aoff[i]; jj < Aoff[i + 1]; ++jj) {
            idx_t j = Acol[jj];
            for (idx_t kk = Boff[j]; kk < Boff[j + 1]; ++kk) {
                idx_t k = Bcol[kk];
                if (mask[k] != i) { mask[k] = i; nnz++; }
            }
        }
    free(mask);
    return nnz;
}

/* Numeric pass: Gustavson with a linked list; sums[k] += a_ij * b_jk in A-row
 * then B-row stored order; exact zeros dropped; row output in reverse
 * first-touch order.  sequential/multiply/csr_spgemm.h:75-150. */
int64_t orc_spgemm(int64_t R, int64_t Cb,
                   const idx_t *Aoff, const idx_t *Acol, const float *Aval,
                   const idx_t *Boff, const idx_t *Bcol, const float *Bval,
                   idx_t *Coff, idx_t *Ccol, float *Cval)
{
    idx_t *next = (idx_t *)malloc((Cb > 0 ? Cb : 1) * sizeof(idx_t));
    float *sums = (float *)calloc(Cb > 0 ? Cb : 1, sizeof(float));
    if (!next || !sums) { free(next); free(sums); return -1; }
    for (int64_t k = 0; k < Cb; ++k) next[k] = -1;
    int64_t nnz = 0;
    Coff[0] = 0;
    for (int64_t i = 0; i < R; ++i) {
        idx_t head = -2, length = 0;
        for (idx_t jj = Aoff[i]; jj < Aoff[i + 1]; ++jj) {
            idx_t j = Acol[jj];
            float v = Aval[jj];
            for (idx_t kk = Boff[j]; kk < Boff[j + 1]; ++kk) {
                idx_t k = Bcol[kk];
                float p = v * Bval[kk];
                sums[k] = sums[k] + p;
                if (next[k] == -1) { next[k] = head; head = k; length++; }
            }
        }
        for (idx_t jj = 0; jj < length; ++jj) {
            if (sums[head] != 0.0f) {
                Ccol[nnz] = head;
                Cval[nnz] = sums[head];
                nnz++;
            }
            idx_t t = head;
            head = next[head];
            next[t] = -1;
            sums[t] = 0.0f;
        }
        Coff[i + 1] = (idx_t)nnz;
    }
    free(next); free(sums);
    return nnz;
}

/* Host CSR -> HYB with a given ELL width: the first `width` entries of each
 * row (stored order) go to the column-major ELL block (index k*pitch + row,
 * padding column -1 / value 0), the rest go to the COO part in CSR order.
 * generic/conversions/csr_to_other.h:226-303.  Returns the COO entry count. */
int64_t orc_csr_to_hyb(int64_t R, const idx_t *off, const idx_t *col, const float *val,
                       int64_t width, int64_t pitch,
                       idx_t *ell_col, float *ell_val,
                       idx_t *coo_row, idx_t *coo_col, float *coo_val)
{
    for (int64_t t = 0; t < width * pitch; ++t) { ell_col[t] = -1; ell_val[t] = 0.0f; }
    int64_t nc = 0;
    for (int64_t i = 0; i < R; ++i) {
        for (idx_t jj = off[i]; jj < off[i + 1]; ++jj) {
            int64_t k = jj - off[i];
            if (k < width) {
                ell_col[k * pitch + i] = col[jj];
                ell_val[k * pitch + i] = val[jj];
            } else {
                coo_row[nc] = (idx_t)i;
                coo_col[nc] = col[jj];
                coo_val[nc] = val[jj];
                nc++;
            }
        }
    }
    return nc;
}

/* COO -> CSR on host: cusp::sort_by_row (the drivers sort before loading,
 * src/Fill_Test.inl:90) then generic coo->csr conversion =
 * indices_to_offsets + copies (generic/conversions/coo_to_other.h:118-125).
 * Duplicates and explicit zeros are kept.  rows/cols/vals are sorted in place;
 * the resulting CSR column/value arrays are cols/vals. */
int orc_coo_to_csr(int64_t R, int64_t n, idx_t *rows, idx_t *cols, float *vals, idx_t *off)
{
    int rc = orc_sort_by_row(n, rows, cols, vals);
    if (rc) return rc;
    orc_indices_to_offsets(n, rows, R, off);
    return 0;
}

/* One in-place update batch on the CPU path: the batch triples are sorted by
 * row (cusp::sort_by_row), converted to CSR B, and C = cusp::add(A, B).  This
 * is the reference driver's CSR update step (src/Fill_Test.inl:228,240-258)
 * on host matrices.  The batch arrays are copied (the caller's are left
 * untouched).  C arrays must hold nnz(A) + n entries.  Returns nnz(C). */
int64_t orc_update(int64_t R, int64_t C,
                   const idx_t *Aoff, const idx_t *Acol, const float *Aval,
                   int64_t n, const idx_t *brow, const idx_t *bcol, const float *bval,
                   idx_t *Coff, idx_t *Ccol, float *Cval)
{
    idx_t *r = (idx_t *)malloc((n > 0 ? n : 1) * sizeof(idx_t));
    idx_t *c = (idx_t *)malloc((n > 0 ? n : 1) * sizeof(idx_t));
    float *v = (float *)malloc((n > 0 ? n : 1) * sizeof(float));
    idx_t *boff = (idx_t *)malloc((R + 1) * sizeof(idx_t));
    if (!r || !c || !v || !boff) { free(r); free(c); free(v); free(boff); return -1; }
    if (n > 0) {
        memcpy(r, brow, n * sizeof(idx_t));
        memcpy(c, bcol, n * sizeof(idx_t));
        memcpy(v, bval, n * sizeof(float));
    }
    int64_t nnz;
    if (orc_coo_to_csr(R, n, r, c, v, boff) != 0) nnz = -2;
    else nnz = orc_csr_add(R, C, Aoff, Acol, Aval, boff, c, v, Coff, Ccol, Cval);
    free(r); free(c); free(v); free(boff);
    return nnz;
}

/* orc_update on `threads` OpenMP threads, for the multi-core CPU baseline:
 * the same batch sort (a stable counting sort by row, done as per-thread
 * counts over contiguous chunks, an exclusive scan in (row, chunk) order and
 * a scatter: the order of equal rows is the input order, as the sequential
 * counting sort gives) and the same per-row cusp::add computation, with the
 * rows split into contiguous ranges, one per thread, each written to a
 * private buffer and then copied into place.  Every row's entries and their
 * order equal orc_update's (pinned bit-exactly against it).  Returns nnz(C),
 * or a negative value on a bad row index or an allocation failure. */
#ifdef _OPENMP
#include <omp.h>
#endif
int64_t orc_update_mt(int64_t R, int64_t C,
                      const idx_t *Aoff, const idx_t *Acol, const float *Aval,
                      int64_t n, const idx_t *brow, const idx_t *bcol, const float *bval,
                      idx_t *Coff, idx_t *Ccol, float *Cval, int threads)
{
    if (threads < 1) threads = 1;
    const int T = threads;
    idx_t *r = (idx_t *)malloc((n > 0 ? n : 1) * sizeof(idx_t));
    idx_t *c = (idx_t *)malloc((n > 0 ? n : 1) * sizeof(idx_t));
    float *v = (float *)malloc((n > 0 ? n : 1) * sizeof(float));
    idx_t *boff = (idx_t *)calloc(R + 1, sizeof(idx_t));
    int64_t *cnt = (int64_t *)calloc((size_t)T * (R + 1), sizeof(int64_t));
    int64_t *rown = (int64_t *)calloc(R + 1, sizeof(int64_t));
    int bad = 0;
    if (!r || !c || !v || !boff || !cnt || !rown) {
        free(r); free(c); free(v); free(boff); free(cnt); free(rown);
        return -1;
    }
    /* counting sort by row: chunk t counts its rows, then positions */
    #pragma omp parallel num_threads(T) reduction(|:bad)
    {
#ifdef _OPENMP
        const int t = omp_get_thread_num();
#else
        const int t = 0;
#endif
        const int64_t i0 = n * t / T, i1 = n * (t + 1) / T;
        int64_t *ct = cnt + (size_t)t * (R + 1);
        for (int64_t i = i0; i < i1; ++i) {
            if (brow[i] < 0 || brow[i] >= R) { bad = 1; continue; }
            ct[brow[i]]++;
        }
    }
    if (bad) { free(r); free(c); free(v); free(boff); free(cnt); free(rown); return -2; }
    {
        int64_t run = 0;
        for (int64_t row = 0; row < R; ++row) {
            boff[row] = (idx_t)run;
            for (int t = 0; t < T; ++t) {
                int64_t *ct = cnt + (size_t)t * (R + 1);
                const int64_t k = ct[row];
                ct[row] = run;
                run += k;
            }
        }
        boff[R] = (idx_t)run;
    }
    #pragma omp parallel num_threads(T)
    {
#ifdef _OPENMP
        const int t = omp_get_thread_num();
#else
        const int t = 0;
#endif
        const int64_t i0 = n * t / T, i1 = n * (t + 1) / T;
        int64_t *ct = cnt + (size_t)t * (R + 1);
        for (int64_t i = i0; i < i1; ++i) {
            const int64_t p = ct[brow[i]]++;
            r[p] = brow[i];
            c[p] = bcol[i];
            v[p] = bval[i];
        }
    }
    /* per-row add, rows split into contiguous ranges */
    idx_t **bufc = (idx_t **)calloc(T, sizeof(idx_t *));
    float **bufv = (float **)calloc(T, sizeof(float *));
    int fail = 0;
    #pragma omp parallel num_threads(T) reduction(|:fail)
    {
#ifdef _OPENMP
        const int t = omp_get_thread_num();
#else
        const int t = 0;
#endif
        const int64_t r0 = R * t / T, r1 = R * (t + 1) / T;
        const int64_t cap = (int64_t)(Aoff[r1] - Aoff[r0]) + (int64_t)(boff[r1] - boff[r0]);
        idx_t *next = (idx_t *)malloc((C > 0 ? C : 1) * sizeof(idx_t));
        float *arow = (float *)calloc(C > 0 ? C : 1, sizeof(float));
        float *brw = (float *)calloc(C > 0 ? C : 1, sizeof(float));
        idx_t *oc = (idx_t *)malloc((cap > 0 ? cap : 1) * sizeof(idx_t));
        float *ov = (float *)malloc((cap > 0 ? cap : 1) * sizeof(float));
        if (!next || !arow || !brw || !oc || !ov) {
            fail = 1;
        } else {
            for (int64_t j = 0; j < C; ++j) next[j] = -1;
            int64_t k = 0;
            for (int64_t i = r0; i < r1; ++i) {
                idx_t head = -2, length = 0;
                for (idx_t jj = Aoff[i]; jj < Aoff[i + 1]; ++jj) {
                    idx_t j = Acol[jj];
                    arow[j] += Aval[jj];
                    if (next[j] == -1) { next[j] = head; head = j; length++; }
                }
                for (idx_t jj = boff[i]; jj < boff[i + 1]; ++jj) {
                    idx_t j = c[jj];
                    brw[j] += v[jj];
                    if (next[j] == -1) { next[j] = head; head = j; length++; }
                }
                const int64_t k0 = k;
                for (idx_t jj = 0; jj < length; ++jj) {
                    float result = arow[head] + brw[head];
                    if (result != 0.0f) {
                        oc[k] = head;
                        ov[k] = result;
                        k++;
                    }
                    idx_t prev = head;
                    head = next[head];
                    next[prev] = -1;
                    arow[prev] = 0.0f;
                    brw[prev] = 0.0f;
                }
                rown[i] = k - k0;
            }
        }
        free(next); free(arow); free(brw);
        bufc[t] = oc;
        bufv[t] = ov;
    }
    int64_t nnz = -1;
    if (!fail) {
        int64_t run = 0;
        Coff[0] = 0;
        for (int64_t i = 0; i < R; ++i) {
            run += rown[i];
            Coff[i + 1] = (idx_t)run;
        }
        nnz = run;
        #pragma omp parallel num_threads(T)
        {
#ifdef _OPENMP
            const int t = omp_get_thread_num();
#else
            const int t = 0;
#endif
            const int64_t r0 = R * t / T, r1 = R * (t + 1) / T;
            const int64_t base = Coff[r0], len = (int64_t)Coff[r1] - base;
            if (len > 0) {
                memcpy(Ccol + base, bufc[t], (size_t)len * sizeof(idx_t));
                memcpy(Cval + base, bufv[t], (size_t)len * sizeof(float));
            }
        }
    }
    for (int t = 0; t < T; ++t) { free(bufc[t]); free(bufv[t]); }
    free(bufc); free(bufv);
    free(r); free(c); free(v); free(boff); free(cnt); free(rown);
    return nnz;
}

/* orc_csr_spmv with rows split over `threads` OpenMP threads (each row is
 * summed in stored order as there: the same result bit for bit). */
void orc_csr_spmv_mt(int64_t R, const idx_t *off, const idx_t *col, const float *val, const float *x, float *y,
                     int threads)
{
    #pragma omp parallel for num_threads(threads < 1 ? 1 : threads) schedule(static)
    for (int64_t i = 0; i < R; ++i) {
        float s = 0.0f;
        for (idx_t jj = off[i]; jj < off[i + 1]; ++jj) s = s + val[jj] * x[col[jj]];
        y[i] = s;
    }
}

/* Stable sort of each row's (col, val) by column, in place (insertion sort
 * for short rows, a stable merge sort through scratch for long ones). */
static void merge_by_col(idx_t *c, float *v, idx_t *tc, float *tv, int64_t n)
{
    if (n <= 24) {
        for (int64_t i = 1; i < n; ++i) {
            idx_t kc = c[i];
            float kv = v[i];
            int64_t j = i - 1;
            while (j >= 0 && c[j] > kc) { c[j + 1] = c[j]; v[j + 1] = v[j]; --j; }
            c[j + 1] = kc;
            v[j + 1] = kv;
        }
        return;
    }
    int64_t h = n / 2;
    merge_by_col(c, v, tc, tv, h);
    merge_by_col(c + h, v + h, tc, tv, n - h);
    int64_t i = 0, j = h, k = 0;
    while (i < h && j < n) {
        if (c[j] < c[i]) { tc[k] = c[j]; tv[k] = v[j]; ++j; }
        else { tc[k] = c[i]; tv[k] = v[i]; ++i; }
        ++k;
    }
    while (i < h) { tc[k] = c[i]; tv[k] = v[i]; ++i; ++k; }
    while (j < n) { tc[k] = c[j]; tv[k] = v[j]; ++j; ++k; }
    memcpy(c, tc, n * sizeof(idx_t));
    memcpy(v, tv, n * sizeof(float));
}

int orc_sort_rows_by_col(int64_t R, const idx_t *off, idx_t *col, float *val)
{
    int64_t maxlen = 1;
    for (int64_t i = 0; i < R; ++i)
        if (off[i + 1] - off[i] > maxlen) maxlen = off[i + 1] - off[i];
    idx_t *tc = (idx_t *)malloc(maxlen * sizeof(idx_t));
    float *tv = (float *)malloc(maxlen * sizeof(float));
    if (!tc || !tv) { free(tc); free(tv); return -1; }
    for (int64_t i = 0; i < R; ++i) merge_by_col(col + off[i], val + off[i], tc, tv, off[i + 1] - off[i]);
    free(tc); free(tv);
    return 0;
}

/* For every stored entry (i, k) of C: with s = sum_j a_ij b_jk and
 * t = sum_j |a_ij b_jk| accumulated in FP64, count |c_ik - s| > rtol * t.
 * Also counts stored entries with no partial product at all.  Returns the
 * number of failing entries (-1 on allocation failure); *first_bad receives
 * the flat index of the first one (or -1). */
int64_t orc_spgemm_check_f64(int64_t R, int64_t Cb,
                             const idx_t *Aoff, const idx_t *Acol, const float *Aval,
                             const idx_t *Boff, const idx_t *Bcol, const float *Bval,
                             const idx_t *Coff, const idx_t *Ccol, const float *Cval,
                             double rtol, int64_t *first_bad)
{
    double *s = (double *)calloc(Cb > 0 ? Cb : 1, sizeof(double));
    double *t = (double *)calloc(Cb > 0 ? Cb : 1, sizeof(double));
    unsigned char *hit = (unsigned char *)calloc(Cb > 0 ? Cb : 1, 1);
    if (!s || !t || !hit) { free(s); free(t); free(hit); return -1; }
    int64_t bad = 0;
    *first_bad = -1;
    for (int64_t i = 0; i < R; ++i) {
        for (idx_t jj = Aoff[i]; jj < Aoff[i + 1]; ++jj) {
            const idx_t j = Acol[jj];
            const double a = (double)Aval[jj];
            for (idx_t kk = Boff[j]; kk < Boff[j + 1]; ++kk) {
                const idx_t k = Bcol[kk];
                const double p = a * (double)Bval[kk];
                s[k] += p;
                t[k] += p < 0 ? -p : p;
                hit[k] = 1;
            }
        }
        for (idx_t e = Coff[i]; e < Coff[i + 1]; ++e) {
            const idx_t k = Ccol[e];
            const double d = (double)Cval[e] - s[k];
            if (!hit[k] || (d < 0 ? -d : d) > rtol * t[k]) {
                if (*first_bad < 0) *first_bad = e;
                bad++;
            }
        }
        for (idx_t jj = Aoff[i]; jj < Aoff[i + 1]; ++jj) {
            const idx_t j = Acol[jj];
            for (idx_t kk = Boff[j]; kk < Boff[j + 1]; ++kk) {
                const idx_t k = Bcol[kk];
                s[k] = 0.0;
                t[k] = 0.0;
                hit[k] = 0;
            }
        }
    }
    free(s); free(t); free(hit);
    return bad;
}

/* y = A x for HYB = ELL (width slots per row, column-major with pitch,
 * padding column -1) + COO (sorted by row): the host path initializes y to 0,
 * adds the ELL slots slot by slot (n outer, rows inner), then the COO entries
 * in stored order.  cusp/system/detail/sequential/multiply/hyb_spmv.h:42-58,
 * ell_spmv.h:54-73, coo_spmv.h:57-73. */
void orc_hyb_spmv(int64_t R, int64_t width, int64_t pitch, const idx_t *ell_col, const float *ell_val,
                  int64_t ncoo, const idx_t *coo_row, const idx_t *coo_col, const float *coo_val,
                  const float *x, float *y)
{
    for (int64_t i = 0; i < R; ++i) y[i] = 0.0f;
    for (int64_t n = 0; n < width; ++n)
        for (int64_t i = 0; i < R; ++i) {
            const idx_t j = ell_col[n * pitch + i];
            if (j != -1) y[i] = y[i] + ell_val[n * pitch + i] * x[j];
        }
    for (int64_t e = 0; e < ncoo; ++e) y[coo_row[e]] = y[coo_row[e]] + coo_val[e] * x[coo_col[e]];
}

/* Append-only HYB update, the reference's UpdateMatrix(hyb&, rows, cols,
 * vals, N) (include/load.h:327-373 -> UpdateMatrix_hyb,
 * include/sparse_update.inl:318-367): the batch is taken in row order
 * (stable: the reference requires it row-sorted, indices_to_offsets); each
 * entry of row r fills ELL slot row_sizes[r] while that is below the width,
 * else it is appended to the COO part; row_sizes[r] counts every entry of
 * the row.  The COO part is then stably sorted by row (thrust::sort_by_key),
 * which leaves a row's earlier overflow entries before the new ones in batch
 * order.  No merging, no zero drop: a coordinate may occur several times and
 * the represented matrix is the sum of its entries.  row_sizes has R + 1
 * entries ([R] = COO count); the COO output (capacity >= ncoo + n) must not
 * alias the input.  Returns the new COO count, or -1. */
int64_t orc_hyb_update(int64_t R, int64_t width, int64_t pitch, idx_t *row_sizes, idx_t *ell_col, float *ell_val,
                       int64_t ncoo, const idx_t *coo_row, const idx_t *coo_col, const float *coo_val,
                       int64_t n, const idx_t *brow, const idx_t *bcol, const float *bval,
                       idx_t *out_row, idx_t *out_col, float *out_val)
{
    idx_t *r = (idx_t *)malloc((n > 0 ? n : 1) * sizeof(idx_t));
    idx_t *c = (idx_t *)malloc((n > 0 ? n : 1) * sizeof(idx_t));
    float *v = (float *)malloc((n > 0 ? n : 1) * sizeof(float));
    idx_t *nr = (idx_t *)malloc((n > 0 ? n : 1) * sizeof(idx_t));
    idx_t *nc = (idx_t *)malloc((n > 0 ? n : 1) * sizeof(idx_t));
    float *nv = (float *)malloc((n > 0 ? n : 1) * sizeof(float));
    if (!r || !c || !v || !nr || !nc || !nv) { free(r); free(c); free(v); free(nr); free(nc); free(nv); return -1; }
    if (n > 0) {
        memcpy(r, brow, n * sizeof(idx_t));
        memcpy(c, bcol, n * sizeof(idx_t));
        memcpy(v, bval, n * sizeof(float));
    }
    orc_sort_by_row(n, r, c, v);
    int64_t m = 0;  /* new overflow entries, in row order */
    for (int64_t e = 0; e < n; ++e) {
        const idx_t row = r[e];
        const idx_t rl = row_sizes[row];
        if (rl < width) {
            ell_col[(int64_t)rl * pitch + row] = c[e];
            ell_val[(int64_t)rl * pitch + row] = v[e];
        } else {
            nr[m] = row; nc[m] = c[e]; nv[m] = v[e]; ++m;
        }
        row_sizes[row] = rl + 1;
    }
    /* stable merge by row: old COO entries of a row before the new ones */
    int64_t i = 0, j = 0, k = 0;
    while (i < ncoo || j < m) {
        const int take_old = j >= m || (i < ncoo && coo_row[i] <= nr[j]);
        if (take_old) { out_row[k] = coo_row[i]; out_col[k] = coo_col[i]; out_val[k] = coo_val[i]; ++i; }
        else { out_row[k] = nr[j]; out_col[k] = nc[j]; out_val[k] = nv[j]; ++j; }
        ++k;
    }
    row_sizes[R] = (idx_t)k;
    free(r); free(c); free(v); free(nr); free(nc); free(nv);
    return k;
}

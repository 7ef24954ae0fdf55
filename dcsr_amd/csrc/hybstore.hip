// hybstore.hip -- the standalone dynamic HYB matrix (ELL + COO overflow).
//
// Replaces the reference's hyb_matrix wrapper (include/dcsr_matrix.h:148-172)
// and the device entry points that fill and grow it:
//   LoadMatrix(hyb&, rows, cols, vals, N)  include/load.h:61-106
//                                          (LoadMatrix_hyb_coo,
//                                          include/sparse_update.inl:569-633)
//   LoadMatrix(hyb&, const csr&)           include/load.h:174-205
//                                          (LoadMatrix_hyb_csr, :635-698)
//   UpdateMatrix(hyb&, rows, cols, vals, N) include/load.h:327-373
//                                          (UpdateMatrix_hyb, :318-367)
// and its SpMV (cusp::multiply on the hyb matrix, src/Fill_Test.inl:167-170).
//
// Layout (the reference's): ELL block of `width` slots per row, column-major
// with pitch = rows rounded up to 32 (slot k of row r at k * pitch + r,
// padding column -1, value 0); COO overflow sorted by row; row_sizes[r] =
// entries of row r (ELL + COO), row_sizes[R] = COO count.  Beside it the
// store keeps coo_off[r], the first COO index of row r (exclusive scan of
// max(0, row_sizes[r] - width)), so SpMV finds a row's overflow directly.
//
// Semantics:
//   load   -- stable by row: the first `width` entries of each row, in input
//             order, fill its ELL slots, the rest go to COO in row order (the
//             host COO -> CSR -> HYB conversion,
//             cusp/system/detail/generic/conversions/csr_to_other.h:226-303).
//             The reference's overflow copy indexes cols[size - width + ...]
//             without the row's start, which this store does not reproduce.
//   update -- append, as UpdateMatrix_hyb: the batch taken in stable row
//             order, an entry of row r fills ELL slot row_sizes[r] while that
//             is below the width, otherwise it is appended to r's overflow
//             after r's earlier overflow entries.  No merge, no zero drop: a
//             coordinate may be stored several times and the matrix is the
//             sum of its entries (SpMV adds them all).
//   spmv   -- per row, 0 + the ELL slots in slot order + the COO entries in
//             order, left to right: the host hyb multiply's order
//             (sequential/multiply/hyb_spmv.h:42-58), so y is bit-identical
//             to it for alpha = 1, beta = 0.
//
// Kernels: every pass is flat over rows or over batch entries (coalesced);
// the COO merge of an update moves each old overflow entry once by the
// number of new overflow entries of the rows before it (a scan), the
// reference instead re-sorted the whole COO part (thrust::sort_by_key).
#include <string.h>
#include "dcsr_internal.h"
#include "prims.h"
#include "wave.h"

struct dcsr_hyb_s {
    int64_t rows, cols, pitch;
    int32_t width;
    int rbits;
    int32_t *ell_col;
    float *ell_val;
    int32_t *row_sizes;  // rows + 1
    uint32_t *coo_off;   // rows + 1
    int32_t *coo_row[2], *coo_col[2];
    float *coo_val[2];
    int cur;             // COO buffer holding the live entries
    int64_t coo_cap, ncoo;
    // grow-only scratch
    void *ws[8];
    size_t ws_bytes[8];
    unsigned int *status;  // device word: bad-index flag and scan diagnostics
    hipStream_t last;
};

namespace {

constexpr int TB = 256;
enum { HW_K0, HW_K1, HW_V0, HW_V1, HW_HIST, HW_PART, HW_CNT, HW_PAIR };

dcsr_status hs_ws(dcsr_hyb h, int slot, size_t bytes, void **out)
{
    if (bytes < 16) bytes = 16;
    if (h->ws_bytes[slot] < bytes) {
        if (h->ws[slot]) DCSR_HIP(hipFree(h->ws[slot]));
        h->ws[slot] = nullptr;
        h->ws_bytes[slot] = 0;
        const size_t nb = bytes + bytes / 4;
        DCSR_HIP(hipMalloc(&h->ws[slot], nb));
        h->ws_bytes[slot] = nb;
        if (slot == HW_PART) {  // look-back control words start at zero
            DCSR_HIP(hipMemset(h->ws[slot], 0, nb));
            DCSR_HIP(hipDeviceSynchronize());
        }
    }
    *out = h->ws[slot];
    return DCSR_OK;
}

__global__ void k_hs_fill_i32(int32_t *p, int64_t n, int32_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_hs_pad(int32_t *__restrict__ ell_col, float *__restrict__ ell_val, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        ell_col[i] = -1;
        ell_val[i] = 0.0f;
    }
}

// Sort key = row (stable sort keeps input order inside a row), value = input
// index; an out-of-range row or column becomes key ~0 and raises the flag.
__global__ void k_hs_pack(const int32_t *__restrict__ rows, const int32_t *__restrict__ cols, int64_t n, int64_t R,
                          int64_t C, uint64_t *__restrict__ key, uint32_t *__restrict__ idx, unsigned int *flag)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (i < n) {
        const int32_t r = rows[i], c = cols[i];
        bad = (r < 0) | (r >= R) | (c < 0) | (c >= C);
        key[i] = bad ? ~0ull : (uint64_t)(uint32_t)r;
        idx[i] = (uint32_t)i;
    }
    if (__ballot(bad) && lane_id() == 0) atomicOr(flag, 1u);
}

// Row runs of the sorted keys: cnt[r] = entries of row r, first[r] = its
// first sorted position (rows without entries keep cnt 0).
__global__ void k_hs_runs(const uint64_t *__restrict__ key, int64_t n, int32_t *__restrict__ cnt,
                          uint32_t *__restrict__ first)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = key[i];
    if (i > 0 && key[i - 1] == k) return;
    int64_t lo = i, hi = n;  // first position with a larger key
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (key[mid] <= k) lo = mid + 1; else hi = mid;
    }
    cnt[k] = (int32_t)(lo - i);
    first[k] = (uint32_t)i;
}

__global__ void k_hs_overflow(int64_t R, const int32_t *__restrict__ len, int32_t width, uint32_t *__restrict__ ov)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R) ov[r] = len[r] > width ? (uint32_t)(len[r] - width) : 0u;
}

// Load: sorted entry i of row r is the row's k-th (k = i - first[r]); slots
// k < width go to the ELL block, the rest to COO at coo_off[r] + k - width.
__global__ void k_hs_place_load(const uint64_t *__restrict__ key, const uint32_t *__restrict__ idx, int64_t n,
                                const uint32_t *__restrict__ first, const uint32_t *__restrict__ coo_off,
                                const int32_t *__restrict__ cols, const float *__restrict__ vals, int32_t width,
                                int64_t pitch, int32_t *__restrict__ ell_col, float *__restrict__ ell_val,
                                int32_t *__restrict__ coo_row, int32_t *__restrict__ coo_col,
                                float *__restrict__ coo_val)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t r = (uint32_t)key[i];
    const uint32_t src = idx[i];
    const uint32_t k = (uint32_t)i - first[r];
    const int32_t c = cols[src];
    const float v = vals[src];
    if (k < (uint32_t)width) {
        ell_col[(int64_t)k * pitch + r] = c;
        ell_val[(int64_t)k * pitch + r] = v;
    } else {
        const uint64_t o = (uint64_t)coo_off[r] + (k - (uint32_t)width);
        coo_row[o] = (int32_t)r;
        coo_col[o] = c;
        coo_val[o] = v;
    }
}

__global__ void k_hs_sizes_load(int64_t R, const int32_t *__restrict__ cnt, int32_t *__restrict__ row_sizes,
                                const uint32_t *__restrict__ coo_off)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R) row_sizes[r] = cnt[r];
    if (r == R) row_sizes[R] = (int32_t)coo_off[R];
}

// Update, per row: (old overflow, new overflow) of the append.  With s old
// entries and c new ones, the row holds max(0, s - width) overflow entries
// before and max(0, s + c - width) after.
__global__ void k_hs_pairs(int64_t R, const int32_t *__restrict__ row_sizes, const int32_t *__restrict__ cnt,
                           int32_t width, uint32_t *__restrict__ pair)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const int64_t s = row_sizes[r], c = cnt[r];
    const int64_t o0 = s > width ? s - width : 0;
    const int64_t o1 = s + c > width ? s + c - width : 0;
    pair[2 * r] = (uint32_t)o0;
    pair[2 * r + 1] = (uint32_t)(o1 - o0);
}

// Old overflow entry i (row r) moves past the new overflow entries of the
// rows before r: dest = i + nb[r] (pre[2r + 1], exclusive).
__global__ void k_hs_move_old(int64_t ncoo, const int32_t *__restrict__ orow, const int32_t *__restrict__ ocol,
                              const float *__restrict__ oval, const uint32_t *__restrict__ pre,
                              int32_t *__restrict__ nrow, int32_t *__restrict__ ncol, float *__restrict__ nval)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncoo) return;
    const int32_t r = orow[i];
    const uint64_t d = (uint64_t)i + pre[2 * (int64_t)r + 1];
    nrow[d] = r;
    ncol[d] = ocol[i];
    nval[d] = oval[i];
}

// New entry i (row r, its k-th new entry, s old entries): ELL slot s + k
// while below the width, else the row's (s + k - max(s, width))-th new
// overflow entry, placed after the row's old ones.
__global__ void k_hs_place_new(const uint64_t *__restrict__ key, const uint32_t *__restrict__ idx, int64_t n,
                               const uint32_t *__restrict__ first, const int32_t *__restrict__ row_sizes,
                               const uint32_t *__restrict__ pre, const int32_t *__restrict__ cols,
                               const float *__restrict__ vals, int32_t width, int64_t pitch,
                               int32_t *__restrict__ ell_col, float *__restrict__ ell_val,
                               int32_t *__restrict__ nrow, int32_t *__restrict__ ncol, float *__restrict__ nval)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t r = (uint32_t)key[i];
    const uint32_t src = idx[i];
    const int64_t k = (int64_t)i - first[r];
    const int64_t s = row_sizes[r];
    const int32_t c = cols[src];
    const float v = vals[src];
    if (s + k < width) {
        ell_col[(s + k) * pitch + r] = c;
        ell_val[(s + k) * pitch + r] = v;
    } else {
        const int64_t o0 = s > width ? s - width : 0;
        const int64_t j = s + k - (s > width ? s : (int64_t)width);
        // old overflow of rows <= r, new overflow of rows < r, then j
        const uint64_t d = (uint64_t)pre[2 * (int64_t)r] + (uint64_t)o0 + pre[2 * (int64_t)r + 1] + (uint64_t)j;
        nrow[d] = (int32_t)r;
        ncol[d] = c;
        nval[d] = v;
    }
}

__global__ void k_hs_sizes_update(int64_t R, const int32_t *__restrict__ cnt, const uint32_t *__restrict__ pre,
                                  int32_t *__restrict__ row_sizes, uint32_t *__restrict__ coo_off)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R) {
        row_sizes[r] += cnt[r];
        coo_off[r] = pre[2 * r] + pre[2 * r + 1];
    }
    if (r == R) {
        const uint32_t tot = pre[2 * R] + pre[2 * R + 1];
        coo_off[R] = tot;
        row_sizes[R] = (int32_t)tot;
    }
}

// y[r] = alpha * (0 + ELL slots in slot order + COO entries in order) +
// beta * y[r]: one lane per row; the ELL loads of a wave are contiguous (64
// consecutive rows of one slot column).
__global__ __launch_bounds__(TB) void k_hs_spmv(int64_t R, int32_t width, int64_t pitch,
                                                const int32_t *__restrict__ ell_col,
                                                const float *__restrict__ ell_val,
                                                const int32_t *__restrict__ row_sizes,
                                                const uint32_t *__restrict__ coo_off,
                                                const int32_t *__restrict__ coo_col,
                                                const float *__restrict__ coo_val, const float *__restrict__ x,
                                                float *__restrict__ y, float alpha, float beta)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const int32_t s = row_sizes[r];
    const int32_t ne = s < width ? s : width;
    float acc = 0.0f;
    int32_t k = 0;
    // four slots in flight
    for (; k + 4 <= ne; k += 4) {
        int32_t c[4];
        float v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            c[q] = ell_col[(int64_t)(k + q) * pitch + r];
            v[q] = ell_val[(int64_t)(k + q) * pitch + r];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) acc = acc + v[q] * x[c[q]];
    }
    for (; k < ne; ++k) acc = acc + ell_val[(int64_t)k * pitch + r] * x[ell_col[(int64_t)k * pitch + r]];
    for (uint32_t e = coo_off[r], e1 = coo_off[r + 1]; e < e1; ++e) acc = acc + coo_val[e] * x[coo_col[e]];
    y[r] = beta == 0.0f ? alpha * acc : alpha * acc + beta * y[r];
}

// Sort a batch of n triples stably by row; *bad (host) receives whether an
// index was out of range (then nothing else ran).  Outputs: sorted keys and
// input indices, per-row counts (zero elsewhere) and first positions.
dcsr_status hs_sort_batch(dcsr_hyb h, const int32_t *rows, const int32_t *cols, int64_t n, hipStream_t s,
                          bool *bad, uint64_t **key, uint32_t **idx, int32_t **cnt, uint32_t **first)
{
    const int64_t R = h->rows;
    uint64_t *k0, *k1;
    uint32_t *v0, *v1, *hist, *part;
    DCSR_TRY(hs_ws(h, HW_K0, (size_t)n * 8, (void **)&k0));
    DCSR_TRY(hs_ws(h, HW_K1, (size_t)n * 8, (void **)&k1));
    DCSR_TRY(hs_ws(h, HW_V0, (size_t)n * 4, (void **)&v0));
    DCSR_TRY(hs_ws(h, HW_V1, (size_t)n * 4, (void **)&v1));
    DCSR_TRY(hs_ws(h, HW_HIST, radix_hist_words(n) * 4, (void **)&hist));
    DCSR_TRY(hs_ws(h, HW_PART, scan_part_words(2 * (R + 1), 2) * 4, (void **)&part));
    DCSR_TRY(hs_ws(h, HW_CNT, (size_t)(R + 1) * 8, (void **)cnt));
    *first = (uint32_t *)(*cnt + (R + 1));
    k_hs_fill_i32<<<1, 64, 0, s>>>((int32_t *)h->status, 1, 0);
    if (n > 0)
        k_hs_pack<<<grid_for(n, TB), TB, 0, s>>>(rows, cols, n, R, h->cols, k0, v0, h->status);
    DCSR_HIP(hipGetLastError());
    unsigned int flag = 0;
    DCSR_HIP(hipMemcpyAsync(&flag, h->status, 4, hipMemcpyDeviceToHost, s));
    DCSR_HIP(hipStreamSynchronize(s));
    *bad = (flag & 1u) != 0;
    if (*bad) return DCSR_OK;
    int which = 0;
    DCSR_HIP(radix_sort_pairs(k0, v0, k1, v1, n, h->rbits, hist, part, &which, s));
    *key = which ? k1 : k0;
    *idx = which ? v1 : v0;
    k_hs_fill_i32<<<grid_for(R, TB * 4), TB, 0, s>>>(*cnt, R, 0);
    if (n > 0) k_hs_runs<<<grid_for(n, TB), TB, 0, s>>>(*key, n, *cnt, *first);
    DCSR_HIP(hipGetLastError());
    return DCSR_OK;
}

// Make the spare COO buffer (and, with keep, the live one) hold `need`
// entries.
dcsr_status hs_coo_reserve(dcsr_hyb h, int64_t need, bool keep, hipStream_t s)
{
    if (need <= h->coo_cap) return DCSR_OK;
    int64_t cap = need + need / 2 + 1024;
    int32_t *nr[2] = {nullptr, nullptr}, *nc[2] = {nullptr, nullptr};
    float *nv[2] = {nullptr, nullptr};
    for (int b = 0; b < 2; ++b) {
        DCSR_HIP(hipMalloc(&nr[b], (size_t)cap * 4));
        DCSR_HIP(hipMalloc(&nc[b], (size_t)cap * 4));
        DCSR_HIP(hipMalloc(&nv[b], (size_t)cap * 4));
    }
    if (keep && h->ncoo > 0) {
        DCSR_HIP(hipMemcpyAsync(nr[h->cur], h->coo_row[h->cur], (size_t)h->ncoo * 4, hipMemcpyDeviceToDevice, s));
        DCSR_HIP(hipMemcpyAsync(nc[h->cur], h->coo_col[h->cur], (size_t)h->ncoo * 4, hipMemcpyDeviceToDevice, s));
        DCSR_HIP(hipMemcpyAsync(nv[h->cur], h->coo_val[h->cur], (size_t)h->ncoo * 4, hipMemcpyDeviceToDevice, s));
    }
    DCSR_HIP(hipStreamSynchronize(s));
    for (int b = 0; b < 2; ++b) {
        (void)hipFree(h->coo_row[b]);
        (void)hipFree(h->coo_col[b]);
        (void)hipFree(h->coo_val[b]);
        h->coo_row[b] = nr[b];
        h->coo_col[b] = nc[b];
        h->coo_val[b] = nv[b];
    }
    h->coo_cap = cap;
    return DCSR_OK;
}

}  // namespace

dcsr_status dcsr_hyb_create(dcsr_hyb *out, int64_t rows, int64_t cols, int32_t width, int64_t coo_cap)
{
    if (!out || rows < 0 || cols < 0 || rows >= (1ll << 31) - 1 || cols >= (1ll << 31) || width < 0 || coo_cap < 0)
        return DCSR_ERR_INVALID;
    const int64_t pitch = (rows + 31) / 32 * 32;
    if ((uint64_t)width * (uint64_t)pitch >= (1ull << 32)) return DCSR_ERR_UNSUPPORTED;
    dcsr_hyb h = (dcsr_hyb)calloc(1, sizeof(struct dcsr_hyb_s));
    if (!h) return DCSR_ERR_OOM;
    h->rows = rows;
    h->cols = cols;
    h->width = width;
    h->pitch = pitch;
    h->rbits = bits_for(rows + 1);
    const size_t ncell = (size_t)width * (size_t)pitch;
    dcsr_status st = DCSR_OK;
    if (hipMalloc(&h->ell_col, (ncell ? ncell : 1) * 4) != hipSuccess ||
        hipMalloc(&h->ell_val, (ncell ? ncell : 1) * 4) != hipSuccess ||
        hipMalloc(&h->row_sizes, (size_t)(rows + 1) * 4) != hipSuccess ||
        hipMalloc(&h->coo_off, (size_t)(rows + 1) * 4) != hipSuccess ||
        hipMalloc(&h->status, 64) != hipSuccess)
        st = DCSR_ERR_OOM;
    if (st == DCSR_OK) {
        // an empty matrix: every slot padding, every row empty
        if (ncell) k_hs_pad<<<grid_for((int64_t)ncell, TB * 4), TB>>>(h->ell_col, h->ell_val, (int64_t)ncell);
        k_hs_fill_i32<<<grid_for(rows + 1, TB * 4), TB>>>(h->row_sizes, rows + 1, 0);
        k_hs_fill_i32<<<grid_for(rows + 1, TB * 4), TB>>>((int32_t *)h->coo_off, rows + 1, 0);
        if (hipGetLastError() != hipSuccess) st = DCSR_ERR_HIP;
    }
    if (st == DCSR_OK) st = hs_coo_reserve(h, coo_cap > 0 ? coo_cap : 1, false, nullptr);
    if (st != DCSR_OK) {
        dcsr_hyb_destroy(h);
        return st;
    }
    *out = h;
    return DCSR_OK;
}

dcsr_status dcsr_hyb_destroy(dcsr_hyb h)
{
    if (!h) return DCSR_ERR_INVALID;
    if (h->last) (void)hipStreamSynchronize(h->last);
    else (void)hipDeviceSynchronize();
    (void)hipFree(h->ell_col);
    (void)hipFree(h->ell_val);
    (void)hipFree(h->row_sizes);
    (void)hipFree(h->coo_off);
    (void)hipFree(h->status);
    for (int b = 0; b < 2; ++b) {
        (void)hipFree(h->coo_row[b]);
        (void)hipFree(h->coo_col[b]);
        (void)hipFree(h->coo_val[b]);
    }
    for (int i = 0; i < 8; ++i)
        if (h->ws[i]) (void)hipFree(h->ws[i]);
    free(h);
    return DCSR_OK;
}

dcsr_status dcsr_hyb_view_get(dcsr_hyb h, dcsr_hyb_view *v)
{
    if (!h || !v) return DCSR_ERR_INVALID;
    memset(v, 0, sizeof(*v));
    v->rows = h->rows;
    v->cols = h->cols;
    v->width = h->width;
    v->pitch = h->pitch;
    v->ell_col = h->ell_col;
    v->ell_val = h->ell_val;
    v->row_sizes = h->row_sizes;
    v->coo_row = h->coo_row[h->cur];
    v->coo_col = h->coo_col[h->cur];
    v->coo_val = h->coo_val[h->cur];
    v->num_overflow = h->ncoo;
    v->coo_cap = h->coo_cap;
    return DCSR_OK;
}

static dcsr_status hyb_load(dcsr_hyb h, const int32_t *rows, const int32_t *cols, const float *vals, int64_t n,
                            hipStream_t s)
{
    const int64_t R = h->rows;
    bool bad = false;
    uint64_t *key = nullptr;
    uint32_t *idx = nullptr, *first = nullptr, *pair = nullptr, *part = nullptr;
    int32_t *cnt = nullptr;
    DCSR_TRY(hs_sort_batch(h, rows, cols, n, s, &bad, &key, &idx, &cnt, &first));
    if (bad) return DCSR_ERR_BAD_INDEX;
    DCSR_TRY(hs_ws(h, HW_PAIR, (size_t)(2 * (R + 1)) * 4, (void **)&pair));
    part = (uint32_t *)h->ws[HW_PART];
    k_hs_overflow<<<grid_for(R, TB), TB, 0, s>>>(R, cnt, h->width, pair);
    DCSR_HIP(scan_exclusive_u32(pair, h->coo_off, R, 1, part, h->coo_off + R, s, h->status));
    uint32_t ncoo = 0;
    DCSR_HIP(hipMemcpyAsync(&ncoo, h->coo_off + R, 4, hipMemcpyDeviceToHost, s));
    DCSR_HIP(hipStreamSynchronize(s));
    DCSR_TRY(hs_coo_reserve(h, (int64_t)ncoo, false, s));
    const int64_t ncell = (int64_t)h->width * h->pitch;
    if (ncell) k_hs_pad<<<grid_for(ncell, TB * 4), TB, 0, s>>>(h->ell_col, h->ell_val, ncell);
    if (n > 0)
        k_hs_place_load<<<grid_for(n, TB), TB, 0, s>>>(key, idx, n, first, h->coo_off, cols, vals, h->width, h->pitch,
                                                       h->ell_col, h->ell_val, h->coo_row[h->cur],
                                                       h->coo_col[h->cur], h->coo_val[h->cur]);
    k_hs_sizes_load<<<grid_for(R + 1, TB), TB, 0, s>>>(R, cnt, h->row_sizes, h->coo_off);
    DCSR_HIP(hipGetLastError());
    h->ncoo = ncoo;
    DCSR_HIP(hipStreamSynchronize(s));
    return DCSR_OK;
}

dcsr_status dcsr_hyb_load_coo(dcsr_hyb h, const int32_t *rows, const int32_t *cols, const float *vals, int64_t n,
                              dcsr_stream_t stream)
{
    if (!h || n < 0 || (n > 0 && (!rows || !cols || !vals)) || n >= (int64_t)0x7ffffff0ll) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    h->last = s;
    return hyb_load(h, rows, cols, vals, n, s);
}

namespace {
__global__ void k_hs_expand_rows(const int32_t *__restrict__ off, int64_t R, int64_t n, int32_t *__restrict__ rows)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t lo = 0, hi = R;  // largest r with off[r] <= i
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid; else hi = mid;
    }
    rows[i] = (int32_t)lo;
}
}  // namespace

dcsr_status dcsr_hyb_load_csr(dcsr_hyb h, const int32_t *row_offsets, const int32_t *cols, const float *vals,
                              dcsr_stream_t stream)
{
    if (!h || !row_offsets) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    h->last = s;
    int32_t ends[2] = {0, 0};
    DCSR_HIP(hipMemcpyAsync(&ends[0], row_offsets, 4, hipMemcpyDeviceToHost, s));
    DCSR_HIP(hipMemcpyAsync(&ends[1], row_offsets + h->rows, 4, hipMemcpyDeviceToHost, s));
    DCSR_HIP(hipStreamSynchronize(s));
    const int64_t n = ends[1];
    if (ends[0] != 0 || n < 0 || (n > 0 && (!cols || !vals))) return DCSR_ERR_INVALID;
    // row index of every entry, for the stable row sort of the COO path
    int32_t *rtmp;
    DCSR_HIP(hipMalloc(&rtmp, (size_t)(n > 0 ? n : 1) * 4));
    if (n > 0 && h->rows > 0) k_hs_expand_rows<<<grid_for(n, TB), TB, 0, s>>>(row_offsets, h->rows, n, rtmp);
    DCSR_HIP(hipGetLastError());
    const dcsr_status st = hyb_load(h, rtmp, cols, vals, n, s);
    (void)hipFree(rtmp);
    return st;
}

dcsr_status dcsr_hyb_update(dcsr_hyb h, const int32_t *rows, const int32_t *cols, const float *vals, int64_t n,
                            dcsr_stream_t stream)
{
    if (!h || n < 0 || (n > 0 && (!rows || !cols || !vals)) || n >= (int64_t)0x7ffffff0ll) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    h->last = s;
    if (n == 0) return DCSR_OK;
    const int64_t R = h->rows;
    bool bad = false;
    uint64_t *key = nullptr;
    uint32_t *idx = nullptr, *first = nullptr, *pair = nullptr, *pre = nullptr;
    int32_t *cnt = nullptr;
    DCSR_TRY(hs_sort_batch(h, rows, cols, n, s, &bad, &key, &idx, &cnt, &first));
    if (bad) return DCSR_ERR_BAD_INDEX;
    DCSR_TRY(hs_ws(h, HW_PAIR, (size_t)(4 * (R + 1)) * 4, (void **)&pair));
    pre = pair + 2 * (R + 1);
    uint32_t *part = (uint32_t *)h->ws[HW_PART];
    k_hs_pairs<<<grid_for(R, TB), TB, 0, s>>>(R, h->row_sizes, cnt, h->width, pair);
    DCSR_HIP(scan_exclusive_u32(pair, pre, R, 2, part, pre + 2 * R, s, h->status));
    uint32_t tot[2] = {0, 0};
    DCSR_HIP(hipMemcpyAsync(tot, pre + 2 * R, 8, hipMemcpyDeviceToHost, s));
    DCSR_HIP(hipStreamSynchronize(s));
    const int64_t ncoo_new = (int64_t)tot[0] + tot[1];
    if (ncoo_new >= (int64_t)0x7ffffff0ll) return DCSR_ERR_UNSUPPORTED;
    DCSR_TRY(hs_coo_reserve(h, ncoo_new, true, s));
    const int a = h->cur, b = a ^ 1;
    if (h->ncoo > 0)
        k_hs_move_old<<<grid_for(h->ncoo, TB), TB, 0, s>>>(h->ncoo, h->coo_row[a], h->coo_col[a], h->coo_val[a], pre,
                                                          h->coo_row[b], h->coo_col[b], h->coo_val[b]);
    k_hs_place_new<<<grid_for(n, TB), TB, 0, s>>>(key, idx, n, first, h->row_sizes, pre, cols, vals, h->width,
                                                  h->pitch, h->ell_col, h->ell_val, h->coo_row[b], h->coo_col[b],
                                                  h->coo_val[b]);
    k_hs_sizes_update<<<grid_for(R + 1, TB), TB, 0, s>>>(R, cnt, pre, h->row_sizes, h->coo_off);
    DCSR_HIP(hipGetLastError());
    h->cur = b;
    h->ncoo = ncoo_new;
    return DCSR_OK;
}

dcsr_status dcsr_hyb_spmv(dcsr_hyb h, const float *x, float *y, float alpha, float beta, dcsr_stream_t stream)
{
    if (!h || (h->rows > 0 && (!x || !y))) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    h->last = s;
    if (h->rows == 0) return DCSR_OK;
    k_hs_spmv<<<grid_for(h->rows, TB), TB, 0, s>>>(h->rows, h->width, h->pitch, h->ell_col, h->ell_val, h->row_sizes,
                                                   h->coo_off, h->coo_col[h->cur], h->coo_val[h->cur], x, y, alpha,
                                                   beta);
    DCSR_HIP(hipGetLastError());
    return DCSR_OK;
}

dcsr_status dcsr_hyb_export(dcsr_hyb h, int32_t *ell_col, float *ell_val, int32_t *row_sizes, int32_t *coo_row,
                            int32_t *coo_col, float *coo_val, int64_t coo_cap, int64_t *ncoo, dcsr_stream_t stream)
{
    if (!h || !ncoo) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    h->last = s;
    *ncoo = h->ncoo;
    if (h->ncoo > coo_cap) return DCSR_ERR_CAPACITY;
    const size_t ncell = (size_t)h->width * (size_t)h->pitch;
    if ((ncell && (!ell_col || !ell_val)) || !row_sizes || (h->ncoo && (!coo_row || !coo_col || !coo_val)))
        return DCSR_ERR_INVALID;
    if (ncell) {
        DCSR_HIP(hipMemcpyAsync(ell_col, h->ell_col, ncell * 4, hipMemcpyDeviceToDevice, s));
        DCSR_HIP(hipMemcpyAsync(ell_val, h->ell_val, ncell * 4, hipMemcpyDeviceToDevice, s));
    }
    DCSR_HIP(hipMemcpyAsync(row_sizes, h->row_sizes, (size_t)(h->rows + 1) * 4, hipMemcpyDeviceToDevice, s));
    if (h->ncoo) {
        DCSR_HIP(hipMemcpyAsync(coo_row, h->coo_row[h->cur], (size_t)h->ncoo * 4, hipMemcpyDeviceToDevice, s));
        DCSR_HIP(hipMemcpyAsync(coo_col, h->coo_col[h->cur], (size_t)h->ncoo * 4, hipMemcpyDeviceToDevice, s));
        DCSR_HIP(hipMemcpyAsync(coo_val, h->coo_val[h->cur], (size_t)h->ncoo * 4, hipMemcpyDeviceToDevice, s));
    }
    DCSR_HIP(hipStreamSynchronize(s));
    return DCSR_OK;
}

dcsr_status dcsr_hyb_to_dcsr(dcsr_hyb h, dcsr_handle dst, dcsr_stream_t stream)
{
    if (!h || !dst || dst->rows != h->rows || dst->cols != h->cols) return DCSR_ERR_INVALID;
    h->last = (hipStream_t)stream;
    return dcsr_load_hyb(dst, h->width, h->pitch, h->ell_col, h->ell_val, h->coo_row[h->cur], h->coo_col[h->cur],
                         h->coo_val[h->cur], h->ncoo, stream);
}

dcsr_status dcsr_hyb_from_dcsr(dcsr_hyb h, dcsr_handle src, dcsr_stream_t stream)
{
    if (!h || !src || src->rows != h->rows || src->cols != h->cols) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    h->last = s;
    dcsr_info info;
    DCSR_TRY(dcsr_query(src, &info));
    const int64_t nnz = info.nnz;
    int32_t *off, *col;
    float *val;
    DCSR_HIP(hipMalloc(&off, (size_t)(h->rows + 1) * 4));
    DCSR_HIP(hipMalloc(&col, (size_t)(nnz > 0 ? nnz : 1) * 4));
    DCSR_HIP(hipMalloc(&val, (size_t)(nnz > 0 ? nnz : 1) * 4));
    int64_t got = 0;
    dcsr_status st = dcsr_export_csr(src, off, col, val, nnz, &got, 0, s);
    if (st == DCSR_OK) st = dcsr_hyb_load_csr(h, off, col, val, s);
    (void)hipFree(off);
    (void)hipFree(col);
    (void)hipFree(val);
    return st;
}

// csrcmp.hip -- CSR-rebuild comparators: the reference's CSR baselines for
// the Dynamic-CSR experiment, rebuilt on this engine's primitives.
//
// * dcsr_csr_add: C = A + B with A a canonical CSR matrix and B a batch of
//   triples, rebuilt whole as cusp's generic add does it (A and B as COO
//   concatenated, sorted by (row, column), reduced by key, zeros removed:
//   cusp/system/detail/generic/elementwise.inl:177-225; the reference driver
//   calls it per update slice, src/Fill_Test.inl:240-258).  Every call reads
//   and writes the whole matrix.
// * dcsr_csr_spgemm: C = A * B of two CSR matrices by global ESC as cusp's
//   generic multiply does it (every partial product expanded, sorted by
//   (row, column), reduced by key: cusp/system/detail/generic/multiply/
//   spgemm.h:43-312, used by src/SPMM_Test.inl:142).
//
// Both reduce each (row, column) run left to right in the host path's order
// and drop exact zeros, so their results equal the oracle and dcsr_update /
// dcsr_spgemm bit for bit: add sums A's entries of a coordinate from 0 in
// stored order, the batch's from 0 in arrival order, and adds the two (the
// host cusp::add order, cusp/system/detail/sequential/elementwise.h:44-131;
// A may hold duplicates, as a loaded DCSR matrix may); the product sums a
// column's partials in A-row order then B-row order
// (sequential/multiply/csr_spgemm.h:104-129).  The stable sort keeps that
// order: for add, each key carries its source in the lowest bit (A's entry
// first); for the product, the expansion writes partials in A-row, B-row
// order.  Temporaries are allocated per call; both calls sync.
#include "dcsr_internal.h"
#include "prims.h"
#include "wave.h"

namespace {

constexpr int CT = 256;

// Row of each entry of a CSR matrix (nnz entries, rows + 1 offsets).
__global__ void k_cmp_rows(const int32_t *__restrict__ off, int64_t rows, int32_t *__restrict__ rowof)
{
    const int64_t r = (int64_t)blockIdx.x * CT + threadIdx.x;
    if (r >= rows) return;
    for (int32_t i = off[r]; i < off[r + 1]; ++i) rowof[i] = (int32_t)r;
}

// add: keys of A's entries (source bit 0) and B's triples (source bit 1);
// an out-of-range triple raises *bad.
__global__ void k_cmp_add_keys(const int32_t *__restrict__ arow, const int32_t *__restrict__ acol,
                               const float *__restrict__ aval, int64_t na, const int32_t *__restrict__ brow,
                               const int32_t *__restrict__ bcol, const float *__restrict__ bval, int64_t nb,
                               int64_t R, int64_t C, int cbits, uint64_t *__restrict__ key,
                               uint32_t *__restrict__ val, unsigned int *bad)
{
    const int64_t i = (int64_t)blockIdx.x * CT + threadIdx.x;
    if (i < na) {
        key[i] = (((uint64_t)(uint32_t)arow[i] << cbits) | (uint32_t)acol[i]) << 1;
        val[i] = __float_as_uint(aval[i]);
    } else if (i < na + nb) {
        const int64_t j = i - na;
        const int32_t r = brow[j], c = bcol[j];
        const bool ok = r >= 0 && r < R && c >= 0 && c < C;
        if (!ok) atomicOr(bad, 1u);
        key[i] = ok ? ((((uint64_t)(uint32_t)r << cbits) | (uint32_t)c) << 1) | 1ull : ~0ull;
        val[i] = __float_as_uint(bval[j]);
    }
}

// spgemm: partial products of A's entry e at pexp[e] .. pexp[e+1], in B-row
// order; key = (row << cbits) | column.
__global__ void k_cmp_expand(const int32_t *__restrict__ arow, const int32_t *__restrict__ acol,
                             const float *__restrict__ aval, int64_t na, const int32_t *__restrict__ boff,
                             const int32_t *__restrict__ bcol, const float *__restrict__ bval,
                             const uint32_t *__restrict__ pexp, int cbits, uint64_t *__restrict__ key,
                             uint32_t *__restrict__ val)
{
    const int64_t e = (int64_t)blockIdx.x * CT + threadIdx.x;
    if (e >= na) return;
    const int32_t k = acol[e];
    const float a = aval[e];
    const uint64_t hi = (uint64_t)(uint32_t)arow[e] << cbits;
    uint64_t p = pexp[e];
    for (int32_t j = boff[k]; j < boff[k + 1]; ++j, ++p) {
        key[p] = hi | (uint32_t)bcol[j];
        val[p] = __float_as_uint(a * bval[j]);
    }
}

// len[e] = partial products of A's entry e; *total (64-bit) their sum.
__global__ void k_cmp_seglen(const int32_t *__restrict__ acol, int64_t na, const int32_t *__restrict__ boff,
                             uint32_t *__restrict__ len, unsigned long long *total)
{
    const int64_t e = (int64_t)blockIdx.x * CT + threadIdx.x;
    uint32_t l = 0;
    if (e < na) {
        l = (uint32_t)(boff[acol[e] + 1] - boff[acol[e]]);
        len[e] = l;
    }
    const uint32_t w = wave_sum_u32(l);
    if (lane_id() == 0 && w) atomicAdd(total, (unsigned long long)w);
}

// Reduce each run of equal coordinates (key >> sbits) left to right.  With
// sbits == 1 (add) the run's A entries (source bit 0, first) and B entries
// are summed separately from 0 and then added, as the host add does:
// (0 + a1 + ...) + (0 + b1 + ...).  keep[i] = 1 for a run head whose result
// is not 0.0; res[i] its value.
__global__ void k_cmp_reduce(const uint64_t *__restrict__ key, const uint32_t *__restrict__ val, int64_t n, int sbits,
                             uint32_t *__restrict__ keep, float *__restrict__ res)
{
    const int64_t i = (int64_t)blockIdx.x * CT + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = key[i];
    const bool head = k != ~0ull && (i == 0 || (key[i - 1] >> sbits) != (k >> sbits));
    uint32_t kp = 0;
    float r = 0.0f;
    if (head) {
        float sa = 0.0f, sb = 0.0f;
        for (int64_t j = i; j < n && (key[j] >> sbits) == (k >> sbits); ++j) {
            const float x = __uint_as_float(val[j]);
            if (sbits == 1 && (key[j] & 1ull) == 0ull) sa = sa + x;
            else sb = sb + x;
        }
        r = sbits == 1 ? sa + sb : sb;
        kp = r != 0.0f ? 1u : 0u;
    }
    keep[i] = kp;
    res[i] = r;
}

__global__ void k_cmp_write(const uint64_t *__restrict__ key, const uint32_t *__restrict__ keep,
                            const uint32_t *__restrict__ pos, const float *__restrict__ res, int64_t n, int sbits,
                            int cbits, int32_t *__restrict__ crow, int32_t *__restrict__ ccol, float *__restrict__ cval)
{
    const int64_t i = (int64_t)blockIdx.x * CT + threadIdx.x;
    if (i >= n || !keep[i]) return;
    const uint64_t k = key[i] >> sbits;
    const uint32_t p = pos[i];
    crow[p] = (int32_t)(k >> cbits);
    ccol[p] = (int32_t)(k & ((1ull << cbits) - 1ull));
    cval[p] = res[i];
}

// off[r] = entries of C in rows < r (crow sorted), r in [0, rows].
__global__ void k_cmp_offsets(const int32_t *__restrict__ crow, int64_t nc, int64_t rows, int32_t *__restrict__ off)
{
    const int64_t r = (int64_t)blockIdx.x * CT + threadIdx.x;
    if (r > rows) return;
    int64_t lo = 0, hi = nc;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (crow[mid] < r) lo = mid + 1; else hi = mid;
    }
    off[r] = (int32_t)lo;
}

// Temporaries of one call, freed on every exit.
struct Tmp {
    static constexpr int MAXP = 24;
    void *p[MAXP] = {};
    int n = 0;
    ~Tmp()
    {
        for (int i = 0; i < n; ++i) (void)hipFree(p[i]);
    }
    template <typename T>
    dcsr_status get(T **out, size_t count)
    {
        void *q = nullptr;
        if (n == MAXP) return DCSR_ERR_INVALID;
        if (hipMalloc(&q, (count ? count : 1) * sizeof(T)) != hipSuccess) return DCSR_ERR_OOM;
        p[n++] = q;
        *out = (T *)q;
        return DCSR_OK;
    }
};

unsigned blocks_of(int64_t n) { return (unsigned)((n + CT - 1) / CT < 1 ? 1 : (n + CT - 1) / CT); }

// Sort (key, val) pairs (n of them, stable) and reduce them into C.  The
// sorted pairs end in (k0, v0) or (k1, v1) as the sort reports.
dcsr_status sort_reduce_write(Tmp &t, uint64_t *k0, uint32_t *v0, int64_t n, int keybits, int sbits, int cbits,
                              int64_t rows, int32_t *c_off, int32_t *c_col, float *c_val, int64_t c_cap,
                              int64_t *c_nnz, hipStream_t s)
{
    uint64_t *k1;
    uint32_t *v1, *hist, *part, *keep, *pos, *tot;
    float *res;
    int32_t *crow;
    DCSR_TRY(t.get(&k1, n));
    DCSR_TRY(t.get(&v1, n));
    DCSR_TRY(t.get(&hist, radix_hist_words(n)));
    DCSR_TRY(t.get(&part, scan_part_words(n, 1)));
    DCSR_TRY(t.get(&keep, n));
    DCSR_TRY(t.get(&pos, n));
    DCSR_TRY(t.get(&tot, 4));
    DCSR_TRY(t.get(&res, n));
    DCSR_HIP(hipMemsetAsync(part, 0, scan_part_words(n, 1) * 4, s));
    int which = 0;
    DCSR_HIP(radix_sort_pairs(k0, v0, k1, v1, n, keybits, hist, part, &which, s, false));
    const uint64_t *key = which ? k1 : k0;
    const uint32_t *val = which ? v1 : v0;
    k_cmp_reduce<<<blocks_of(n), CT, 0, s>>>(key, val, n, sbits, keep, res);
    DCSR_HIP(scan_exclusive_u32(keep, pos, n, 1, part, tot, s, nullptr));
    uint32_t total = 0;
    DCSR_HIP(hipMemcpyAsync(&total, tot, 4, hipMemcpyDeviceToHost, s));
    DCSR_HIP(hipStreamSynchronize(s));
    *c_nnz = total;
    if ((int64_t)total > c_cap) return DCSR_ERR_CAPACITY;
    DCSR_TRY(t.get(&crow, total));
    k_cmp_write<<<blocks_of(n), CT, 0, s>>>(key, keep, pos, res, n, sbits, cbits, crow, c_col, c_val);
    k_cmp_offsets<<<blocks_of(rows + 1), CT, 0, s>>>(crow, total, rows, c_off);
    DCSR_HIP(hipGetLastError());
    DCSR_HIP(hipStreamSynchronize(s));
    return DCSR_OK;
}

}  // namespace

extern "C" dcsr_status dcsr_csr_add(int64_t rows, int64_t cols, const int32_t *a_off, const int32_t *a_col,
                                    const float *a_val, const int32_t *b_row, const int32_t *b_col,
                                    const float *b_val, int64_t nb, int32_t *c_off, int32_t *c_col, float *c_val,
                                    int64_t c_cap, int64_t *c_nnz, dcsr_stream_t stream)
{
    if (rows < 0 || cols < 0 || rows >= (1ll << 31) || cols >= (1ll << 31) || nb < 0 || !a_off || !c_off || !c_nnz)
        return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    int32_t na32 = 0;
    DCSR_HIP(hipMemcpyAsync(&na32, a_off + rows, 4, hipMemcpyDeviceToHost, s));
    DCSR_HIP(hipStreamSynchronize(s));
    const int64_t na = na32, n = na + nb;
    if ((na && (!a_col || !a_val)) || (nb && (!b_row || !b_col || !b_val))) return DCSR_ERR_INVALID;
    if (n >= (int64_t)0x7ffffff0ll) return DCSR_ERR_INVALID;
    const int cbits = bits_for(cols), rbits = bits_for(rows);
    Tmp t;
    int32_t *arow;
    uint64_t *k0;
    uint32_t *v0, *bad;
    DCSR_TRY(t.get(&arow, na));
    DCSR_TRY(t.get(&k0, n));
    DCSR_TRY(t.get(&v0, n));
    DCSR_TRY(t.get(&bad, 1));
    DCSR_HIP(hipMemsetAsync(bad, 0, 4, s));
    k_cmp_rows<<<blocks_of(rows), CT, 0, s>>>(a_off, rows, arow);
    k_cmp_add_keys<<<blocks_of(n), CT, 0, s>>>(arow, a_col, a_val, na, b_row, b_col, b_val, nb, rows, cols, cbits,
                                                k0, v0, bad);
    uint32_t hbad = 0;
    DCSR_HIP(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, s));
    DCSR_HIP(hipStreamSynchronize(s));
    if (hbad) return DCSR_ERR_BAD_INDEX;
    if (n == 0) {
        *c_nnz = 0;
        k_cmp_offsets<<<blocks_of(rows + 1), CT, 0, s>>>(nullptr, 0, rows, c_off);
        DCSR_HIP(hipStreamSynchronize(s));
        return DCSR_OK;
    }
    return sort_reduce_write(t, k0, v0, n, rbits + cbits + 1, 1, cbits, rows, c_off, c_col, c_val, c_cap, c_nnz, s);
}

extern "C" dcsr_status dcsr_csr_spgemm(int64_t a_rows, int64_t a_cols, int64_t b_cols, const int32_t *a_off,
                                       const int32_t *a_col, const float *a_val, const int32_t *b_off,
                                       const int32_t *b_col, const float *b_val, int32_t *c_off, int32_t *c_col,
                                       float *c_val, int64_t c_cap, int64_t *c_nnz, dcsr_stream_t stream)
{
    if (a_rows < 0 || a_cols < 0 || b_cols < 0 || a_rows >= (1ll << 31) || b_cols >= (1ll << 31) || !a_off ||
        !b_off || !c_off || !c_nnz)
        return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    int32_t na32 = 0;
    DCSR_HIP(hipMemcpyAsync(&na32, a_off + a_rows, 4, hipMemcpyDeviceToHost, s));
    DCSR_HIP(hipStreamSynchronize(s));
    const int64_t na = na32;
    const int cbits = bits_for(b_cols), rbits = bits_for(a_rows);
    Tmp t;
    int32_t *arow;
    uint32_t *len, *pexp, *tot, *part;
    unsigned long long *p64;
    DCSR_TRY(t.get(&arow, na));
    DCSR_TRY(t.get(&len, na + 1));
    DCSR_TRY(t.get(&pexp, na + 1));
    DCSR_TRY(t.get(&tot, 4));
    DCSR_TRY(t.get(&p64, 1));
    DCSR_TRY(t.get(&part, scan_part_words(na + 1, 1)));
    DCSR_HIP(hipMemsetAsync(part, 0, scan_part_words(na + 1, 1) * 4, s));
    DCSR_HIP(hipMemsetAsync(p64, 0, 8, s));
    k_cmp_rows<<<blocks_of(a_rows), CT, 0, s>>>(a_off, a_rows, arow);
    k_cmp_seglen<<<blocks_of(na), CT, 0, s>>>(a_col, na, b_off, len, p64);
    unsigned long long P64 = 0;
    DCSR_HIP(hipMemcpyAsync(&P64, p64, 8, hipMemcpyDeviceToHost, s));
    DCSR_HIP(hipStreamSynchronize(s));
    // the sort and scans index partials with 32 bits
    if (P64 >= 0x7ffffff0ull) return DCSR_ERR_UNSUPPORTED;
    DCSR_HIP(scan_exclusive_u32(len, pexp, na, 1, part, tot, s, nullptr));
    const uint32_t P = (uint32_t)P64;
    if (P == 0) {
        *c_nnz = 0;
        k_cmp_offsets<<<blocks_of(a_rows + 1), CT, 0, s>>>(nullptr, 0, a_rows, c_off);
        DCSR_HIP(hipStreamSynchronize(s));
        return DCSR_OK;
    }
    uint64_t *k0;
    uint32_t *v0;
    DCSR_TRY(t.get(&k0, P));
    DCSR_TRY(t.get(&v0, P));
    k_cmp_expand<<<blocks_of(na), CT, 0, s>>>(arow, a_col, a_val, na, b_off, b_col, b_val, pexp, cbits, k0, v0);
    return sort_reduce_write(t, k0, v0, P, rbits + cbits, 0, cbits, a_rows, c_off, c_col, c_val, c_cap, c_nnz, s);
}

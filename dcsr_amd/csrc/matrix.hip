// matrix.hip -- handle lifetime, COO/CSR -> DCSR load, DCSR -> CSR export,
// compaction (defrag), pool growth, hub-row list, status plumbing.
//
// Reference counterparts: dcsr_matrix::resize (include/dcsr_matrix.h:41-108),
// device::Initialize_Matrix (include/load.h:8-22), device::LoadMatrix
// (include/load.h:27-59, 207-253), device::SortMatrixRow (include/load.h:404-581),
// device::BinRows (include/load.h:583-610).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "dcsr_internal.h"
#include "prims.h"
#include "segments.h"

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
namespace {

constexpr int TB = 256;

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Empty rows with capacity L: start = r*L (InitializeMatrix_dcsr,
// include/sparse_update.inl:1155-1176).
__global__ void k_init_rows(int64_t R, uint32_t L, uint32_t *start, int32_t *len, int32_t *cap, DevState *ds)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R) {
        start[r] = (uint32_t)(r * L);
        len[r] = 0;
        cap[r] = (int32_t)L;
    }
    if (r == 0) {
        ds->bump = (unsigned long long)R * L;
        ds->nnz = 0;
        ds->garbage = 0;
        ds->status &= ST_SCAN_FALLBACK;
        ds->hub_count = 0;
        ds->hub_chunks = 0;
    }
}

// Pack (row, col) into a sort key; invalid indices become ~0 and raise
// ST_BAD_INDEX.
__global__ void k_pack_keys(const int32_t *__restrict__ rows, const int32_t *__restrict__ cols,
                            const float *__restrict__ vals, int64_t n, int64_t R, int64_t C, int cbits,
                            uint64_t *__restrict__ key, uint32_t *__restrict__ val, DevState *ds)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (i < n) {
        int32_t r = rows[i], c = cols[i];
        bad = (r < 0) | (r >= R) | (c < 0) | (c >= C);
        key[i] = bad ? ~0ull : (((uint64_t)(uint32_t)r << cbits) | (uint32_t)c);
        val[i] = __float_as_uint(vals[i]);
    }
    if (__ballot(bad) && lane_id() == 0) atomicOr(&ds->status, ST_BAD_INDEX);
}

// Expand CSR offsets into row indices (one thread per entry, binary search).
__global__ void k_expand_rows(const int32_t *__restrict__ off, int64_t R, int64_t n, int32_t *__restrict__ rows)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t lo = 0, hi = R;  // largest r with off[r] <= i
    while (hi - lo > 1) {
        int64_t mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid; else hi = mid;
    }
    rows[i] = (int32_t)lo;
}

__global__ void k_check_offsets(const int32_t *__restrict__ off, int64_t R, DevState *ds)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (r < R) bad = off[r + 1] < off[r];
    if (r == 0) bad |= off[0] != 0;
    if (__ballot(bad) && lane_id() == 0) atomicOr(&ds->status, ST_BAD_INDEX);
}

// After the sort: per row, its length and first sorted position.  Also
// detects duplicate coordinates and explicit zeros (non-canonical input).
__global__ void k_row_runs(const uint64_t *__restrict__ key, const uint32_t *__restrict__ val, int64_t n, int cbits,
                           int32_t *__restrict__ row_len, uint32_t *__restrict__ first, uint32_t *flag_noncanon)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool nc = false;
    if (i < n) {
        uint64_t k = key[i];
        uint32_t r = (uint32_t)(k >> cbits);
        bool head = (i == 0) || ((uint32_t)(key[i - 1] >> cbits) != r);
        nc = (i > 0 && key[i - 1] == k) || (__uint_as_float(val[i]) == 0.0f);
        if (head) {
            uint64_t nextkey = (uint64_t)(r + 1) << cbits;
            int64_t lo = i, hi = n;  // first index with key >= nextkey
            while (lo < hi) {
                int64_t mid = (lo + hi) >> 1;
                if (key[mid] < nextkey) lo = mid + 1; else hi = mid;
            }
            row_len[r] = (int32_t)(lo - i);
            first[r] = (uint32_t)i;
        }
    }
    if (__ballot(nc) && lane_id() == 0) atomicOr(flag_noncanon, 1u);
}

// Write sorted entries into the pool at row_start[r] + (i - first[r]).
__global__ void k_fill_pool(const uint64_t *__restrict__ key, const uint32_t *__restrict__ val, int64_t n, int cbits,
                            const uint32_t *__restrict__ start, const uint32_t *__restrict__ first,
                            int32_t *__restrict__ pcol, float *__restrict__ pval)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t k = key[i];
    uint32_t r = (uint32_t)(k >> cbits);
    uint64_t p = (uint64_t)start[r] + (uint64_t)(i - first[r]);
    pcol[p] = (int32_t)(k & ((1ull << cbits) - 1ull));
    pval[p] = __uint_as_float(val[i]);
}

__global__ void k_set_loaded(int64_t R, const uint32_t *__restrict__ start, const int32_t *__restrict__ len,
                             int32_t *__restrict__ cap, DevState *ds, uint64_t nnz)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R) cap[r] = len[r];
    if (r == 0) {
        ds->bump = nnz;
        ds->nnz = (long long)nnz;
        ds->garbage = 0;
        ds->hub_count = 0;
        ds->hub_chunks = 0;
    }
}

// Flat copy of every row into destination offsets (compaction / export).
// dst_off[r] is the row's new start; src rows are (start, len).  Each wave
// owns CP_SPAN destination entries, in chunks of at most 64 rows: a start
// map of the chunk's rows (ranked to rows through `tab`: empty rows share a
// start) gives every entry its row without a search, and all of a lane's
// loads are in flight before its first store.
constexpr uint32_t CP_SPAN = 1024;
constexpr uint32_t CP_PER = CP_SPAN / 64;
__global__ __launch_bounds__(TB) void k_copy_rows(const uint32_t *__restrict__ dst_off, uint32_t R, uint32_t total,
                                                  const uint32_t *__restrict__ src_start,
                                                  const int32_t *__restrict__ src_col, const float *__restrict__ src_val,
                                                  int32_t *__restrict__ dst_col, float *__restrict__ dst_val)
{
    struct Lds {
        uint32_t off[65];
        uint32_t src[64];
        uint32_t pmap[CP_SPAN / 32];
        uint32_t tab[64];
    };
    __shared__ Lds sh[TB / 64];
    Lds &S = sh[threadIdx.x >> 6];
    const uint32_t lane = (uint32_t)lane_id();
    const uint32_t dev_total = dst_off[R];  // the offsets decide, not the caller's count
    if (total > dev_total) total = dev_total;
    const uint64_t nwaves = (uint64_t)gridDim.x * (TB / 64);
    for (uint64_t wave = (uint64_t)blockIdx.x * (TB / 64) + (threadIdx.x >> 6); wave * CP_SPAN < total;
         wave += nwaves) {
        const uint32_t e0 = (uint32_t)(wave * CP_SPAN);
        const uint32_t e1 = total - e0 < CP_SPAN ? total : e0 + CP_SPAN;
        uint32_t j = wave_find_last_le(dst_off, R, e0);
        uint32_t e = e0;
        while (e < e1) {
            const uint32_t t = j + lane;
            const uint32_t off_t = t <= R ? dst_off[t] : 0xffffffffu;
            const uint32_t off_64 = j + 64u <= R ? dst_off[j + 64u] : 0xffffffffu;
            const uint32_t src = t < R ? src_start[t] : 0u;
            S.off[lane] = off_t;
            if (lane == 0) S.off[64] = off_64;
            S.src[lane] = src;
            if (lane < CP_SPAN / 32) S.pmap[lane] = 0u;
            __builtin_amdgcn_wave_barrier();
            const uint32_t wend = S.off[64] < e1 ? S.off[64] : e1;
            const uint32_t cnt = wend - e;
            const int j0 = find_seg64(S.off, e);  // the row holding e
            // rows with entries starting in the chunk
            const bool starts_here = t < R && off_t >= e && off_t < wend && S.off[lane + 1] > off_t;
            if (starts_here) atomicOr(&S.pmap[(off_t - e) >> 5], 1u << ((off_t - e) & 31u));
            __builtin_amdgcn_wave_barrier();
            const StartMap pm = start_map_load(S.pmap, CP_SPAN / 64);
            const uint32_t rank = start_map_rank(pm, starts_here ? off_t - e : 0u);  // every lane takes part
            if (starts_here) S.tab[rank] = lane;
            __builtin_amdgcn_wave_barrier();
            int32_t c[CP_PER];
            float v[CP_PER];
#pragma unroll
            for (uint32_t q = 0; q < CP_PER; ++q) {
                const uint32_t p = q * 64u + lane;
                const uint32_t nst = start_map_count(pm, (int)q);
                c[q] = 0;
                v[q] = 0.0f;
                if (p < cnt) {
                    const uint32_t jr = nst == 0u ? (uint32_t)j0 : S.tab[nst - 1u];
                    const uint64_t sp = (uint64_t)S.src[jr] + (e + p - S.off[jr]);
                    c[q] = src_col[sp];
                    v[q] = src_val[sp];
                }
            }
#pragma unroll
            for (uint32_t q = 0; q < CP_PER; ++q) {
                const uint32_t p = q * 64u + lane;
                if (p < cnt) {
                    dst_col[e + p] = c[q];
                    dst_val[e + p] = v[q];
                }
            }
            __builtin_amdgcn_wave_barrier();
            e = wend;
            j += 64;
        }
    }
}

__global__ void k_len_u32(const int32_t *__restrict__ len, int64_t R, uint32_t *__restrict__ out)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R) out[r] = (uint32_t)len[r];
}

__global__ void k_u32_to_i32_offsets(const uint32_t *__restrict__ in, int64_t n, int32_t *__restrict__ out)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) out[r] = (int32_t)in[r];
}

// New row table after compaction: start = off[r], cap = len.
__global__ void k_set_compacted(int64_t R, const uint32_t *__restrict__ off, uint32_t *__restrict__ start,
                                const int32_t *__restrict__ len, int32_t *__restrict__ cap, DevState *ds,
                                const uint32_t *__restrict__ total)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R) {
        start[r] = off[r];
        cap[r] = len[r];
    }
    if (r == 0) {
        ds->bump = total[0];
        ds->nnz = total[0];
        ds->garbage = 0;
    }
}

// Canonical form of a natural-order CSR: emit one entry per run of equal
// columns with value (0 + a1) + a2 + ... (stored order) when it is not 0.
// Pass 1 flags heads that emit.
__global__ void k_canon_flags(const int32_t *__restrict__ off_i, int64_t R, const int32_t *__restrict__ col,
                              const float *__restrict__ val, int64_t n, uint32_t *__restrict__ emit,
                              float *__restrict__ sum)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // row of entry i
    int64_t lo = 0, hi = R;
    while (hi - lo > 1) {
        int64_t mid = (lo + hi) >> 1;
        if (off_i[mid] <= i) lo = mid; else hi = mid;
    }
    int64_t rs = off_i[lo], re = off_i[lo + 1];
    bool head = (i == rs) || (col[i - 1] != col[i]);
    uint32_t e = 0;
    if (head) {
        float s = 0.0f + val[i];
        for (int64_t m = i + 1; m < re && col[m] == col[i]; ++m) s = s + val[m];
        e = (s != 0.0f) ? 1u : 0u;
        sum[i] = s;
    }
    emit[i] = e;
}

__global__ void k_canon_write(const int32_t *__restrict__ col, int64_t n, const uint32_t *__restrict__ emit_flag,
                              const uint32_t *__restrict__ pos, const float *__restrict__ sum,
                              int32_t *__restrict__ ocol, float *__restrict__ oval)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (emit_flag[i]) {
        ocol[pos[i]] = col[i];
        oval[pos[i]] = sum[i];
    }
}

__global__ void k_canon_offsets(const int32_t *__restrict__ off_in, int64_t R, const uint32_t *__restrict__ pos,
                                const uint32_t *__restrict__ total, int32_t *__restrict__ off_out, int64_t n)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r <= R) {
        int64_t o = off_in[r];
        off_out[r] = (int32_t)(o < n ? pos[o] : total[0]);
    }
}

// Number of rows longer than DCSR_HUB_LEN (reported by dcsr_query).
__global__ void k_count_hubs(int64_t R, const int32_t *__restrict__ len, DevState *ds)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool hub = r < R && len[r] > DCSR_HUB_LEN;
    uint64_t m = __ballot(hub);
    if (m && lane_id() == 0) atomicAdd(&ds->hub_count, (uint32_t)__popcll(m));
}

__global__ void k_zero_hubs(DevState *ds) { ds->hub_count = 0; }

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
dcsr_status dcsr_hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    if (getenv("DCSR_VERBOSE"))
        fprintf(stderr, "[dcsr] HIP error %d (%s) in %s at %s:%d\n", (int)e, hipGetErrorString(e), what, file, line);
    return e == hipErrorOutOfMemory ? DCSR_ERR_OOM : DCSR_ERR_HIP;
}

const char *dcsr_status_string(dcsr_status s)
{
    switch (s) {
    case DCSR_OK: return "ok";
    case DCSR_ERR_INVALID: return "invalid argument";
    case DCSR_ERR_BAD_INDEX: return "row or column index out of range";
    case DCSR_ERR_POOL_EXHAUSTED: return "entry pool exhausted";
    case DCSR_ERR_SEGMENT_OVERFLOW: return "row segment overflow";
    case DCSR_ERR_HIP: return "HIP runtime error";
    case DCSR_ERR_OOM: return "device out of memory";
    case DCSR_ERR_CAPACITY: return "output buffer too small";
    case DCSR_ERR_UNSUPPORTED: return "unsupported";
    case DCSR_ERR_TOO_LARGE: return "request exceeds the allocation-free entry point's fixed scratch";
    }
    return "unknown";
}

namespace {
__global__ void k_zero_words(uint32_t *p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}
}  // namespace

static void release_retired(dcsr_handle h);

// Grow a slot.  The old buffer may still be read by work queued on any
// stream, so it is retired (freed at the next state read-back or at
// destroy), never freed here: hipFree would wait for the whole device.
static dcsr_status ws_grow(dcsr_handle h, WsSlot slot, size_t bytes)
{
    if (h->ws.ptr[slot]) {
        if (h->nretired == (int)(sizeof(h->retired) / sizeof(h->retired[0]))) {
            DCSR_HIP(hipDeviceSynchronize());
            release_retired(h);
        }
        h->retired[h->nretired++] = h->ws.ptr[slot];
    }
    h->ws.ptr[slot] = nullptr;
    h->ws.bytes[slot] = 0;
    const size_t nb = bytes + bytes / 4;  // headroom for growing batches
    DCSR_HIP(hipMalloc(&h->ws.ptr[slot], nb));
    h->ws.bytes[slot] = nb;
    return DCSR_OK;
}

dcsr_status ws_get(dcsr_handle h, WsSlot slot, size_t bytes, void **out)
{
    if (bytes == 0) bytes = 16;
    if (h->ws.bytes[slot] < bytes) DCSR_TRY(ws_grow(h, slot, bytes));
    *out = h->ws.ptr[slot];
    return DCSR_OK;
}

dcsr_status ws_get_zeroed(dcsr_handle h, WsSlot slot, size_t bytes, void **out, hipStream_t s)
{
    if (bytes == 0) bytes = 16;
    if (h->ws.bytes[slot] < bytes) {
        DCSR_TRY(ws_grow(h, slot, bytes));
        // stream-ordered before the slot's first use on s
        const int64_t words = (int64_t)(h->ws.bytes[slot] / 4);
        int64_t blocks = (words + 255) / 256;
        if (blocks > 1024) blocks = 1024;
        k_zero_words<<<(unsigned)blocks, 256, 0, s>>>((uint32_t *)h->ws.ptr[slot], words);
        DCSR_HIP(hipGetLastError());
    }
    *out = h->ws.ptr[slot];
    return DCSR_OK;
}


dcsr_status pull_state(dcsr_handle h, hipStream_t s)
{
    DCSR_HIP(hipMemcpyAsync(h->hs, h->ds, sizeof(DevState), hipMemcpyDeviceToHost, s));
    DCSR_HIP(hipStreamSynchronize(s));
    if (h->nretired) release_retired(h);
    h->nnz_h = h->hs->nnz;
    h->bump_h = (int64_t)h->hs->bump;
    h->garbage_h = h->hs->garbage;
    state_exact(h);
    return DCSR_OK;
}

void state_exact(dcsr_handle h)
{
    h->state_known = true;
    h->snap_valid_from = h->snap_seq + 1;
}

void state_unknown(dcsr_handle h, int64_t added, int64_t claimed)
{
    const int64_t nnz = live_bound(h) + added, bump = bump_bound(h) + claimed;
    // neither can pass the pool: live entries sit in it and the device
    // rejects a batch whose claims would
    h->nnz_ub = nnz < h->pool_cap ? nnz : h->pool_cap;
    h->bump_ub = bump < h->pool_cap ? bump : h->pool_cap;
    h->cum_add += added;
    h->cum_claim += claimed;
    h->state_known = false;
    h->hubs_dirty = true;
}

// The next NOSYNC update's sequence number (its write kernel stores the
// snapshot under it).
uint64_t snapshot_next(dcsr_handle h) { return ++h->snap_seq; }

// After update seq's launches and its state_unknown: the running sums that
// its snapshot's values are to be read against.
void snapshot_note(dcsr_handle h, uint64_t seq)
{
    h->ring_add[seq % SNAP_RING] = h->cum_add;
    h->ring_claim[seq % SNAP_RING] = h->cum_claim;
}

// The newest landed snapshot k bounds the state: its values plus the sums
// added after update k.  The write kernel stores the values before k (a
// release at system scope), so values read after k was seen are those of k
// or of a later update j, whose sums after j are at least k's: the bound
// holds either way.  A snapshot older than the ring, or than the last change
// the sums do not count, is skipped.
void refresh_bounds(dcsr_handle h)
{
    const uint64_t k = __atomic_load_n(&h->hsnap->seq, __ATOMIC_ACQUIRE);
    if (k == h->snap_used || k < h->snap_valid_from || k > h->snap_seq || k + SNAP_RING <= h->snap_seq) return;
    h->snap_used = k;
    if (h->state_known) return;
    const int64_t nnz = (int64_t)__atomic_load_n(&h->hsnap->nnz, __ATOMIC_RELAXED) +
                        (h->cum_add - h->ring_add[k % SNAP_RING]);
    const int64_t bump = (int64_t)__atomic_load_n(&h->hsnap->bump, __ATOMIC_RELAXED) +
                         (h->cum_claim - h->ring_claim[k % SNAP_RING]);
    if (nnz < h->nnz_ub) h->nnz_ub = nnz;
    if (bump < h->bump_ub) h->bump_ub = bump;
}

static dcsr_status status_from_bits(unsigned bits)
{
    if (bits & ST_BAD_INDEX) return DCSR_ERR_BAD_INDEX;
    if (bits & ST_SEGMENT_OVERFLOW) return DCSR_ERR_SEGMENT_OVERFLOW;
    if (bits & ST_POOL_EXHAUSTED) return DCSR_ERR_POOL_EXHAUSTED;
    return DCSR_OK;
}

// Clears the error bits (the scan-fallback diagnostic stays for the handle's life).
__global__ void k_clear_status(DevState *ds) { ds->status &= ST_SCAN_FALLBACK; }

dcsr_status dcsr_get_status(dcsr_handle h)
{
    if (!h) return DCSR_ERR_INVALID;
    DCSR_TRY(pull_state(h, h->last_stream));
    unsigned bits = h->hs->status;
    k_clear_status<<<1, 1, 0, h->last_stream>>>(h->ds);
    DCSR_HIP(hipGetLastError());
    return status_from_bits(bits);
}

static dcsr_status alloc_rows(dcsr_handle h)
{
    size_t R = (size_t)(h->rows > 0 ? h->rows : 1);
    DCSR_HIP(hipMalloc(&h->row_start, R * sizeof(uint32_t)));
    DCSR_HIP(hipMalloc(&h->row_len, R * sizeof(int32_t)));
    DCSR_HIP(hipMalloc(&h->row_cap, R * sizeof(int32_t)));
    return DCSR_OK;
}

dcsr_status alloc_pool(dcsr_handle h, int64_t cap)
{
    if (cap < h->cfg.min_pool) cap = h->cfg.min_pool;
    if (cap > (int64_t)0xfffffff0ll) cap = (int64_t)0xfffffff0ll;
    if (h->pool_col) DCSR_HIP(hipFree(h->pool_col));
    if (h->pool_val) DCSR_HIP(hipFree(h->pool_val));
    h->pool_col = nullptr;
    h->pool_val = nullptr;
    DCSR_HIP(hipMalloc(&h->pool_col, (size_t)cap * sizeof(int32_t)));
    DCSR_HIP(hipMalloc(&h->pool_val, (size_t)cap * sizeof(float)));
    h->pool_cap = cap;
    return DCSR_OK;
}

dcsr_status dcsr_create(dcsr_handle *out, int64_t rows, int64_t cols, const dcsr_config *cfg)
{
    if (!out || rows < 0 || cols < 0 || rows >= (1ll << 31) || cols >= (1ll << 31)) return DCSR_ERR_INVALID;
    dcsr_handle h = (dcsr_handle)calloc(1, sizeof(struct dcsr_matrix_s));
    if (!h) return DCSR_ERR_OOM;
    h->rows = rows;
    h->cols = cols;
    h->rbits = bits_for(rows);
    h->cbits = bits_for(cols);
    h->cfg.pool_factor = 2.0;
    h->cfg.min_pool = 1024;
    h->cfg.allow_grow = 1;
    if (cfg) {
        if (cfg->pool_factor >= 1.0) h->cfg.pool_factor = cfg->pool_factor;
        if (cfg->min_pool > 0) h->cfg.min_pool = cfg->min_pool;
        h->cfg.allow_grow = cfg->allow_grow;
    }
    h->canonical = true;
    h->hubs_dirty = false;
    h->flat_dirty = true;
    state_exact(h);
    // SpMV column panels: 1.5M columns (6 MB of x; measured best on a 10M x
    // 10M power-law matrix among 1M, 1.25M, 1.5M, 2M and 3M with non-temporal
    // entry loads: 1.91, 1.76, 1.74, 1.80, 2.13 ms) for matrices of at least
    // 16M entries; both can be overridden for testing
    h->pnl_cols = 1536 << 10;
    h->pnl_prev_calls = DCSR_PNL_REUSE;  // no evidence yet: a fresh matrix is assumed to be reused
    h->pnl_min_nnz = 16ll << 20;
    if (const char *e = getenv("DCSR_SPMV_PANEL_COLS")) h->pnl_cols = atoll(e) > 0 ? atoll(e) : h->pnl_cols;
    if (const char *e = getenv("DCSR_SPMV_PANEL_MIN_NNZ")) h->pnl_min_nnz = atoll(e) >= 0 ? atoll(e) : h->pnl_min_nnz;
    // the update tag's starting value (a test of its wrap-around)
    if (const char *e = getenv("DCSR_UPDATE_TAG_START")) h->upd_tag = (uint32_t)strtoul(e, nullptr, 0);
    dcsr_status st = alloc_rows(h);
    if (st == DCSR_OK) st = alloc_pool(h, h->cfg.min_pool);
    if (st == DCSR_OK && hipMalloc(&h->ds, sizeof(DevState)) != hipSuccess) st = DCSR_ERR_OOM;
    if (st == DCSR_OK && hipHostMalloc(&h->hs, sizeof(DevState), hipHostMallocDefault) != hipSuccess) st = DCSR_ERR_OOM;
    // the NOSYNC snapshot: coherent, so the write kernel's stores reach the
    // host without a copy
    if (st == DCSR_OK &&
        hipHostMalloc(&h->hsnap, sizeof(HostSnap), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
        st = DCSR_ERR_OOM;
    if (st == DCSR_OK) {
        memset(h->hsnap, 0, sizeof(HostSnap));
        h->snap_valid_from = 1;
        if (hipHostGetDevicePointer((void **)&h->dsnap, h->hsnap, 0) != hipSuccess) st = DCSR_ERR_HIP;
    }
    if (st == DCSR_OK && hipMemset(h->ds, 0, sizeof(DevState)) != hipSuccess) st = DCSR_ERR_HIP;
    if (st == DCSR_OK) {
        // all rows empty
        k_init_rows<<<grid_for(rows, TB), TB>>>(rows, 0, h->row_start, h->row_len, h->row_cap, h->ds);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) st = DCSR_ERR_HIP;
    }
    if (st != DCSR_OK) {
        dcsr_destroy(h);
        return st;
    }
    *out = h;
    return DCSR_OK;
}

dcsr_status dcsr_destroy(dcsr_handle h)
{
    if (!h) return DCSR_ERR_INVALID;
    if (h->last_stream) (void)hipStreamSynchronize(h->last_stream);
    else (void)hipDeviceSynchronize();
    (void)hipFree(h->row_start);
    (void)hipFree(h->row_len);
    (void)hipFree(h->row_cap);
    (void)hipFree(h->pool_col);
    (void)hipFree(h->pool_val);
    (void)hipFree(h->spare_col);
    (void)hipFree(h->spare_val);
    release_retired(h);
    (void)hipFree(h->ds);
    if (h->hs) (void)hipHostFree(h->hs);
    if (h->hsnap) (void)hipHostFree(h->hsnap);
    for (int i = 0; i < WS_COUNT; ++i)
        if (h->ws.ptr[i]) (void)hipFree(h->ws.ptr[i]);
    free(h);
    return DCSR_OK;
}

dcsr_status dcsr_initialize(dcsr_handle h, int32_t bin_length, dcsr_stream_t stream)
{
    if (!h || bin_length < 0) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    h->last_stream = s;
    int64_t need = h->rows * (int64_t)bin_length;
    if (need > (int64_t)0xfffffff0ll) return DCSR_ERR_INVALID;
    if (need > h->pool_cap) {
        DCSR_HIP(hipStreamSynchronize(s));
        DCSR_TRY(alloc_pool(h, (int64_t)(need * h->cfg.pool_factor)));
    }
    k_init_rows<<<grid_for(h->rows, TB), TB, 0, s>>>(h->rows, (uint32_t)bin_length, h->row_start, h->row_len,
                                                       h->row_cap, h->ds);
    DCSR_HIP(hipGetLastError());
    h->nnz_h = 0;
    h->bump_h = need;
    h->garbage_h = 0;
    h->hub_rows_h = 0;
    h->hub_chunks_h = 0;
    h->canonical = true;
    h->hubs_dirty = true;
    h->flat_dirty = true;
    state_exact(h);
    return DCSR_OK;
}

// Load sorted-by-key triples into a fresh natural-order pool.
static dcsr_status load_sorted(dcsr_handle h, const uint64_t *key, const uint32_t *val, int64_t n, hipStream_t s)
{
    const int64_t R = h->rows;
    uint32_t *first, *scan_part, *flag;
    DCSR_TRY(ws_get(h, WS_TMP0, (size_t)(R + 1) * sizeof(uint32_t), (void **)&first));
    DCSR_TRY(ws_get_zeroed(h, WS_PART, scan_part_words(R + 1, 1) * sizeof(uint32_t), (void **)&scan_part, s));
    DCSR_TRY(ws_get(h, WS_TMP2, 64, (void **)&flag));
    k_fill_i32<<<grid_for(R, TB * 4), TB, 0, s>>>(h->row_len, R, 0);
    k_fill_i32<<<1, 64, 0, s>>>((int32_t *)flag, 2, 0);
    k_row_runs<<<grid_for(n, TB), TB, 0, s>>>(key, val, n, h->cbits, h->row_len, first, flag);
    // row starts = exclusive scan of lengths (natural order, no slack)
    uint32_t *lens;
    DCSR_TRY(ws_get(h, WS_TMP1, (size_t)(R + 1) * sizeof(uint32_t), (void **)&lens));
    k_len_u32<<<grid_for(R, TB), TB, 0, s>>>(h->row_len, R, lens);
    DCSR_HIP(scan_exclusive_u32(lens, h->row_start, R, 1, scan_part, flag + 1, s, fault_word(h)));
    int64_t want = (int64_t)(n * h->cfg.pool_factor);
    if (want < n) want = n;
    if (want > h->pool_cap || n > h->pool_cap) {
        DCSR_HIP(hipStreamSynchronize(s));
        DCSR_TRY(alloc_pool(h, want));
    }
    // the spare pool (compaction target) is sized with the main pool so that
    // compactions during updates allocate nothing
    if (h->spare_cap < h->pool_cap) {
        DCSR_HIP(hipStreamSynchronize(s));
        if (h->spare_col) (void)hipFree(h->spare_col);
        if (h->spare_val) (void)hipFree(h->spare_val);
        h->spare_col = nullptr;
        h->spare_val = nullptr;
        h->spare_cap = 0;
        DCSR_HIP(hipMalloc(&h->spare_col, (size_t)h->pool_cap * 4));
        DCSR_HIP(hipMalloc(&h->spare_val, (size_t)h->pool_cap * 4));
        h->spare_cap = h->pool_cap;
    }
    k_fill_pool<<<grid_for(n, TB), TB, 0, s>>>(key, val, n, h->cbits, h->row_start, first, h->pool_col, h->pool_val);
    k_set_loaded<<<grid_for(R > 0 ? R : 1, TB), TB, 0, s>>>(R, h->row_start, h->row_len, h->row_cap, h->ds,
                                                            (uint64_t)n);
    DCSR_HIP(hipGetLastError());
    uint32_t hflag[2];
    DCSR_HIP(hipMemcpyAsync(hflag, flag, sizeof(hflag), hipMemcpyDeviceToHost, s));
    DCSR_TRY(pull_state(h, s));
    h->canonical = hflag[0] == 0;
    h->hubs_dirty = true;
    h->flat_dirty = true;
    h->hub_rows_h = -1;  // unknown until the hub list is rebuilt
    return DCSR_OK;
}

static dcsr_status load_coo_impl(dcsr_handle h, const int32_t *rows, const int32_t *cols, const float *vals,
                                 int64_t n, hipStream_t s)
{
    if (n >= (int64_t)0xfffffff0ll) return DCSR_ERR_INVALID;
    uint64_t *k0, *k1;
    uint32_t *v0, *v1, *hist, *part;
    DCSR_TRY(ws_get(h, WS_KEY_A, (size_t)n * 8, (void **)&k0));
    DCSR_TRY(ws_get(h, WS_KEY_B, (size_t)n * 8, (void **)&k1));
    DCSR_TRY(ws_get(h, WS_VAL_A, (size_t)n * 4, (void **)&v0));
    DCSR_TRY(ws_get(h, WS_VAL_B, (size_t)n * 4, (void **)&v1));
    DCSR_TRY(ws_get(h, WS_HIST, radix_hist_words(n) * 4, (void **)&hist));
    DCSR_TRY(ws_get_zeroed(h, WS_SCAN, radix_part_words(n) * 4, (void **)&part, s));
    k_clear_status<<<1, 1, 0, s>>>(h->ds);
    if (n > 0)
        k_pack_keys<<<grid_for(n, TB), TB, 0, s>>>(rows, cols, vals, n, h->rows, h->cols, h->cbits, k0, v0, h->ds);
    DCSR_HIP(hipGetLastError());
    // reject out-of-range indices before anything is modified
    DCSR_TRY(pull_state(h, s));
    if (h->hs->status & ST_BAD_INDEX) {
        k_clear_status<<<1, 1, 0, s>>>(h->ds);
        return DCSR_ERR_BAD_INDEX;
    }
    int which = 0;
    DCSR_HIP(radix_sort_pairs(k0, v0, k1, v1, n, h->rbits + h->cbits, hist, part, &which, s));
    return load_sorted(h, which ? k1 : k0, which ? v1 : v0, n, s);
}

dcsr_status load_coo_internal(dcsr_handle h, const int32_t *rows, const int32_t *cols, const float *vals, int64_t n,
                              hipStream_t s)
{
    return load_coo_impl(h, rows, cols, vals, n, s);
}

dcsr_status dcsr_load_coo(dcsr_handle h, const int32_t *rows, const int32_t *cols, const float *vals, int64_t n,
                          dcsr_stream_t stream)
{
    if (!h || n < 0 || (n > 0 && (!rows || !cols || !vals))) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    h->last_stream = s;
    return load_coo_impl(h, rows, cols, vals, n, s);
}

dcsr_status dcsr_load_csr(dcsr_handle h, const int32_t *row_offsets, const int32_t *cols, const float *vals,
                          dcsr_stream_t stream)
{
    if (!h || !row_offsets) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    h->last_stream = s;
    int32_t nnz = 0;
    DCSR_HIP(hipMemcpyAsync(&nnz, row_offsets + h->rows, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    k_clear_status<<<1, 1, 0, s>>>(h->ds);
    if (h->rows > 0) k_check_offsets<<<grid_for(h->rows, TB), TB, 0, s>>>(row_offsets, h->rows, h->ds);
    DCSR_TRY(pull_state(h, s));
    if (nnz < 0 || (h->hs->status & ST_BAD_INDEX)) return DCSR_ERR_INVALID;
    if (nnz > 0 && (!cols || !vals)) return DCSR_ERR_INVALID;
    int32_t *rows;
    DCSR_TRY(ws_get(h, WS_CANON, (size_t)nnz * 4, (void **)&rows));
    if (nnz > 0) k_expand_rows<<<grid_for(nnz, TB), TB, 0, s>>>(row_offsets, h->rows, nnz, rows);
    return load_coo_impl(h, rows, cols, vals, nnz, s);
}

// Exclusive offsets of the current row lengths (u32, R+1 entries) into `off`.
static dcsr_status row_offsets_u32(dcsr_handle h, uint32_t *off, hipStream_t s)
{
    uint32_t *part;
    const int64_t R = h->rows;
    DCSR_TRY(ws_get_zeroed(h, WS_PART, scan_part_words(R + 1, 1) * 4, (void **)&part, s));
    // lengths are non-negative int32: scan them as uint32
    DCSR_HIP(scan_exclusive_u32((const uint32_t *)h->row_len, off, R, 1, part, off + R, s, fault_word(h)));
    return DCSR_OK;
}

void launch_copy_rows(const uint32_t *dst_off, int64_t R, int64_t total, const uint32_t *src_start,
                             const int32_t *scol, const float *sval, int32_t *dcol, float *dval, hipStream_t s)
{
    if (total <= 0 || R <= 0) return;
    const int64_t waves = (total + CP_SPAN - 1) / CP_SPAN;
    int64_t blocks = (waves + 3) / 4;
    if (blocks > 2048) blocks = 2048;  // waves stride over the rest
    k_copy_rows<<<(unsigned)blocks, TB, 0, s>>>(dst_off, (uint32_t)R, (uint32_t)total, src_start, scol, sval, dcol,
                                                dval);
}

dcsr_status dcsr_export_csr(dcsr_handle h, int32_t *row_offsets, int32_t *cols, float *vals, int64_t cap,
                            int64_t *nnz_out, uint32_t flags, dcsr_stream_t stream)
{
    if (!h || !row_offsets || !nnz_out) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    h->last_stream = s;
    const int64_t R = h->rows;
    if (!h->state_known) DCSR_TRY(pull_state(h, s));
    const int64_t nnz = h->nnz_h;
    bool canon = (flags & DCSR_EXPORT_CANONICAL) && !h->canonical;
    if (!canon && cap < nnz) return DCSR_ERR_CAPACITY;
    if (nnz > 0 && (!cols || !vals)) return DCSR_ERR_INVALID;
    uint32_t *off;
    DCSR_TRY(ws_get(h, WS_TMP0, (size_t)(R + 1) * 4, (void **)&off));
    DCSR_TRY(row_offsets_u32(h, off, s));
    if (!canon) {
        k_u32_to_i32_offsets<<<grid_for(R + 1, TB), TB, 0, s>>>(off, R + 1, row_offsets);
        launch_copy_rows(off, R, nnz, h->row_start, h->pool_col, h->pool_val, cols, vals, s);
        DCSR_HIP(hipGetLastError());
        DCSR_HIP(hipStreamSynchronize(s));
        *nnz_out = nnz;
        return DCSR_OK;
    }
    // canonical export of a matrix with duplicates / zeros
    int32_t *tcol, *toff;
    float *tval, *tsum;
    uint32_t *emit, *pos, *part, *tot;
    DCSR_TRY(ws_get(h, WS_UCOL, (size_t)nnz * 4 + 4, (void **)&tcol));
    DCSR_TRY(ws_get(h, WS_UVAL, (size_t)nnz * 4 + 4, (void **)&tval));
    DCSR_TRY(ws_get(h, WS_UPOS, (size_t)nnz * 4 + 4, (void **)&tsum));
    DCSR_TRY(ws_get(h, WS_UCODE, (size_t)nnz * 4 + 4, (void **)&emit));
    DCSR_TRY(ws_get(h, WS_UPRE, (size_t)nnz * 4 + 4, (void **)&pos));
    DCSR_TRY(ws_get(h, WS_TOLDPRE, (size_t)(R + 1) * 4, (void **)&toff));
    DCSR_TRY(ws_get_zeroed(h, WS_SCAN, scan_part_words(nnz + 1, 1) * 4, (void **)&part, s));
    DCSR_TRY(ws_get(h, WS_TMP2, 64, (void **)&tot));
    k_u32_to_i32_offsets<<<grid_for(R + 1, TB), TB, 0, s>>>(off, R + 1, toff);
    launch_copy_rows(off, R, nnz, h->row_start, h->pool_col, h->pool_val, tcol, tval, s);
    if (nnz > 0) {
        k_canon_flags<<<grid_for(nnz, TB), TB, 0, s>>>(toff, R, tcol, tval, nnz, emit, tsum);
        DCSR_HIP(scan_exclusive_u32(emit, pos, nnz, 1, part, tot, s, fault_word(h)));
    } else {
        DCSR_HIP(scan_exclusive_u32(emit, pos, 0, 1, part, tot, s, fault_word(h)));
    }
    uint32_t total = 0;
    DCSR_HIP(hipMemcpyAsync(&total, tot, 4, hipMemcpyDeviceToHost, s));
    DCSR_HIP(hipStreamSynchronize(s));
    if ((int64_t)total > cap) return DCSR_ERR_CAPACITY;
    if (nnz > 0) k_canon_write<<<grid_for(nnz, TB), TB, 0, s>>>(tcol, nnz, emit, pos, tsum, cols, vals);
    k_canon_offsets<<<grid_for(R + 1, TB), TB, 0, s>>>(toff, R, pos, tot, row_offsets, nnz);
    DCSR_HIP(hipGetLastError());
    DCSR_HIP(hipStreamSynchronize(s));
    *nnz_out = total;
    return DCSR_OK;
}

// Rewrite every row contiguously into a pool of capacity new_cap (>= nnz).
// The destination is the handle's spare pool (allocated only when it is too
// small); the old pool becomes the spare, so steady-state compaction makes no
// allocation and no host synchronization beyond the state read-back.
// Copy the live rows, in natural order, into the spare pool (grown to
// new_cap when smaller) and swap pools.  Needs no read-back: the copy is
// sized from the host's bound of the live entries and stops at the device's
// count; the device's bump pointer becomes its live count.
// Free the retired spare buffers (the caller knows the stream is idle).
static void release_retired(dcsr_handle h)
{
    for (int i = 0; i < h->nretired; ++i) (void)hipFree(h->retired[i]);
    h->nretired = 0;
}

static dcsr_status compact_into(dcsr_handle h, int64_t new_cap, hipStream_t s)
{
    const int64_t R = h->rows;
    const int64_t nnz = live_bound(h);
    if (new_cap < nnz) new_cap = nnz;
    if (new_cap < h->cfg.min_pool) new_cap = h->cfg.min_pool;
    if (new_cap > (int64_t)0xfffffff0ll) return DCSR_ERR_POOL_EXHAUSTED;
    uint32_t *off;
    DCSR_TRY(ws_get(h, WS_TMP0, (size_t)(R + 1) * 4, (void **)&off));
    DCSR_TRY(row_offsets_u32(h, off, s));
    if (h->spare_cap < new_cap) {
        // kernels queued earlier may still read the old spare (the pool of
        // the compaction before): it is retired, not freed, so that growing
        // never waits for the stream
        if (h->nretired + 2 > (int)(sizeof(h->retired) / sizeof(h->retired[0]))) {
            DCSR_HIP(hipStreamSynchronize(s));
            release_retired(h);
        }
        if (h->spare_col) h->retired[h->nretired++] = h->spare_col;
        if (h->spare_val) h->retired[h->nretired++] = h->spare_val;
        h->spare_col = nullptr;
        h->spare_val = nullptr;
        h->spare_cap = 0;
        DCSR_HIP(hipMalloc(&h->spare_col, (size_t)new_cap * 4));
        DCSR_HIP(hipMalloc(&h->spare_val, (size_t)new_cap * 4));
        h->spare_cap = new_cap;
    }
    launch_copy_rows(off, R, nnz, h->row_start, h->pool_col, h->pool_val, h->spare_col, h->spare_val, s);
    ++h->ncompact;
    if (R > 0) k_set_compacted<<<grid_for(R, TB), TB, 0, s>>>(R, off, h->row_start, h->row_len, h->row_cap, h->ds, off + R);
    else k_set_compacted<<<1, 1, 0, s>>>(0, off, h->row_start, h->row_len, h->row_cap, h->ds, off);
    DCSR_HIP(hipGetLastError());
    int32_t *oc = h->pool_col;
    float *ov = h->pool_val;
    int64_t ocap = h->pool_cap;
    h->pool_col = h->spare_col;
    h->pool_val = h->spare_val;
    h->pool_cap = h->spare_cap;
    h->spare_col = oc;
    h->spare_val = ov;
    h->spare_cap = ocap;
    if (h->state_known) {
        h->bump_h = nnz;
        h->garbage_h = 0;
    } else {
        // the device's bump pointer is now its live count; a snapshot in
        // flight predates the swap
        h->bump_ub = h->nnz_ub;
        h->snap_valid_from = h->snap_seq + 1;
    }
    return DCSR_OK;
}

dcsr_status compact_in_order(dcsr_handle h, hipStream_t s) { return compact_into(h, h->pool_cap, s); }

dcsr_status grow_in_order(dcsr_handle h, int64_t new_cap, hipStream_t s)
{
    if (new_cap < h->pool_cap) new_cap = h->pool_cap;
    return compact_into(h, new_cap, s);
}

dcsr_status dcsr_compact(dcsr_handle h, dcsr_stream_t stream)
{
    if (!h) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    h->last_stream = s;
    if (!h->state_known) DCSR_TRY(pull_state(h, s));
    // keep the capacity, release retired chunks
    return compact_into(h, h->pool_cap, s);
}

// Make room for at least min_free entries past the bump pointer: compact
// (releasing retired chunks) and grow geometrically when live data plus the
// request would fill more than 2/3 of the pool.
dcsr_status pool_reserve(dcsr_handle h, int64_t min_free, hipStream_t s, const int32_t *)
{
    if (!h->state_known) DCSR_TRY(pull_state(h, s));
    if (h->pool_cap - h->bump_h >= min_free) return DCSR_OK;
    int64_t cap = h->pool_cap;
    int64_t want = h->nnz_h + min_free;
    if (want + want / 2 > cap) {
        int64_t grown = (int64_t)(want * h->cfg.pool_factor);
        if (grown < 2 * cap) grown = 2 * cap;
        cap = grown;
    }
    return compact_into(h, cap, s);
}

// Rewrite a non-canonical matrix (duplicates / explicit zeros from the load)
// into canonical rows, as the first host-path add does for every row
// (cusp/system/detail/sequential/elementwise.h:44-131 visits all rows of A).
dcsr_status canonicalize(dcsr_handle h, hipStream_t s)
{
    if (h->canonical) return DCSR_OK;
    const int64_t R = h->rows;
    int64_t nnz = h->nnz_h;
    int32_t *off, *ocol;
    float *oval;
    DCSR_HIP(hipMalloc(&off, (size_t)(R + 1) * 4));
    DCSR_HIP(hipMalloc(&ocol, (size_t)(nnz > 0 ? nnz : 1) * 4));
    DCSR_HIP(hipMalloc(&oval, (size_t)(nnz > 0 ? nnz : 1) * 4));
    int64_t cn = 0;
    dcsr_status st = dcsr_export_csr(h, off, ocol, oval, nnz, &cn, DCSR_EXPORT_CANONICAL, s);
    if (st == DCSR_OK) st = dcsr_load_csr(h, off, ocol, oval, s);
    (void)hipFree(off);
    (void)hipFree(ocol);
    (void)hipFree(oval);
    if (st == DCSR_OK) h->canonical = true;
    return st;
}

dcsr_status dcsr_rebin(dcsr_handle h, dcsr_stream_t stream)
{
    if (!h) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    h->last_stream = s;
    h->flat_dirty = true;
    DCSR_TRY(ensure_flat(h, s));
    k_zero_hubs<<<1, 1, 0, s>>>(h->ds);
    if (h->rows > 0) k_count_hubs<<<grid_for(h->rows, TB), TB, 0, s>>>(h->rows, h->row_len, h->ds);
    DCSR_HIP(hipGetLastError());
    DCSR_TRY(pull_state(h, s));
    h->hub_rows_h = h->hs->hub_count;
    h->hubs_dirty = false;
    return DCSR_OK;
}

dcsr_status dcsr_query(dcsr_handle h, dcsr_info *info)
{
    if (!h || !info) return DCSR_ERR_INVALID;
    DCSR_TRY(pull_state(h, h->last_stream));
    memset(info, 0, sizeof(*info));
    info->rows = h->rows;
    info->cols = h->cols;
    info->nnz = h->nnz_h;
    info->pool_cap = h->pool_cap;
    info->pool_used = h->bump_h;
    info->garbage = h->garbage_h;
    info->hub_rows = h->hubs_dirty ? -1 : h->hub_rows_h;
    info->last_inserted = h->last_ins;
    info->last_added = h->last_add;
    info->last_deleted = h->last_del;
    info->last_touched_rows = h->last_touched;
    info->canonical = h->canonical ? 1 : 0;
    info->max_segments = 1;
    info->scan_fallback = (h->hs->status & ST_SCAN_FALLBACK) ? 1 : 0;
    info->spmv_paneled = h->pnl_valid ? 1 : 0;
    info->compactions = h->ncompact;
    return DCSR_OK;
}

dcsr_status dcsr_export_layout(dcsr_handle h, uint32_t *row_start, int32_t *row_len, int32_t *row_cap,
                               dcsr_stream_t stream)
{
    if (!h || (h->rows > 0 && (!row_start || !row_len || !row_cap))) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    h->last_stream = s;
    const size_t R = (size_t)h->rows;
    if (R > 0) {
        DCSR_HIP(hipMemcpyAsync(row_start, h->row_start, R * 4, hipMemcpyDeviceToDevice, s));
        DCSR_HIP(hipMemcpyAsync(row_len, h->row_len, R * 4, hipMemcpyDeviceToDevice, s));
        DCSR_HIP(hipMemcpyAsync(row_cap, h->row_cap, R * 4, hipMemcpyDeviceToDevice, s));
    }
    return pull_state(h, s);
}

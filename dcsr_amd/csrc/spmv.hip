// spmv.hip -- y = alpha * A x + beta * y on the DCSR layout.
//
// Replaces device::spmv(const dcsr&, x, y, streams) (include/spmv.h:121-248)
// and its seven binned kernels on seven streams (include/spmv.inl:364-622),
// which read a row length table that is stale after an update.  Here:
//
//  * k_spmv_rows: one wave per 64 consecutive rows.  The wave prefix-sums
//    the 64 live row lengths (DPP scan), then streams the concatenated
//    entries 64 at a time; each lane finds its row by a 6-step search in LDS,
//    loads (col, val) and gathers x, and a DPP segmented scan folds the
//    products per row.  No binning pass is needed and short rows are packed
//    64-per-wave, so loads stay coalesced for any length mix up to the hub
//    threshold.
//  * k_spmv_hub: rows longer than DCSR_HUB_LEN are cut into DCSR_HUB_CHUNK
//    pieces; each workgroup reduces one piece in a fixed order into a partial
//    and the last piece to finish sums the partials in piece order.
//
// Every reduction has a fixed association order, so y is bitwise
// reproducible for a given matrix (no float atomics).
#include "dcsr_internal.h"
#include "segments.h"

namespace {

constexpr int TB = 256;

// Row source: DCSR (start, len) or plain CSR offsets.
struct DcsrRows {
    const uint32_t *start;
    const int32_t *len;
    __device__ __forceinline__ void get(uint32_t r, uint32_t &s, uint32_t &l) const
    {
        s = start[r];
        l = (uint32_t)len[r];
    }
};
struct CsrRows {
    const int32_t *off;
    __device__ __forceinline__ void get(uint32_t r, uint32_t &s, uint32_t &l) const
    {
        int32_t a = off[r];
        s = (uint32_t)a;
        l = (uint32_t)(off[r + 1] - a);
    }
};

template <class Rows, bool SKIP_HUBS>
__global__ __launch_bounds__(TB) void k_spmv_rows(Rows rows, uint32_t R, const int32_t *__restrict__ col,
                                                  const float *__restrict__ val, const float *__restrict__ x,
                                                  float *__restrict__ y, float alpha, float beta)
{
    __shared__ uint32_t s_ex[TB / 64][64];
    __shared__ uint32_t s_st[TB / 64][64];
    __shared__ float s_acc[TB / 64][64];
    const int w = threadIdx.x >> 6;
    const int lane = lane_id();
    const uint32_t r = (blockIdx.x * (TB / 64) + w) * 64u + lane;
    uint32_t st = 0, L = 0;
    bool hub = false;
    if (r < R) {
        rows.get(r, st, L);
        if (SKIP_HUBS && L > DCSR_HUB_LEN) {
            hub = true;
            L = 0;
        }
    }
    const uint32_t inc = wave_incl_scan_u32(L);
    const uint32_t ex = inc - L;
    const uint32_t tot = wave_read_lane_u32(inc, 63);
    s_ex[w][lane] = ex;
    s_st[w][lane] = st;
    s_acc[w][lane] = 0.0f;
    __builtin_amdgcn_wave_barrier();
    for (uint32_t b = 0; b < tot; b += 64) {
        const uint32_t e = b + lane;
        const bool valid = e < tot;
        uint32_t j = 64;
        float p = 0.0f;
        if (valid) {
            j = (uint32_t)find_seg64(s_ex[w], e);
            uint64_t pos = (uint64_t)s_st[w][j] + (e - s_ex[w][j]);
            p = val[pos] * x[col[pos]];
        }
        p = wave_seg_incl_scan_f32(p, j);
        uint32_t jn = __shfl_down(j, 1);
        bool last = (lane == 63) || (jn != j);
        if (valid && last) s_acc[w][j] += p;
        __builtin_amdgcn_wave_barrier();
    }
    if (r < R && !hub) {
        float a = s_acc[w][lane];
        y[r] = beta == 0.0f ? alpha * a : alpha * a + beta * y[r];
    }
}

__global__ __launch_bounds__(TB) void k_spmv_hub(const DevState *__restrict__ ds, const uint32_t *__restrict__ hub_rows,
                                                 const uint32_t *__restrict__ hub_pre, uint32_t *__restrict__ hub_done,
                                                 float *__restrict__ part, const uint32_t *__restrict__ row_start,
                                                 const int32_t *__restrict__ row_len, const int32_t *__restrict__ col,
                                                 const float *__restrict__ val, const float *__restrict__ x,
                                                 float *__restrict__ y, float alpha, float beta)
{
    __shared__ float s_w[TB / 64];
    __shared__ uint32_t s_last;
    const uint32_t nh = ds->hub_count;
    const uint32_t nc = ds->hub_chunks;
    for (uint32_t c = blockIdx.x; c < nc; c += gridDim.x) {
        uint32_t lo = 0, hi = nh;  // largest h with hub_pre[h] <= c
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (hub_pre[mid] <= c) lo = mid; else hi = mid;
        }
        const uint32_t h = lo;
        const uint32_t r = hub_rows[h];
        const uint32_t L = (uint32_t)row_len[r];
        const uint64_t st = row_start[r];
        const uint32_t k0 = (c - hub_pre[h]) * DCSR_HUB_CHUNK;
        const uint32_t k1 = k0 + DCSR_HUB_CHUNK < L ? k0 + DCSR_HUB_CHUNK : L;
        float acc = 0.0f;
        for (uint32_t k = k0 + threadIdx.x; k < k1; k += TB) acc = acc + val[st + k] * x[col[st + k]];
        acc = wave_sum_f32(acc);
        if (lane_id() == 0) s_w[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            float tot = ((s_w[0] + s_w[1]) + (s_w[2] + s_w[3]));
            __hip_atomic_store(&part[c], tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence();
            uint32_t nchunks = hub_pre[h + 1] - hub_pre[h];
            uint32_t done = atomicAdd(&hub_done[h], 1u);
            s_last = (done + 1 == nchunks) ? 1u : 0u;
            if (s_last) {
                __threadfence();
                float sum = 0.0f;
                for (uint32_t q = hub_pre[h]; q < hub_pre[h + 1]; ++q)
                    sum = sum + __hip_atomic_load(&part[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                y[r] = beta == 0.0f ? alpha * sum : alpha * sum + beta * y[r];
                hub_done[h] = 0;
            }
        }
        __syncthreads();
    }
}

}  // namespace

dcsr_status dcsr_spmv(dcsr_handle h, const float *x, float *y, float alpha, float beta, dcsr_stream_t stream)
{
    if (!h || (h->rows > 0 && (!x || !y))) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    h->last_stream = s;
    if (h->rows == 0) return DCSR_OK;
    DCSR_TRY(ensure_hub_list(h, s));
    const uint32_t R = (uint32_t)h->rows;
    DcsrRows rows{h->row_start, h->row_len};
    k_spmv_rows<DcsrRows, true><<<grid_for(R, TB), TB, 0, s>>>(rows, R, h->pool_col, h->pool_val, x, y, alpha, beta);
    if (h->hub_chunks_h > 0) {
        float *part;
        DCSR_TRY(ws_get(h, WS_HUB_PART, (size_t)h->hub_chunks_h * 4, (void **)&part));
        unsigned g = (unsigned)(h->hub_chunks_h < 4096 ? h->hub_chunks_h : 4096);
        k_spmv_hub<<<g, TB, 0, s>>>(h->ds, (const uint32_t *)h->ws.ptr[WS_HUB_ROWS], (const uint32_t *)h->ws.ptr[WS_HUB_PRE],
                                    (uint32_t *)h->ws.ptr[WS_HUB_DONE], part, h->row_start, h->row_len, h->pool_col,
                                    h->pool_val, x, y, alpha, beta);
    }
    DCSR_HIP(hipGetLastError());
    return DCSR_OK;
}

dcsr_status dcsr_csr_spmv(int64_t rows, const int32_t *row_offsets, const int32_t *cols, const float *vals,
                          const float *x, float *y, dcsr_stream_t stream)
{
    if (rows < 0 || rows >= (1ll << 31)) return DCSR_ERR_INVALID;
    if (rows == 0) return DCSR_OK;
    if (!row_offsets || !x || !y) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    CsrRows rr{row_offsets};
    k_spmv_rows<CsrRows, false><<<grid_for(rows, TB), TB, 0, s>>>(rr, (uint32_t)rows, cols, vals, x, y, 1.0f, 0.0f);
    DCSR_HIP(hipGetLastError());
    return DCSR_OK;
}

// spmv.hip -- y = alpha * A x + beta * y on the DCSR layout.
//
// Replaces device::spmv(const dcsr&, x, y, streams) (include/spmv.h:121-248)
// and its seven row-length-binned kernels on seven streams
// (include/spmv.inl:364-622), which read a row-length table that goes stale
// after an update.
//
// Design: merge-based ("flat") SpMV.  The rows' live lengths are
// prefix-summed once per structural change (foff, natural row order).  Each
// wave owns a fixed range of DCSR_SPMV_PW consecutive nonzeros of that flat
// order, whatever rows they belong to, so power-law hubs and empty rows cost
// the same per nonzero.  A wave walks its range 64 rows at a time: lanes find
// their row by a 6-step search over the window's offsets in LDS, load
// (col, val) from the row's pool chunk, gather x, and a DPP segmented scan
// folds the products per row; four 64-wide chunks are in flight per
// iteration.  A row that starts inside the range is written by that wave;
// the partial of a row that started in an earlier wave is left in a carry
// slot, and a second small kernel adds each row's carry chain in wave order.
// All reductions have a fixed association order: y is bitwise reproducible.
#include "dcsr_internal.h"
#include "prims.h"
#include "segments.h"

namespace {

constexpr int TB = 256;
constexpr int ILP = 4;
constexpr uint32_t NONE = 0xffffffffu;

__device__ __forceinline__ float combine(float alpha, float acc, float beta, const float *y, uint32_t r)
{
    return beta == 0.0f ? alpha * acc : alpha * acc + beta * y[r];
}

__global__ __launch_bounds__(TB) void k_spmv_flat(uint32_t R, const uint32_t *__restrict__ foff,
                                                  const uint32_t *__restrict__ row_start,
                                                  const int32_t *__restrict__ col, const float *__restrict__ val,
                                                  const float *__restrict__ x, float *__restrict__ y, float alpha,
                                                  float beta, uint32_t per_wave, uint32_t *__restrict__ carry_row,
                                                  float *__restrict__ carry_val)
{
    __shared__ uint32_t s_off[TB / 64][65];
    __shared__ uint32_t s_st[TB / 64][64];
    __shared__ float s_acc[TB / 64][64];
    const int w = threadIdx.x >> 6;
    const int lane = lane_id();
    const uint32_t wave = blockIdx.x * (TB / 64) + w;
    const uint32_t total = foff[R];
    const uint32_t lastw = total ? (total - 1) / per_wave : 0;
    if (lane == 0) carry_row[wave] = NONE;
    if (wave > lastw) return;
    const uint32_t e0 = wave * per_wave;
    const uint32_t e1 = (total - e0) < per_wave ? total : e0 + per_wave;
    const bool last = wave == lastw;

    // first row not entirely before e0 (smallest r with foff[r+1] >= e0) =
    // largest i in [0, R] with foff[i] < e0, or 0 when e0 == 0
    uint32_t j = e0 == 0 ? 0u : wave_find_last_le(foff, R + 1, e0 - 1);
    while (j < R) {
        const uint32_t r = j + lane;
        s_off[w][lane] = r <= R ? foff[r] : total;
        if (lane == 0) s_off[w][64] = (j + 64 <= R) ? foff[j + 64] : total;
        s_st[w][lane] = r < R ? row_start[r] : 0u;
        s_acc[w][lane] = 0.0f;
        __builtin_amdgcn_wave_barrier();
        const uint32_t wb = s_off[w][0] > e0 ? s_off[w][0] : e0;
        const uint32_t wend = s_off[w][64] < e1 ? s_off[w][64] : e1;
        for (uint32_t b = wb; b < wend; b += 64 * ILP) {
            uint32_t jj[ILP], e[ILP];
            uint64_t pos[ILP];
            int32_t c[ILP];
            float v[ILP], xv[ILP];
#pragma unroll
            for (int u = 0; u < ILP; ++u) {
                e[u] = b + u * 64 + lane;
                jj[u] = 64;
                pos[u] = 0;
                if (e[u] < wend) {
                    jj[u] = (uint32_t)find_seg64(s_off[w], e[u]);
                    pos[u] = (uint64_t)s_st[w][jj[u]] + (e[u] - s_off[w][jj[u]]);
                }
            }
#pragma unroll
            for (int u = 0; u < ILP; ++u) {
                c[u] = 0;
                v[u] = 0.0f;
                if (jj[u] < 64) {
                    c[u] = col[pos[u]];
                    v[u] = val[pos[u]];
                }
            }
#pragma unroll
            for (int u = 0; u < ILP; ++u) xv[u] = jj[u] < 64 ? x[c[u]] : 0.0f;
#pragma unroll
            for (int u = 0; u < ILP; ++u) {
                float p = v[u] * xv[u];
                p = wave_seg_incl_scan_f32(p, jj[u]);
                uint32_t jn = __shfl_down(jj[u], 1);
                if (jj[u] < 64 && (lane == 63 || jn != jj[u])) s_acc[w][jj[u]] += p;
                __builtin_amdgcn_wave_barrier();
            }
        }
        // flush the window's rows
        if (r < R) {
            const uint32_t rs = s_off[w][lane], re = s_off[w][lane + 1];
            const float a = s_acc[w][lane];
            if (re > rs) {
                if (rs >= e0 && rs < e1) y[r] = combine(alpha, a, beta, y, r);
                else if (rs < e0 && re > e0) {
                    carry_row[wave] = r;
                    carry_val[wave] = a;
                }
            } else if ((rs >= e0 && rs < e1) || (last && rs == total)) {
                y[r] = beta == 0.0f ? 0.0f : beta * y[r];
            }
        }
        const uint32_t next_start = s_off[w][64];
        __builtin_amdgcn_wave_barrier();
        j += 64;
        if (!(next_start < e1 || (last && j < R))) break;
    }
}

// Each carry chain (consecutive waves carrying into the same row) is summed
// by one wave in a fixed order and added to the row's head value.
__global__ __launch_bounds__(TB) void k_spmv_carry(uint32_t nwaves, const uint32_t *__restrict__ carry_row,
                                                   const float *__restrict__ carry_val, float *__restrict__ y,
                                                   float alpha)
{
    const uint32_t w = blockIdx.x * (TB / 64) + (threadIdx.x >> 6);
    if (w >= nwaves) return;
    const uint32_t r = carry_row[w];
    if (r == NONE || (w > 0 && carry_row[w - 1] == r)) return;
    const int lane = lane_id();
    float s = 0.0f;
    for (uint32_t k = w + lane;; k += 64) {
        bool in = k < nwaves && carry_row[k] == r;
        if (in) s = s + carry_val[k];
        if (!__ballot(in)) break;
    }
    s = wave_sum_f32(s);
    if (lane == 0) y[r] = y[r] + alpha * s;
}

// Plain CSR comparator: one wave per 64 consecutive rows (the wave packs the
// rows' entries 64 at a time), cusp CSR-vector style
// (include/spmv.inl:708-781 of the reference).
__global__ __launch_bounds__(TB) void k_csr_rows(uint32_t R, const int32_t *__restrict__ off,
                                                 const int32_t *__restrict__ col, const float *__restrict__ val,
                                                 const float *__restrict__ x, float *__restrict__ y)
{
    __shared__ uint32_t s_ex[TB / 64][64];
    __shared__ uint32_t s_st[TB / 64][64];
    __shared__ float s_acc[TB / 64][64];
    const int w = threadIdx.x >> 6;
    const int lane = lane_id();
    const uint32_t r = (blockIdx.x * (TB / 64) + w) * 64u + lane;
    uint32_t st = 0, L = 0;
    if (r < R) {
        st = (uint32_t)off[r];
        L = (uint32_t)(off[r + 1] - off[r]);
    }
    const uint32_t inc = wave_incl_scan_u32(L);
    const uint32_t tot = wave_read_lane_u32(inc, 63);
    s_ex[w][lane] = inc - L;
    s_st[w][lane] = st;
    s_acc[w][lane] = 0.0f;
    __builtin_amdgcn_wave_barrier();
    for (uint32_t b = 0; b < tot; b += 64) {
        const uint32_t e = b + lane;
        uint32_t j = 64;
        float p = 0.0f;
        if (e < tot) {
            j = (uint32_t)find_seg64(s_ex[w], e);
            uint64_t pos = (uint64_t)s_st[w][j] + (e - s_ex[w][j]);
            p = val[pos] * x[col[pos]];
        }
        p = wave_seg_incl_scan_f32(p, j);
        uint32_t jn = __shfl_down(j, 1);
        if (j < 64 && (lane == 63 || jn != j)) s_acc[w][j] += p;
        __builtin_amdgcn_wave_barrier();
    }
    if (r < R) y[r] = s_acc[w][lane];
}

}  // namespace

// Flat offsets of the live row lengths (natural row order), rebuilt lazily
// after a structural change.  This is the engine's "binning" step.
dcsr_status ensure_flat(dcsr_handle h, hipStream_t s)
{
    if (!h->flat_dirty) return DCSR_OK;
    const int64_t R = h->rows;
    uint32_t *foff, *part;
    DCSR_TRY(ws_get(h, WS_FOFF, (size_t)(R + 1) * 4, (void **)&foff));
    DCSR_TRY(ws_get(h, WS_PART, scan_part_words(R + 1, 1) * 4, (void **)&part));
    // row lengths are non-negative int32: scan them as uint32 in place
    DCSR_HIP(scan_exclusive_u32((const uint32_t *)h->row_len, foff, R, 1, part, foff + R, s));
    h->flat_dirty = false;
    return DCSR_OK;
}

dcsr_status dcsr_spmv(dcsr_handle h, const float *x, float *y, float alpha, float beta, dcsr_stream_t stream)
{
    if (!h || (h->rows > 0 && (!x || !y))) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    h->last_stream = s;
    if (h->rows == 0) return DCSR_OK;
    if (!h->state_known) DCSR_TRY(pull_state(h, s));
    DCSR_TRY(ensure_flat(h, s));
    const uint32_t R = (uint32_t)h->rows;
    const uint32_t pw = DCSR_SPMV_PW;
    int64_t nwaves = (h->nnz_h + pw - 1) / pw;
    if (nwaves < 1) nwaves = 1;
    unsigned blocks = (unsigned)((nwaves + 3) / 4);
    nwaves = (int64_t)blocks * 4;
    uint32_t *crow;
    float *cval;
    DCSR_TRY(ws_get(h, WS_CARRY_ROW, (size_t)nwaves * 4, (void **)&crow));
    DCSR_TRY(ws_get(h, WS_CARRY_VAL, (size_t)nwaves * 4, (void **)&cval));
    const uint32_t *foff = (const uint32_t *)h->ws.ptr[WS_FOFF];
    k_spmv_flat<<<blocks, TB, 0, s>>>(R, foff, h->row_start, h->pool_col, h->pool_val, x, y, alpha, beta, pw, crow,
                                      cval);
    k_spmv_carry<<<blocks, TB, 0, s>>>((uint32_t)nwaves, crow, cval, y, alpha);
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
    k_csr_rows<<<grid_for(rows, TB), TB, 0, s>>>((uint32_t)rows, row_offsets, cols, vals, x, y);
    DCSR_HIP(hipGetLastError());
    return DCSR_OK;
}

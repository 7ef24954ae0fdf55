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
#include "lookback.h"
#include "prims.h"
#include "rowscan.h"
#include "segments.h"

namespace {

constexpr int TB = 256;
constexpr int ILP = 4;
constexpr uint32_t NONE = 0xffffffffu;
constexpr uint32_t GRAN = RS_GRAN;  // row-map granule (nonzeros)

// rowmap[i] = first row not entirely before flat position i*GRAN, i.e. the
// smallest r with foff[r+1] >= i*GRAN.  Lets any GRAN-aligned range find its
// first row with one load.
__global__ void k_rowmap(uint32_t R, const uint32_t *__restrict__ foff, uint32_t *__restrict__ rowmap, uint32_t cap)
{
    const uint32_t total = foff[R];
    const uint32_t ng = total / GRAN + 1;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ng || i >= cap) return;
    const uint32_t e = i * GRAN;
    uint32_t lo = 0, hi = R;  // smallest r in [0, R) with foff[r+1] >= e
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (foff[mid + 1] >= e) hi = mid; else lo = mid + 1;
    }
    rowmap[i] = lo;
}

// Flat offsets and row map of the row lengths (rowscan.h), one tile per
// workgroup.  Measured on 325K rows against 8 and 32 rows per thread: 55 vs
// 57 and 59 us per SpMV after an update.
__global__ __launch_bounds__(TB) void k_scan_rows(uint32_t R, const int32_t *__restrict__ len,
                                                  uint32_t *__restrict__ foff, uint32_t *__restrict__ rowmap,
                                                  uint32_t *__restrict__ part, unsigned int *fault, int limit)
{
    __shared__ RowScanLds L;
    auto lens = [&](int64_t r0, bool full, uint32_t(&v)[RS_IPT]) {
        if (full) {
            const int4 *p = (const int4 *)(len + r0);
#pragma unroll
            for (int q = 0; q < RS_IPT / 4; ++q) {
                const int4 t = p[q];
                v[4 * q] = (uint32_t)t.x;
                v[4 * q + 1] = (uint32_t)t.y;
                v[4 * q + 2] = (uint32_t)t.z;
                v[4 * q + 3] = (uint32_t)t.w;
            }
        } else {
#pragma unroll
            for (int j = 0; j < RS_IPT; ++j) v[j] = r0 + j < (int64_t)R ? (uint32_t)len[r0 + j] : 0u;
        }
    };
    // fallback: the lengths of every earlier tile's rows
    auto prefix = [&](int64_t n) { return lb_wave_sum_prefix((const uint32_t *)len, n, 1); };
    rowscan_tile(R, lens, prefix, foff, rowmap, part, fault, limit, gridDim.x, L);
}

// An entry load; NT: non-temporal, so that streaming the matrix does not
// evict the dense operand from L2 (used when that operand exceeds one XCD's
// L2; where it fits, plain loads also keep the pool warm for the next update).
template <bool NT, typename T>
__device__ __forceinline__ T ld_entry(const T *p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

__device__ __forceinline__ float combine(float alpha, float acc, float beta, const float *y, uint64_t i)
{
    return beta == 0.0f ? alpha * acc : alpha * acc + beta * y[i];
}

// Flat kernel for a column block of width KB (1: SpMV, 2 or 4: narrow SpMM).
// X is row-major with leading dimension ldx, Y with ldy (SpMV: both 1).
// NT: non-temporal entry loads (ld_entry); on the paneled copy of a
// 10M-row / 160M-entry matrix 1.87 -> 1.80 ms per SpMV.
template <int KB, bool NT = false>
__global__ __launch_bounds__(TB) void k_flat(uint32_t R, const uint32_t *__restrict__ foff,
                                             const uint32_t *__restrict__ rowmap,
                                             const uint32_t *__restrict__ row_start, const int32_t *__restrict__ col,
                                             const float *__restrict__ val, const float *__restrict__ X, int64_t ldx,
                                             float *__restrict__ Y, int64_t ldy, float alpha, float beta,
                                             uint32_t per_wave, uint32_t *__restrict__ carry_row,
                                             float *__restrict__ carry_val)
{
    __shared__ uint32_t s_off[TB / 64][65];
    __shared__ uint32_t s_st[TB / 64][64];
    __shared__ float s_acc[TB / 64][KB][64];
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

    // first row not entirely before e0 (per_wave is a multiple of GRAN)
    uint32_t j = rowmap[e0 / GRAN];
    while (j < R) {
        const uint32_t r = j + lane;
        s_off[w][lane] = r <= R ? foff[r] : total;
        if (lane == 0) s_off[w][64] = (j + 64 <= R) ? foff[j + 64] : total;
        s_st[w][lane] = r < R ? row_start[r] : 0u;
#pragma unroll
        for (int q = 0; q < KB; ++q) s_acc[w][q][lane] = 0.0f;
        __builtin_amdgcn_wave_barrier();
        const uint32_t wb = s_off[w][0] > e0 ? s_off[w][0] : e0;
        const uint32_t wend = s_off[w][64] < e1 ? s_off[w][64] : e1;
        for (uint32_t b = wb; b < wend; b += 64 * ILP) {
            uint32_t jj[ILP];
            uint64_t pos[ILP];
            int32_t c[ILP];
            float v[ILP], xv[ILP][KB];
#pragma unroll
            for (int u = 0; u < ILP; ++u) {
                const uint32_t e = b + u * 64 + lane;
                jj[u] = 64;
                pos[u] = 0;
                if (e < wend) {
                    jj[u] = (uint32_t)find_seg64(s_off[w], e);
                    pos[u] = (uint64_t)s_st[w][jj[u]] + (e - s_off[w][jj[u]]);
                }
            }
#pragma unroll
            for (int u = 0; u < ILP; ++u) {
                c[u] = 0;
                v[u] = 0.0f;
                if (jj[u] < 64) {
                    c[u] = ld_entry<NT>(&col[pos[u]]);
                    v[u] = ld_entry<NT>(&val[pos[u]]);
                }
            }
#pragma unroll
            for (int u = 0; u < ILP; ++u)
#pragma unroll
                for (int q = 0; q < KB; ++q) xv[u][q] = jj[u] < 64 ? X[(int64_t)c[u] * ldx + q] : 0.0f;
#pragma unroll
            for (int u = 0; u < ILP; ++u) {
                const uint32_t jn = __shfl_down(jj[u], 1);
                const bool tail = jj[u] < 64 && (lane == 63 || jn != jj[u]);
#pragma unroll
                for (int q = 0; q < KB; ++q) {
                    float p = v[u] * xv[u][q];
                    p = wave_seg_incl_scan_f32(p, jj[u]);
                    if (tail) s_acc[w][q][jj[u]] += p;
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
        // flush the window's rows
        if (r < R) {
            const uint32_t rs = s_off[w][lane], re = s_off[w][lane + 1];
            if (re > rs) {
                if (rs >= e0 && rs < e1) {
#pragma unroll
                    for (int q = 0; q < KB; ++q)
                        Y[(uint64_t)r * ldy + q] = combine(alpha, s_acc[w][q][lane], beta, Y, (uint64_t)r * ldy + q);
                } else if (rs < e0 && re > e0) {
                    carry_row[wave] = r;
#pragma unroll
                    for (int q = 0; q < KB; ++q) carry_val[(uint64_t)wave * KB + q] = s_acc[w][q][lane];
                }
            } else if ((rs >= e0 && rs < e1) || (last && rs == total)) {
#pragma unroll
                for (int q = 0; q < KB; ++q) {
                    const uint64_t i = (uint64_t)r * ldy + q;
                    Y[i] = beta == 0.0f ? 0.0f : beta * Y[i];
                }
            }
        }
        const uint32_t next_start = s_off[w][64];
        __builtin_amdgcn_wave_barrier();
        j += 64;
        if (!(next_start < e1 || (last && j < R))) break;
    }
}

// Wide column blocks (KB = 8, 16, 32) on the flat decomposition: a wave owns
// DCSR_SPMV_PW consecutive nonzeros like the flat kernel and walks them 64
// rows at a time.  The window's (col, val, row) are first staged in LDS by
// all 64 lanes (one coalesced pass, no per-row dependent loads); then the
// wave's 64/KB lane groups take contiguous slices of the staged entries,
// each of a group's KB lanes accumulating one column (the X-row gathers are
// contiguous KB*4-byte runs, several entries in flight).  Rows interior to a
// slice are written directly; a slice's first and last partial rows are
// combined across groups in slice order by group 0, so every row is summed
// in entry order.  The row that started before the wave goes to the carry
// slot, as in the flat kernel.
constexpr int WIDE_UN = 4;
template <int KB, bool NT = false>
__global__ __launch_bounds__(TB) void k_spmm_wide(uint32_t R, const uint32_t *__restrict__ foff,
                                                  const uint32_t *__restrict__ rowmap,
                                                  const uint32_t *__restrict__ row_start,
                                                  const int32_t *__restrict__ col, const float *__restrict__ val,
                                                  const float *__restrict__ X, int64_t ldx, float *__restrict__ Y,
                                                  int64_t ldy, float alpha, float beta,
                                                  uint32_t *__restrict__ carry_row, float *__restrict__ carry_val)
{
    constexpr int G = 64 / KB;
    constexpr uint32_t PW = DCSR_SPMV_PW;
    __shared__ uint32_t s_off[TB / 64][65];
    __shared__ uint32_t s_st[TB / 64][64];
    __shared__ int32_t s_col[TB / 64][PW];
    __shared__ float s_val[TB / 64][PW];
    __shared__ uint8_t s_row[TB / 64][PW];
    __shared__ float s_part[TB / 64][2][G][KB];
    __shared__ int32_t s_prow[TB / 64][2][G];
    const int w = threadIdx.x >> 6;
    const int lane = lane_id();
    const int cl = lane % KB;
    const int g = lane / KB;
    const uint32_t wave = blockIdx.x * (TB / 64) + w;
    const uint32_t total = foff[R];
    const uint32_t lastw = total ? (total - 1) / PW : 0;
    if (lane == 0) carry_row[wave] = NONE;
    if (wave > lastw) return;
    const uint32_t e0 = wave * PW;
    const uint32_t e1 = (total - e0) < PW ? total : e0 + PW;
    const bool last = wave == lastw;

    uint32_t j = rowmap[e0 / GRAN];
    while (j < R) {
        const uint32_t r = j + lane;
        s_off[w][lane] = r <= R ? foff[r] : total;
        if (lane == 0) s_off[w][64] = (j + 64 <= R) ? foff[j + 64] : total;
        s_st[w][lane] = r < R ? row_start[r] : 0u;
        __builtin_amdgcn_wave_barrier();
        const uint32_t wb = s_off[w][0] > e0 ? s_off[w][0] : e0;
        const uint32_t wend = s_off[w][64] < e1 ? s_off[w][64] : e1;
        const uint32_t n = wend > wb ? wend - wb : 0u;
        // stage the window's entries
        for (uint32_t b = 0; b < n; b += 64) {
            const uint32_t i = b + lane;
            if (i < n) {
                const uint32_t e = wb + i;
                const uint32_t jj = (uint32_t)find_seg64(s_off[w], e);
                const uint64_t pos = (uint64_t)s_st[w][jj] + (e - s_off[w][jj]);
                s_col[w][i] = ld_entry<NT>(&col[pos]);
                s_val[w][i] = ld_entry<NT>(&val[pos]);
                s_row[w][i] = (uint8_t)jj;
            }
        }
        __builtin_amdgcn_wave_barrier();
        // this group's slice of the staged entries
        const uint32_t L = (n + G - 1) / G;
        const uint32_t c0 = (uint32_t)g * L < n ? (uint32_t)g * L : n;
        const uint32_t c1 = c0 + L < n ? c0 + L : n;
        int cur = -1, nseg = 0;
        float acc = 0.0f;
        for (uint32_t i = c0; i < c1; i += WIDE_UN) {
            int32_t cc[WIDE_UN];
            float vv[WIDE_UN], xv[WIDE_UN];
            int rr[WIDE_UN];
#pragma unroll
            for (int u = 0; u < WIDE_UN; ++u) {
                const uint32_t idx = i + u;
                const bool ok = idx < c1;
                cc[u] = ok ? s_col[w][idx] : 0;
                vv[u] = ok ? s_val[w][idx] : 0.0f;
                rr[u] = ok ? (int)s_row[w][idx] : -1;
            }
#pragma unroll
            for (int u = 0; u < WIDE_UN; ++u) xv[u] = rr[u] >= 0 ? X[(int64_t)cc[u] * ldx + cl] : 0.0f;
#pragma unroll
            for (int u = 0; u < WIDE_UN; ++u) {
                if (rr[u] < 0) continue;
                if (rr[u] != cur) {
                    if (cur >= 0) {
                        if (nseg == 0) {
                            // first segment of the slice: may continue the previous slice
                            s_part[w][0][g][cl] = acc;
                            if (cl == 0) s_prow[w][0][g] = cur;
                        } else {
                            // interior row: starts and ends in this slice, starts in the wave
                            const uint64_t yi = (uint64_t)(j + cur) * ldy + cl;
                            Y[yi] = combine(alpha, acc, beta, Y, yi);
                        }
                        ++nseg;
                    }
                    cur = rr[u];
                    acc = 0.0f;
                }
                acc = acc + vv[u] * xv[u];
            }
        }
        // the slice's last segment (or its only one)
        if (nseg == 0) {
            s_part[w][0][g][cl] = acc;
            if (cl == 0) {
                s_prow[w][0][g] = cur;  // -1 for an empty slice
                s_prow[w][1][g] = -1;
            }
        } else {
            s_part[w][1][g][cl] = acc;
            if (cl == 0) s_prow[w][1][g] = cur;
        }
        __builtin_amdgcn_wave_barrier();
        // combine slice boundary partials in slice order (group 0)
        if (g == 0) {
            int crow = -1;
            float cacc = 0.0f;
            for (int gg = 0; gg < G; ++gg) {
#pragma unroll
                for (int part = 0; part < 2; ++part) {
                    const int row = s_prow[w][part][gg];
                    if (row < 0) continue;
                    const float a = s_part[w][part][gg][cl];
                    if (row == crow) {
                        cacc = cacc + a;
                        continue;
                    }
                    if (crow >= 0) {
                        const uint32_t rs = s_off[w][crow];
                        if (rs >= e0) {
                            const uint64_t yi = (uint64_t)(j + crow) * ldy + cl;
                            Y[yi] = combine(alpha, cacc, beta, Y, yi);
                        } else {
                            if (cl == 0) carry_row[wave] = j + crow;
                            carry_val[(uint64_t)wave * KB + cl] = cacc;
                        }
                    }
                    crow = row;
                    cacc = a;
                }
            }
            if (crow >= 0) {
                const uint32_t rs = s_off[w][crow];
                if (rs >= e0) {
                    const uint64_t yi = (uint64_t)(j + crow) * ldy + cl;
                    Y[yi] = combine(alpha, cacc, beta, Y, yi);
                } else {
                    if (cl == 0) carry_row[wave] = j + crow;
                    carry_val[(uint64_t)wave * KB + cl] = cacc;
                }
            }
        }
        // empty rows that start in the wave's range (and trailing empty rows)
        if (r < R) {
            const uint32_t rs = s_off[w][lane], re = s_off[w][lane + 1];
            if (re == rs && ((rs >= e0 && rs < e1) || (last && rs == total))) {
#pragma unroll
                for (int q = 0; q < KB; ++q) {
                    const uint64_t yi = (uint64_t)r * ldy + q;
                    Y[yi] = beta == 0.0f ? 0.0f : beta * Y[yi];
                }
            }
        }
        const uint32_t next_start = s_off[w][64];
        __builtin_amdgcn_wave_barrier();
        j += 64;
        if (!(next_start < e1 || (last && j < R))) break;
    }
}

// Y[0..3] = alpha * a + beta * Y[0..3] (beta == 0: Y is not read); Y is
// 16-byte aligned.
__device__ __forceinline__ void store4(float *Y, float4 a, float alpha, float beta)
{
    float4 o;
    if (beta == 0.0f) {
        o = make_float4(alpha * a.x, alpha * a.y, alpha * a.z, alpha * a.w);
    } else {
        const float4 y = *reinterpret_cast<const float4 *>(Y);
        o = make_float4(alpha * a.x + beta * y.x, alpha * a.y + beta * y.y, alpha * a.z + beta * y.z,
                        alpha * a.w + beta * y.w);
    }
    *reinterpret_cast<float4 *>(Y) = o;
}

// The same kernel with float4 column access (KB/4 lanes per group): used when
// X and Y rows are 16-byte aligned, it issues a quarter of the loads.
template <int KB, bool NT = false>
__global__ __launch_bounds__(TB) void k_spmm_wide4(uint32_t R, const uint32_t *__restrict__ foff,
                                                  const uint32_t *__restrict__ rowmap,
                                                  const uint32_t *__restrict__ row_start,
                                                  const int32_t *__restrict__ col, const float *__restrict__ val,
                                                  const float *__restrict__ X, int64_t ldx, float *__restrict__ Y,
                                                  int64_t ldy, float alpha, float beta,
                                                  uint32_t *__restrict__ carry_row, float *__restrict__ carry_val)
{
    constexpr int LPG = KB / 4;  // lanes per group, 4 columns (one float4) each
    constexpr int G = 64 / LPG;
    constexpr uint32_t PW = DCSR_SPMV_PW;
    __shared__ uint32_t s_off[TB / 64][65];
    __shared__ uint32_t s_st[TB / 64][64];
    __shared__ int32_t s_col[TB / 64][PW];
    __shared__ float s_val[TB / 64][PW];
    __shared__ uint8_t s_row[TB / 64][PW];
    __shared__ float4 s_part[TB / 64][2][G][LPG];
    __shared__ int32_t s_prow[TB / 64][2][G];
    const int w = threadIdx.x >> 6;
    const int lane = lane_id();
    const int cl = lane % LPG;
    const int g = lane / LPG;
    const uint32_t wave = blockIdx.x * (TB / 64) + w;
    const uint32_t total = foff[R];
    const uint32_t lastw = total ? (total - 1) / PW : 0;
    if (lane == 0) carry_row[wave] = NONE;
    if (wave > lastw) return;
    const uint32_t e0 = wave * PW;
    const uint32_t e1 = (total - e0) < PW ? total : e0 + PW;
    const bool last = wave == lastw;

    uint32_t j = rowmap[e0 / GRAN];
    while (j < R) {
        const uint32_t r = j + lane;
        s_off[w][lane] = r <= R ? foff[r] : total;
        if (lane == 0) s_off[w][64] = (j + 64 <= R) ? foff[j + 64] : total;
        s_st[w][lane] = r < R ? row_start[r] : 0u;
        __builtin_amdgcn_wave_barrier();
        const uint32_t wb = s_off[w][0] > e0 ? s_off[w][0] : e0;
        const uint32_t wend = s_off[w][64] < e1 ? s_off[w][64] : e1;
        const uint32_t n = wend > wb ? wend - wb : 0u;
        // stage the window's entries
        for (uint32_t b = 0; b < n; b += 64) {
            const uint32_t i = b + lane;
            if (i < n) {
                const uint32_t e = wb + i;
                const uint32_t jj = (uint32_t)find_seg64(s_off[w], e);
                const uint64_t pos = (uint64_t)s_st[w][jj] + (e - s_off[w][jj]);
                s_col[w][i] = ld_entry<NT>(&col[pos]);
                s_val[w][i] = ld_entry<NT>(&val[pos]);
                s_row[w][i] = (uint8_t)jj;
            }
        }
        __builtin_amdgcn_wave_barrier();
        // this group's slice of the staged entries
        const uint32_t L = (n + G - 1) / G;
        const uint32_t c0 = (uint32_t)g * L < n ? (uint32_t)g * L : n;
        const uint32_t c1 = c0 + L < n ? c0 + L : n;
        int cur = -1, nseg = 0;
        float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        for (uint32_t i = c0; i < c1; i += WIDE_UN) {
            int32_t cc[WIDE_UN];
            float vv[WIDE_UN];
            float4 xv[WIDE_UN];
            int rr[WIDE_UN];
#pragma unroll
            for (int u = 0; u < WIDE_UN; ++u) {
                const uint32_t idx = i + u;
                const bool ok = idx < c1;
                cc[u] = ok ? s_col[w][idx] : 0;
                vv[u] = ok ? s_val[w][idx] : 0.0f;
                rr[u] = ok ? (int)s_row[w][idx] : -1;
            }
#pragma unroll
            for (int u = 0; u < WIDE_UN; ++u)
                xv[u] = rr[u] >= 0 ? *reinterpret_cast<const float4 *>(X + (int64_t)cc[u] * ldx + 4 * cl)
                                   : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
            for (int u = 0; u < WIDE_UN; ++u) {
                if (rr[u] < 0) continue;
                if (rr[u] != cur) {
                    if (cur >= 0) {
                        if (nseg == 0) {
                            // first segment of the slice: may continue the previous slice
                            s_part[w][0][g][cl] = acc;
                            if (cl == 0) s_prow[w][0][g] = cur;
                        } else {
                            // interior row: starts and ends in this slice, starts in the wave
                            store4(Y + (uint64_t)(j + cur) * ldy + 4 * cl, acc, alpha, beta);
                        }
                        ++nseg;
                    }
                    cur = rr[u];
                    acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                }
                acc.x = acc.x + vv[u] * xv[u].x;
                acc.y = acc.y + vv[u] * xv[u].y;
                acc.z = acc.z + vv[u] * xv[u].z;
                acc.w = acc.w + vv[u] * xv[u].w;
            }
        }
        // the slice's last segment (or its only one)
        if (nseg == 0) {
            s_part[w][0][g][cl] = acc;
            if (cl == 0) {
                s_prow[w][0][g] = cur;  // -1 for an empty slice
                s_prow[w][1][g] = -1;
            }
        } else {
            s_part[w][1][g][cl] = acc;
            if (cl == 0) s_prow[w][1][g] = cur;
        }
        __builtin_amdgcn_wave_barrier();
        // combine slice boundary partials in slice order (group 0)
        if (g == 0) {
            int crow = -1;
            float4 cacc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            for (int gg = 0; gg < G; ++gg) {
#pragma unroll
                for (int part = 0; part < 2; ++part) {
                    const int row = s_prow[w][part][gg];
                    if (row < 0) continue;
                    const float4 a = s_part[w][part][gg][cl];
                    if (row == crow) {
                        cacc.x = cacc.x + a.x;
                        cacc.y = cacc.y + a.y;
                        cacc.z = cacc.z + a.z;
                        cacc.w = cacc.w + a.w;
                        continue;
                    }
                    if (crow >= 0) {
                        const uint32_t rs = s_off[w][crow];
                        if (rs >= e0) {
                            store4(Y + (uint64_t)(j + crow) * ldy + 4 * cl, cacc, alpha, beta);
                        } else {
                            if (cl == 0) carry_row[wave] = j + crow;
                            *reinterpret_cast<float4 *>(carry_val + (uint64_t)wave * KB + 4 * cl) = cacc;
                        }
                    }
                    crow = row;
                    cacc = a;
                }
            }
            if (crow >= 0) {
                const uint32_t rs = s_off[w][crow];
                if (rs >= e0) {
                    store4(Y + (uint64_t)(j + crow) * ldy + 4 * cl, cacc, alpha, beta);
                } else {
                    if (cl == 0) carry_row[wave] = j + crow;
                    *reinterpret_cast<float4 *>(carry_val + (uint64_t)wave * KB + 4 * cl) = cacc;
                }
            }
        }
        // empty rows that start in the wave's range (and trailing empty rows)
        if (r < R) {
            const uint32_t rs = s_off[w][lane], re = s_off[w][lane + 1];
            if (re == rs && ((rs >= e0 && rs < e1) || (last && rs == total))) {
#pragma unroll
                for (int q = 0; q < KB; ++q) {
                    const uint64_t yi = (uint64_t)r * ldy + q;
                    Y[yi] = beta == 0.0f ? 0.0f : beta * Y[yi];
                }
            }
        }
        const uint32_t next_start = s_off[w][64];
        __builtin_amdgcn_wave_barrier();
        j += 64;
        if (!(next_start < e1 || (last && j < R))) break;
    }
}

// Carry chains: consecutive slots (waves) carrying into the same row, added
// to the row's head value.  A wave scans 64 slots.  Chains shorter than 64
// slots: with one column (SpMV) a segmented wave scan sums every chain of
// the 128 slots from the wave's base at once (a chain whose head is in the
// wave ends within them); with KB columns, lane groups sum them (64/KB
// groups, one lane per column, one chain per group, serially in slot
// order).  A chain of 64 or more slots (a hub row spread over many waves) is
// summed by the whole wave, 64 slots at a time, then reduced in a fixed
// tree.  Either way the order depends only on the chain, so results are
// reproducible.
template <int KB>
__global__ __launch_bounds__(TB) void k_carry(uint32_t nunits, const uint32_t *__restrict__ carry_row,
                                              const float *__restrict__ carry_val, float *__restrict__ Y,
                                              int64_t ldy, float alpha)
{
    constexpr int G = 64 / KB;
    const int lane = lane_id();
    const int cl = lane % KB;
    const int g = lane / KB;
    const uint64_t base = ((uint64_t)blockIdx.x * (TB / 64) + (threadIdx.x >> 6)) * 64;
    if (base >= nunits) return;
    const uint64_t w = base + lane;
    uint32_t r = NONE;
    bool head = false, lng = false;
    if constexpr (KB == 1) {
        // every load first: this wave's slots, the next 64, the slot before
        const uint64_t w2 = w + 64;
        r = w < nunits ? carry_row[w] : NONE;
        const uint32_t r2 = w2 < nunits ? carry_row[w2] : NONE;
        const uint32_t rb = base > 0 ? carry_row[base - 1] : NONE;
        float v = w < nunits ? carry_val[w] : 0.0f;
        float v2 = w2 < nunits ? carry_val[w2] : 0.0f;
        if (r == NONE) v = 0.0f;
        if (r2 == NONE) v2 = 0.0f;
        // every lane takes part in the shuffle (a ternary would skip lane 0,
        // and a shuffle from an inactive lane reads 0)
        const uint32_t ru = (uint32_t)__shfl_up((int)r, 1);
        const uint32_t rl = lane == 0 ? rb : ru;
        head = r != NONE && rl != r;
        // slot w + 63: lane 63 of this wave for lane 0, else lane - 1 of the next
        const uint32_t r63 = (uint32_t)__shfl((int)r, 63);
        const uint32_t r2p = (uint32_t)__shfl((int)r2, (lane + 63) & 63);
        lng = head && (lane == 0 ? r63 : r2p) == r;
        const uint64_t longs = __ballot(lng);
        const uint64_t heads = __ballot(head);
        const uint64_t starts = __ballot(r != NONE && (lane == 0 || rl != r));
        const float s1 = wave_seg_incl_scan_f32(v, r);
        const float s2 = wave_seg_incl_scan_f32(v2, r2);
        // the sum of the next 64 slots' first run (a chain continuing past this wave)
        const uint32_t r20 = (uint32_t)__shfl((int)r2, 0);
        const int e2 = __popcll(__ballot(r2 == r20)) - 1;
        const float t2 = __shfl(s2, e2 < 0 ? 0 : e2);
        const uint32_t rn = (uint32_t)__shfl_down((int)r, 1);
        if (r != NONE && (lane == 63 || rn != r)) {
            // this lane ends a run; the run started at the highest start at or below it
            const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
            const int f = 63 - __clzll(starts & upto);
            if (((heads >> f) & 1ull) && !((longs >> f) & 1ull)) {
                const float acc = lane == 63 && r20 == r ? s1 + t2 : s1;
                const uint64_t yi = (uint64_t)r * ldy;
                Y[yi] = Y[yi] + alpha * acc;
            }
        }
    } else {
        if (w < nunits) {
            r = carry_row[w];
            head = r != NONE && (w == 0 || carry_row[w - 1] != r);
            // chains are contiguous: 64+ slots iff slot w+63 still carries r
            lng = head && w + 63 < nunits && carry_row[w + 63] == r;
        }
        const uint64_t longs = __ballot(lng);
        const uint64_t shorts = __ballot(head) & ~longs;
        const int nshort = __popcll(shorts);
        for (int i0 = 0; i0 < nshort; i0 += G) {
            const int i = i0 + g;
            // the i-th short head of this wave
            uint64_t m = shorts;
            for (int t = 0; t < i && m; ++t) m &= m - 1;
            const int hl = m ? __ffsll((long long)m) - 1 : 0;
            const uint32_t hr = (uint32_t)__shfl((int)r, hl);
            if (i < nshort) {
                const uint64_t hw = base + hl;
                float acc = 0.0f;
                for (uint64_t k = hw; k < nunits && carry_row[k] == hr; ++k) acc = acc + carry_val[k * KB + cl];
                const uint64_t yi = (uint64_t)hr * ldy + cl;
                Y[yi] = Y[yi] + alpha * acc;
            }
        }
    }
    uint64_t lm = __ballot(lng);
    while (lm) {
        const int hl = __ffsll((long long)lm) - 1;
        lm &= lm - 1;
        const uint32_t hr = (uint32_t)__shfl((int)r, hl);
        float sacc[KB];
#pragma unroll
        for (int q = 0; q < KB; ++q) sacc[q] = 0.0f;
        for (uint64_t k = base + hl + lane;; k += 64) {
            const bool in = k < nunits && carry_row[k] == hr;
            if (in) {
#pragma unroll
                for (int q = 0; q < KB; ++q) sacc[q] = sacc[q] + carry_val[k * KB + q];
            }
            if (__ballot(in) != ~0ull) break;
        }
#pragma unroll
        for (int q = 0; q < KB; ++q) {
            const float t = wave_sum_f32(sacc[q]);
            if (lane == 0) {
                const uint64_t yi = (uint64_t)hr * ldy + q;
                Y[yi] = Y[yi] + alpha * t;
            }
        }
    }
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

// ---------------------------------------------------------------------------
// Column panels.  With uniformly spread columns, every nonzero's x[col]
// gather misses a 4 MB L2 once x is much larger, and costs a whole 128-byte
// line from the last-level cache or HBM.  The paneled copy regroups the
// entries of each row by column panel (PANEL consecutive columns): panel p
// holds, for every row, the row's entries with columns in panel p, in
// column order.  Seen as one "virtual" matrix of P*(R+1) rows (row p*(R+1)+r
// = row r's panel-p entries, row p*(R+1)+R an empty separator), its flat
// order is panel-major, so the waves resident at any time work inside one
// panel and its slice of x stays in L2.  The flat kernel writes one partial
// per virtual row; a reduction adds a row's partials in panel order.
// ---------------------------------------------------------------------------

// Per row: boundaries of its panel segments (binary searches in the row's
// sorted columns); cnt[p*(R+1)+r] = entries of row r in panel p (separator
// slots 0), bnd[p*R+r] = index of the first of them within the row.
__global__ void k_pnl_bounds(uint32_t R, int P, int64_t panel, const uint32_t *__restrict__ row_start,
                             const int32_t *__restrict__ row_len, const int32_t *__restrict__ col,
                             uint32_t *__restrict__ cnt, uint32_t *__restrict__ bnd)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r == 0)
        for (int p = 0; p < P; ++p) cnt[(uint64_t)p * (R + 1) + R] = 0;
    if (r >= R) return;
    const uint64_t st = row_start[r];
    const int32_t L = row_len[r];
    int32_t prev = 0;
    for (int p = 0; p < P; ++p) {
        int32_t hi = L;
        if (p + 1 < P) {
            const int64_t lim = (int64_t)(p + 1) * panel;
            int32_t lo = prev, h2 = L;  // first entry with col >= lim
            while (lo < h2) {
                const int32_t mid = (lo + h2) >> 1;
                if ((int64_t)col[st + mid] < lim) lo = mid + 1; else h2 = mid;
            }
            hi = lo;
        }
        cnt[(uint64_t)p * (R + 1) + r] = (uint32_t)(hi - prev);
        bnd[(uint64_t)p * R + r] = (uint32_t)prev;
        prev = hi;
    }
}

// Copy every live entry to its panel position: poff = exclusive scan of the
// per-(panel, row) counts.
__global__ void k_pnl_fill(uint32_t nnz, uint32_t R, int64_t panel, const uint32_t *__restrict__ foff,
                           const uint32_t *__restrict__ rowmap, const uint32_t *__restrict__ row_start,
                           const int32_t *__restrict__ col, const float *__restrict__ val,
                           const uint32_t *__restrict__ poff, const uint32_t *__restrict__ bnd,
                           int32_t *__restrict__ pcol, float *__restrict__ pval)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz || e >= foff[R]) return;
    uint32_t r = rowmap[e / GRAN];
    while (foff[r + 1] <= e) ++r;
    const uint32_t k = e - foff[r];
    const uint64_t pos = (uint64_t)row_start[r] + k;
    const int32_t c = col[pos];
    const uint32_t p = (uint32_t)((int64_t)c / panel);
    const uint64_t dst = (uint64_t)poff[(uint64_t)p * (R + 1) + r] + (k - bnd[(uint64_t)p * R + r]);
    pcol[dst] = c;
    pval[dst] = val[pos];
}

// y[r] = alpha * (sum of row r's panel partials, in panel order) + beta * y[r]
__global__ void k_pnl_reduce(uint32_t R, int P, const float *__restrict__ part, float *__restrict__ y, float alpha,
                             float beta)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    float acc = 0.0f;
    for (int p = 0; p < P; ++p) acc = acc + part[(uint64_t)p * (R + 1) + r];
    y[r] = combine(alpha, acc, beta, y, r);
}

}  // namespace

// Carry slots for any nnz up to the pool capacity (one per wave of
// DCSR_SPMV_PW entries, rounded to whole blocks), so the buffers are not
// reallocated while the matrix grows within its pool.
static size_t carry_slots(dcsr_handle h)
{
    const int64_t cap = h->pool_cap > h->nnz_h ? h->pool_cap : h->nnz_h;
    const int64_t waves = (cap + DCSR_SPMV_PW - 1) / DCSR_SPMV_PW + 4;
    return (size_t)((waves + 3) / 4 * 4);
}

// Flat offsets of the live row lengths (natural row order), rebuilt lazily
// after a structural change.  This is the engine's "binning" step.
dcsr_status flat_buffers(dcsr_handle h, hipStream_t s, FlatBufs *fb)
{
    const int64_t R = h->rows;
    DCSR_TRY(ws_get(h, WS_FOFF, (size_t)(R + 1) * 4, (void **)&fb->foff));
    // look-back block: 4 control words and one 64-bit status word per tile
    DCSR_TRY(ws_get_zeroed(h, WS_PART, rowscan_part_bytes(R), (void **)&fb->part, s));
    // the row map covers the live entries, which the pool bounds: sized by
    // the pool, it needs no reallocation while the matrix grows within it
    // (no read-back, also after NOSYNC updates)
    const int64_t live = live_bound(h);
    const int64_t ngcap = (h->pool_cap > live ? h->pool_cap : live) / GRAN + 1;
    DCSR_TRY(ws_get(h, WS_ROWMAP, (size_t)ngcap * 4, (void **)&fb->rowmap));
    return DCSR_OK;
}

void flat_rebuilt(dcsr_handle h)
{
    h->pnl_valid = false;
    if (h->pnl_calls > 0) h->pnl_prev_calls = h->pnl_calls;
    h->pnl_calls = 0;
    h->flat_dirty = false;
}

dcsr_status ensure_flat(dcsr_handle h, hipStream_t s)
{
    if (!h->flat_dirty) return DCSR_OK;
    const int64_t R = h->rows;
    FlatBufs fb;
    DCSR_TRY(flat_buffers(h, s, &fb));
    if (R > 0)
        k_scan_rows<<<(unsigned)rowscan_tiles(R), TB, 0, s>>>((uint32_t)R, h->row_len, fb.foff, fb.rowmap, fb.part,
                                                              fault_word(h), lb_spin_limit());
    DCSR_HIP(hipGetLastError());
    flat_rebuilt(h);
    return DCSR_OK;
}

// Stream the matrix with non-temporal loads when the dense operand's
// columns (kb floats per column of A) exceed one XCD's 4 MB L2.
static bool stream_entries(dcsr_handle h, int kb) { return h->cols * (int64_t)kb * 4 > (4ll << 20); }

template <int KB>
static dcsr_status launch_flat(dcsr_handle h, const float *X, int64_t ldx, float *Y, int64_t ldy, float alpha,
                               float beta, hipStream_t s)
{
    const uint32_t R = (uint32_t)h->rows;
    const uint32_t pw = DCSR_SPMV_PW;
    int64_t nwaves = (live_bound(h) + pw - 1) / pw;
    if (nwaves < 1) nwaves = 1;
    const unsigned blocks = (unsigned)((nwaves + 3) / 4);
    nwaves = (int64_t)blocks * 4;
    uint32_t *crow;
    float *cval;
    const size_t cap_slots = carry_slots(h);
    DCSR_TRY(ws_get(h, WS_CARRY_ROW, cap_slots * 4, (void **)&crow));
    DCSR_TRY(ws_get(h, WS_CARRY_VAL, cap_slots * 4 * KB, (void **)&cval));
    const uint32_t *foff = (const uint32_t *)h->ws.ptr[WS_FOFF];
    const uint32_t *rowmap = (const uint32_t *)h->ws.ptr[WS_ROWMAP];
    if (stream_entries(h, KB))
        k_flat<KB, true><<<blocks, TB, 0, s>>>(R, foff, rowmap, h->row_start, h->pool_col, h->pool_val, X, ldx, Y,
                                               ldy, alpha, beta, pw, crow, cval);
    else
        k_flat<KB><<<blocks, TB, 0, s>>>(R, foff, rowmap, h->row_start, h->pool_col, h->pool_val, X, ldx, Y, ldy,
                                         alpha, beta, pw, crow, cval);
    k_carry<KB><<<grid_for(nwaves, TB), TB, 0, s>>>((uint32_t)nwaves, crow, cval, Y, ldy, alpha);
    DCSR_HIP(hipGetLastError());
    return DCSR_OK;
}

template <int KB>
static dcsr_status launch_wide(dcsr_handle h, const float *X, int64_t ldx, float *Y, int64_t ldy, float alpha,
                               float beta, hipStream_t s)
{
    const uint32_t R = (uint32_t)h->rows;
    int64_t nwaves = (live_bound(h) + DCSR_SPMV_PW - 1) / DCSR_SPMV_PW;
    if (nwaves < 1) nwaves = 1;
    const unsigned blocks = (unsigned)((nwaves + 3) / 4);
    nwaves = (int64_t)blocks * 4;
    uint32_t *crow;
    float *cval;
    const size_t cap_slots = carry_slots(h);
    DCSR_TRY(ws_get(h, WS_SPMM_CROW, cap_slots * 4, (void **)&crow));
    DCSR_TRY(ws_get(h, WS_SPMM_CVAL, cap_slots * 4 * KB, (void **)&cval));
    const uint32_t *foff = (const uint32_t *)h->ws.ptr[WS_FOFF];
    const uint32_t *rowmap = (const uint32_t *)h->ws.ptr[WS_ROWMAP];
    const bool vec4 = ((uintptr_t)X % 16 == 0) && ((uintptr_t)Y % 16 == 0) && ldx % 4 == 0 && ldy % 4 == 0;
    const bool nt = stream_entries(h, KB);
    if (vec4 && nt)
        k_spmm_wide4<KB, true><<<blocks, TB, 0, s>>>(R, foff, rowmap, h->row_start, h->pool_col, h->pool_val, X, ldx,
                                                     Y, ldy, alpha, beta, crow, cval);
    else if (vec4)
        k_spmm_wide4<KB><<<blocks, TB, 0, s>>>(R, foff, rowmap, h->row_start, h->pool_col, h->pool_val, X, ldx, Y,
                                               ldy, alpha, beta, crow, cval);
    else if (nt)
        k_spmm_wide<KB, true><<<blocks, TB, 0, s>>>(R, foff, rowmap, h->row_start, h->pool_col, h->pool_val, X, ldx,
                                                    Y, ldy, alpha, beta, crow, cval);
    else
        k_spmm_wide<KB><<<blocks, TB, 0, s>>>(R, foff, rowmap, h->row_start, h->pool_col, h->pool_val, X, ldx, Y,
                                              ldy, alpha, beta, crow, cval);
    k_carry<KB><<<grid_for(nwaves, TB), TB, 0, s>>>((uint32_t)nwaves, crow, cval, Y, ldy, alpha);
    DCSR_HIP(hipGetLastError());
    return DCSR_OK;
}

static dcsr_status prepare_compute(dcsr_handle h, hipStream_t s)
{
    h->last_stream = s;
    return ensure_flat(h, s);
}

// Paneled copy worth building: column-sorted rows, x several panels wide, a
// large matrix, and the virtual row count within 32 bits.
static bool pnl_eligible(dcsr_handle h)
{
    if (!h->state_known || !h->canonical || h->cols < 2 * h->pnl_cols || h->nnz_h < h->pnl_min_nnz) return false;
    const int64_t P = (h->cols + h->pnl_cols - 1) / h->pnl_cols;
    return P * (h->rows + 1) < (1ll << 31);
}

static dcsr_status build_panels(dcsr_handle h, hipStream_t s)
{
    const int64_t R = h->rows;
    const int P = (int)((h->cols + h->pnl_cols - 1) / h->pnl_cols);
    const int64_t Rv = (int64_t)P * (R + 1);
    const int64_t nnz = h->nnz_h;
    uint32_t *poff, *bnd, *pmap, *part;
    int32_t *pcol;
    float *pval, *ppart;
    DCSR_TRY(ws_get(h, WS_PNL_OFF, (size_t)(Rv + 1) * 4, (void **)&poff));
    DCSR_TRY(ws_get(h, WS_PNL_BND, (size_t)P * R * 4, (void **)&bnd));
    DCSR_TRY(ws_get(h, WS_PNL_COL, (size_t)(nnz + 1) * 4, (void **)&pcol));
    DCSR_TRY(ws_get(h, WS_PNL_VAL, (size_t)(nnz + 1) * 4, (void **)&pval));
    DCSR_TRY(ws_get(h, WS_PNL_MAP, (size_t)(nnz / GRAN + 1) * 4, (void **)&pmap));
    DCSR_TRY(ws_get(h, WS_PNL_PART, (size_t)Rv * 4, (void **)&ppart));
    DCSR_TRY(ws_get_zeroed(h, WS_PART, scan_part_words(Rv + 1, 1) * 4, (void **)&part, s));
    const uint32_t *foff = (const uint32_t *)h->ws.ptr[WS_FOFF];
    const uint32_t *rowmap = (const uint32_t *)h->ws.ptr[WS_ROWMAP];
    // per-(panel, row) counts go to the partial buffer (free until the first
    // paneled SpMV), then are scanned into poff
    uint32_t *pcnt = (uint32_t *)ppart;
    k_pnl_bounds<<<grid_for(R, TB), TB, 0, s>>>((uint32_t)R, P, h->pnl_cols, h->row_start, h->row_len, h->pool_col,
                                                pcnt, bnd);
    DCSR_HIP(scan_exclusive_u32(pcnt, poff, Rv, 1, part, poff + Rv, s, fault_word(h)));
    if (nnz > 0)
        k_pnl_fill<<<grid_for(nnz, TB), TB, 0, s>>>((uint32_t)nnz, (uint32_t)R, h->pnl_cols, foff, rowmap,
                                                    h->row_start, h->pool_col, h->pool_val, poff, bnd, pcol, pval);
    const int64_t ng = nnz / GRAN + 1;
    k_rowmap<<<grid_for(ng, TB), TB, 0, s>>>((uint32_t)Rv, poff, pmap, (uint32_t)ng);
    DCSR_HIP(hipGetLastError());
    h->pnl_count = P;
    h->pnl_valid = true;
    return DCSR_OK;
}

static dcsr_status spmv_paneled(dcsr_handle h, const float *x, float *y, float alpha, float beta, hipStream_t s)
{
    const int64_t R = h->rows;
    const int P = h->pnl_count;
    const uint32_t Rv = (uint32_t)((int64_t)P * (R + 1));
    const uint32_t pw = DCSR_SPMV_PW;
    int64_t nwaves = (live_bound(h) + pw - 1) / pw;
    if (nwaves < 1) nwaves = 1;
    const unsigned blocks = (unsigned)((nwaves + 3) / 4);
    nwaves = (int64_t)blocks * 4;
    uint32_t *crow;
    float *cval;
    const size_t cap_slots = carry_slots(h);
    DCSR_TRY(ws_get(h, WS_CARRY_ROW, cap_slots * 4, (void **)&crow));
    DCSR_TRY(ws_get(h, WS_CARRY_VAL, cap_slots * 4, (void **)&cval));
    const uint32_t *poff = (const uint32_t *)h->ws.ptr[WS_PNL_OFF];
    const uint32_t *pmap = (const uint32_t *)h->ws.ptr[WS_PNL_MAP];
    const int32_t *pcol = (const int32_t *)h->ws.ptr[WS_PNL_COL];
    const float *pval = (const float *)h->ws.ptr[WS_PNL_VAL];
    float *ppart = (float *)h->ws.ptr[WS_PNL_PART];
    // the copy is contiguous in flat order: a virtual row starts at its offset
    k_flat<1, true><<<blocks, TB, 0, s>>>(Rv, poff, pmap, poff, pcol, pval, x, 1, ppart, 1, 1.0f, 0.0f, pw, crow, cval);
    k_carry<1><<<grid_for(nwaves, TB), TB, 0, s>>>((uint32_t)nwaves, crow, cval, ppart, 1, 1.0f);
    k_pnl_reduce<<<grid_for(R, TB), TB, 0, s>>>((uint32_t)R, P, ppart, y, alpha, beta);
    DCSR_HIP(hipGetLastError());
    return DCSR_OK;
}

// SpMV.  A matrix whose x spans several L2-sized column panels is multiplied
// through its paneled copy once its structure has seen enough SpMV calls to
// pay for the copy (about one extra pass over the matrix), so
// update-then-multiply patterns keep the direct flat kernel.
dcsr_status dcsr_spmv(dcsr_handle h, const float *x, float *y, float alpha, float beta, dcsr_stream_t stream)
{
    if (!h || (h->rows > 0 && (!x || !y))) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    h->last_stream = s;
    if (h->rows == 0) return DCSR_OK;
    DCSR_TRY(prepare_compute(h, s));
    if (pnl_eligible(h)) {
        // the copy costs about two thirds of a direct SpMV and saves about a
        // third of one per later call: build it on the second call of a
        // structure when the last structure that saw SpMVs saw
        // DCSR_PNL_REUSE or more (a stream that updates every few calls
        // never pays for it), else on the DCSR_PNL_REUSE-th call
        const int c = ++h->pnl_calls;
        if (!h->pnl_valid && c >= 2 && (h->pnl_prev_calls >= DCSR_PNL_REUSE || c >= DCSR_PNL_REUSE))
            DCSR_TRY(build_panels(h, s));
        if (h->pnl_valid) return spmv_paneled(h, x, y, alpha, beta, s);
    }
    return launch_flat<1>(h, x, 1, y, 1, alpha, beta, s);
}

dcsr_status dcsr_spmm_dense(dcsr_handle h, const float *X, int64_t k, int64_t ldx, float *Y, int64_t ldy, float alpha,
                            float beta, dcsr_stream_t stream)
{
    if (!h || k < 1 || ldx < k || ldy < k) return DCSR_ERR_INVALID;
    if (h->rows > 0 && (!X || !Y)) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    h->last_stream = s;
    if (h->rows == 0) return DCSR_OK;
    DCSR_TRY(prepare_compute(h, s));
    // column blocks of width 32, 16, 8, 4, 2, 1
    for (int64_t cb = 0; cb < k;) {
        const int64_t rem = k - cb;
        const float *Xb = X + cb;
        float *Yb = Y + cb;
        if (rem >= 32) { DCSR_TRY(launch_wide<32>(h, Xb, ldx, Yb, ldy, alpha, beta, s)); cb += 32; }
        else if (rem >= 16) { DCSR_TRY(launch_wide<16>(h, Xb, ldx, Yb, ldy, alpha, beta, s)); cb += 16; }
        else if (rem >= 8) { DCSR_TRY(launch_wide<8>(h, Xb, ldx, Yb, ldy, alpha, beta, s)); cb += 8; }
        else if (rem >= 4) { DCSR_TRY(launch_flat<4>(h, Xb, ldx, Yb, ldy, alpha, beta, s)); cb += 4; }
        else if (rem >= 2) { DCSR_TRY(launch_flat<2>(h, Xb, ldx, Yb, ldy, alpha, beta, s)); cb += 2; }
        else { DCSR_TRY(launch_flat<1>(h, Xb, ldx, Yb, ldy, alpha, beta, s)); cb += 1; }
    }
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

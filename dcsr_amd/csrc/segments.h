// segments.h -- device helpers for work that is flat over the entries of many
// variable-length segments (rows).  A wave owns a fixed range of entries, so
// a row of millions of entries is spread over many waves and short rows are
// packed 64-at-a-time: the load balance does not depend on the row-length
// distribution (power-law hubs included).
#pragma once
#include "wave.h"

// Largest j in [0, 64) with s[j] <= e (s non-decreasing, s[0] <= e).
__device__ __forceinline__ int find_seg64(const uint32_t *s, uint32_t e)
{
    int lo = 0, hi = 64;
#pragma unroll
    for (int it = 0; it < 6; ++it) {
        int mid = (lo + hi) >> 1;
        if (s[mid] <= e) lo = mid; else hi = mid;
    }
    return lo;
}

// Largest j in [0, n) with off[j*S] <= e (global array, non-decreasing,
// off[0] <= e).  64-ary search by the whole wave: every round loads 64
// probes at once and a ballot picks the sub-range, so the number of dependent
// memory round trips is ceil(log64(n)) instead of log2(n).  All lanes must
// call it; every lane gets the result.
template <int S = 1>
__device__ __forceinline__ uint32_t wave_find_last_le(const uint32_t *off, uint32_t n, uint32_t e)
{
    const int lane = lane_id();
    uint32_t lo = 0, hi = n;  // answer in [lo, hi)
    while (hi - lo > 64) {
        const uint32_t step = (hi - lo + 63) / 64;
        const uint32_t idx = lo + (uint32_t)lane * step;
        const bool ok = idx < hi && off[(uint64_t)idx * S] <= e;
        const uint64_t m = __ballot(ok);  // prefix of lanes (monotone), lane 0 always set
        const uint32_t k = 63u - (uint32_t)__clzll((long long)m);
        lo = lo + k * step;
        const uint32_t nh = lo + step;
        hi = nh < hi ? nh : hi;
    }
    const uint32_t idx = lo + (uint32_t)lane;
    const bool ok = idx < hi && off[(uint64_t)idx * S] <= e;
    const uint64_t m = __ballot(ok);
    return lo + (63u - (uint32_t)__clzll((long long)m));
}

// Visit entries [e0, e1) of the flat space given by the exclusive offsets
// off[0..nseg] (element j at off[j*S]; off[nseg*S] = total).  For every entry, calls
//   f(seg, k, e)   with k = e - off[seg]   (all lanes of the wave call
// into the loop together; lanes without an entry get valid == false).
// s_off: 65 words of LDS owned by this wave.
template <int S = 1, class F>
__device__ __forceinline__ void wave_visit(const uint32_t *__restrict__ off, uint32_t nseg, uint32_t e0, uint32_t e1,
                                           uint32_t *s_off, F &&f)
{
    const int lane = lane_id();
    if (e0 >= e1) return;
    uint32_t j = wave_find_last_le<S>(off, nseg, e0);
    uint32_t e = e0;
    while (e < e1) {
        // window: segments j .. j+63, boundaries off[j .. j+64]
        uint32_t idx = j + lane;
        s_off[lane] = idx <= nseg ? off[(uint64_t)idx * S] : 0xffffffffu;
        if (lane == 0) s_off[64] = (j + 64 <= nseg) ? off[(uint64_t)(j + 64) * S] : 0xffffffffu;
        __builtin_amdgcn_wave_barrier();
        uint32_t wend = s_off[64];
        if (wend > e1) wend = e1;
        for (uint32_t b = e; b < wend; b += 64) {
            uint32_t ee = b + lane;
            bool valid = ee < wend;
            int sj = valid ? find_seg64(s_off, ee) : 0;
            f(valid, j + sj, valid ? ee - s_off[sj] : 0u, ee);
        }
        e = wend;
        j += 64;
        __builtin_amdgcn_wave_barrier();
    }
}

// Start map of up to 64 * 64 items: bit i set when item i starts a segment
// (segments in order, each starting at a distinct item).  Loaded from LDS
// words (two per 64-bit word) into lane w = word w, with the count of set
// bits in the words before it; all lanes must call it.
struct StartMap {
    uint64_t word;
    uint32_t before;
};
__device__ __forceinline__ StartMap start_map_load(const uint32_t *m, int nwords64)
{
    const int lane = lane_id();
    uint64_t w = 0;
    if (lane < nwords64) w = (uint64_t)m[2 * lane] | ((uint64_t)m[2 * lane + 1] << 32);
    const uint32_t c = (uint32_t)__popcll(w);
    return {w, wave_incl_scan_u32(c) - c};
}

// Set bits of the map at items <= q * 64 + lane (the calling lane); q must be
// the same in every lane (a compile-time index in unrolled loops).
__device__ __forceinline__ uint32_t start_map_count(const StartMap &sm, int q)
{
    const uint64_t w = ((uint64_t)wave_read_lane_u32((uint32_t)(sm.word >> 32), q) << 32) |
                       wave_read_lane_u32((uint32_t)sm.word, q);
    const uint32_t lane = (uint32_t)lane_id();
    return wave_read_lane_u32(sm.before, q) + (uint32_t)__popcll(w & ((2ull << lane) - 1ull));
}

// Set bits of the map strictly before item i (any i per lane).
__device__ __forceinline__ uint32_t start_map_rank(const StartMap &sm, uint32_t i)
{
    const int wi = (int)(i >> 6);
    const uint64_t w = ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(sm.word >> 32), wi) << 32) |
                       (uint32_t)__shfl((int)(uint32_t)sm.word, wi);
    return (uint32_t)__shfl((int)sm.before, wi) + (uint32_t)__popcll(w & ((1ull << (i & 63)) - 1ull));
}

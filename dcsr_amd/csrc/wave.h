// wave.h -- wave64 cross-lane primitives for gfx950 (CDNA4).
//
// Scans use DPP row shifts inside each 16-lane row, then the gfx9 row
// broadcasts (row_bcast:15 / row_bcast:31) to carry across rows; ballots give
// 64-bit lane masks.  These replace the reference's warp32 shuffle scans
// (include/scan.h:18-96, inline PTX shfl.up) and the cub WarpReduce /
// BlockRadixSort building blocks it used (include/spmm.inl:197-202, 489-490).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define DCSR_WAVE 64

// DPP control words (GFX9 encoding).
#define DPP_ROW_SHR(n) (0x110 + (n))
#define DPP_ROW_BCAST15 0x142
#define DPP_ROW_BCAST31 0x143

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (DCSR_WAVE - 1)); }

__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t old, uint32_t src)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)src, CTRL, ROW_MASK, BANK_MASK, false);
}

template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf>
__device__ __forceinline__ float dpp_f32(float old, float src)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), CTRL,
                                                      ROW_MASK, BANK_MASK, false));
}

// Inclusive prefix sum across the 64 lanes (all lanes must be active).
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v)
{
    v += dpp_u32<DPP_ROW_SHR(1)>(0u, v);
    v += dpp_u32<DPP_ROW_SHR(2)>(0u, v);
    v += dpp_u32<DPP_ROW_SHR(4)>(0u, v);
    v += dpp_u32<DPP_ROW_SHR(8)>(0u, v);
    v += dpp_u32<DPP_ROW_BCAST15, 0xa>(0u, v);
    v += dpp_u32<DPP_ROW_BCAST31, 0xc>(0u, v);
    return v;
}

__device__ __forceinline__ uint32_t wave_read_lane_u32(uint32_t v, int lane)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}

__device__ __forceinline__ float wave_read_lane_f32(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) { return wave_read_lane_u32(wave_incl_scan_u32(v), 63); }

// Segmented inclusive float sum: lanes with equal `seg` (non-decreasing along
// the lanes) form one segment.  The association order is fixed (Hillis-Steele
// over the lane index), so results are deterministic.
__device__ __forceinline__ float wave_seg_incl_scan_f32(float v, uint32_t seg)
{
    float t;
    uint32_t s;
    t = dpp_f32<DPP_ROW_SHR(1)>(0.f, v); s = dpp_u32<DPP_ROW_SHR(1)>(0xffffffffu, seg); v += (s == seg) ? t : 0.f;
    t = dpp_f32<DPP_ROW_SHR(2)>(0.f, v); s = dpp_u32<DPP_ROW_SHR(2)>(0xffffffffu, seg); v += (s == seg) ? t : 0.f;
    t = dpp_f32<DPP_ROW_SHR(4)>(0.f, v); s = dpp_u32<DPP_ROW_SHR(4)>(0xffffffffu, seg); v += (s == seg) ? t : 0.f;
    t = dpp_f32<DPP_ROW_SHR(8)>(0.f, v); s = dpp_u32<DPP_ROW_SHR(8)>(0xffffffffu, seg); v += (s == seg) ? t : 0.f;
    t = dpp_f32<DPP_ROW_BCAST15, 0xa>(0.f, v); s = dpp_u32<DPP_ROW_BCAST15, 0xa>(0xffffffffu, seg); v += (s == seg) ? t : 0.f;
    t = dpp_f32<DPP_ROW_BCAST31, 0xc>(0.f, v); s = dpp_u32<DPP_ROW_BCAST31, 0xc>(0xffffffffu, seg); v += (s == seg) ? t : 0.f;
    return v;
}

// Segmented inclusive uint32 sum (segments as in wave_seg_incl_scan_f32).
__device__ __forceinline__ uint32_t wave_seg_incl_scan_u32(uint32_t v, uint32_t seg)
{
    uint32_t t, s;
    t = dpp_u32<DPP_ROW_SHR(1)>(0u, v); s = dpp_u32<DPP_ROW_SHR(1)>(0xffffffffu, seg); v += (s == seg) ? t : 0u;
    t = dpp_u32<DPP_ROW_SHR(2)>(0u, v); s = dpp_u32<DPP_ROW_SHR(2)>(0xffffffffu, seg); v += (s == seg) ? t : 0u;
    t = dpp_u32<DPP_ROW_SHR(4)>(0u, v); s = dpp_u32<DPP_ROW_SHR(4)>(0xffffffffu, seg); v += (s == seg) ? t : 0u;
    t = dpp_u32<DPP_ROW_SHR(8)>(0u, v); s = dpp_u32<DPP_ROW_SHR(8)>(0xffffffffu, seg); v += (s == seg) ? t : 0u;
    t = dpp_u32<DPP_ROW_BCAST15, 0xa>(0u, v); s = dpp_u32<DPP_ROW_BCAST15, 0xa>(0xffffffffu, seg); v += (s == seg) ? t : 0u;
    t = dpp_u32<DPP_ROW_BCAST31, 0xc>(0u, v); s = dpp_u32<DPP_ROW_BCAST31, 0xc>(0xffffffffu, seg); v += (s == seg) ? t : 0u;
    return v;
}

// Lanes holding the same 8-bit digit (the active-lane mask is `active`).
__device__ __forceinline__ uint64_t wave_match_digit8(uint32_t d, uint64_t active)
{
    uint64_t peers = active;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        uint64_t m = __ballot((d >> b) & 1u);
        peers &= ((d >> b) & 1u) ? m : ~m;
    }
    return peers;
}

// Block-wide exclusive scan of one uint32 per thread (256 threads, four
// waves); *block_total receives the block sum.  lds4: four words of LDS.
__device__ __forceinline__ uint32_t block_excl_scan_256(uint32_t v, uint32_t *lds4, uint32_t *block_total)
{
    const int w = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan_u32(v);
    if (lane_id() == 63) lds4[w] = inc;
    __syncthreads();
    uint32_t carry = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t t = lds4[i];
        if (i < w) carry += t;
        tot += t;
    }
    __syncthreads();
    *block_total = tot;
    return carry + inc - v;
}

// block_excl_scan_256 for N columns at once (one LDS exchange and one pair
// of barriers for all of them): v[i] becomes the exclusive prefix of column
// i over the workgroup, tot[i] its total.  lds: [4][N] words.
template <int N>
__device__ __forceinline__ void block_excl_scan_n(uint32_t (&v)[N], uint32_t (*lds)[N], uint32_t (&tot)[N])
{
    const int w = threadIdx.x >> 6;
    uint32_t inc[N];
#pragma unroll
    for (int i = 0; i < N; ++i) inc[i] = wave_incl_scan_u32(v[i]);
    if (lane_id() == 63) {
#pragma unroll
        for (int i = 0; i < N; ++i) lds[w][i] = inc[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < N; ++i) {
        uint32_t carry = 0, t = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t x = lds[q][i];
            if (q < w) carry += x;
            t += x;
        }
        tot[i] = t;
        v[i] = carry + inc[i] - v[i];
    }
    __syncthreads();
}

// Workgroup totals of N columns at once: v[i] becomes column i's total.
template <int N>
__device__ __forceinline__ void block_sum_n(uint32_t (&v)[N], uint32_t (*lds)[N])
{
    uint32_t t[N];
#pragma unroll
    for (int i = 0; i < N; ++i) t[i] = wave_read_lane_u32(wave_incl_scan_u32(v[i]), 63);
    if (lane_id() == 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) lds[threadIdx.x >> 6][i] = t[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = (lds[0][i] + lds[1][i]) + (lds[2][i] + lds[3][i]);
    __syncthreads();
}

// Fixed-order float sum over the wave (tree over lane index).
__device__ __forceinline__ float wave_sum_f32(float v)
{
    v += __shfl_xor(v, 32);
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 8);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 1);
    return v;
}

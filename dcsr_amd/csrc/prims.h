// prims.h -- device-wide primitives (host launchers).  Implemented in
// prims.hip with wave64 DPP scans and ballot-based digit ranking.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Exclusive prefix sum of NC interleaved uint32 components (in[i*NC + c]),
// n elements, out of place (in != out: a tile whose bounded look-back wait
// gives up re-reads the inputs of the tiles before it).  `total` (device, NC
// words) receives the sums.
// `part` is scratch of scan_part_words(n, NC) words, zero-filled when first
// allocated; calls that share a `part` buffer must be stream-ordered.
// `fault` (device, may be null) is the owner's sticky status word: a tile
// that took that exact fallback sets LB_FALLBACK_BIT (0x8) in it.
size_t scan_part_words(int64_t n, int nc);
hipError_t scan_exclusive_u32(const uint32_t *in, uint32_t *out, int64_t n, int nc, uint32_t *part,
                              uint32_t *total, hipStream_t s, unsigned int *fault);

// Entries per tile of the radix sort's count table (digit-major: the count
// of digit d in tile t sits at hist[d * tiles + t]).
constexpr int RADIX_TILE = 2048;

// Stable LSD radix sort of (key, value) pairs on key bits [0, nbits).
// Ping-pongs between (k0, v0) and (k1, v1); returns in *which the buffer
// index (0 or 1) holding the result.  `hist` is scratch of
// radix_hist_words(n) words; `part` is scan scratch of radix_part_words(n)
// words (zero-filled when first allocated, as for scan_exclusive_u32).
size_t radix_hist_words(int64_t n);
size_t radix_part_words(int64_t n);
// counted: the caller already wrote the first pass's (bits 0..7) counts
// into hist, per RADIX_TILE tile of (k0, v0).
hipError_t radix_sort_pairs(uint64_t *k0, uint32_t *v0, uint64_t *k1, uint32_t *v1, int64_t n, int nbits,
                            uint32_t *hist, uint32_t *part, int *which, hipStream_t s, bool counted = false);

// Row-group sort (gsort.hip) of pairs keyed (row << cbits) | col, ~0 for an
// invalid key: digit d = key >> lb (lb = gshift + cbits: the row's group of
// 2^gshift rows), ~0 -> digit ndig - 1.  The caller counts the digits of each
// RADIX_TILE tile of (k0, v0) into hist (blocked: see gs_hist_index); the
// stable result ends in (k0, v0), (k1, v1) is scratch.
constexpr int GS_MAX_DIGITS = 2048;  // groups + 1
#ifndef DCSR_GS_CAP
#define DCSR_GS_CAP 3072
#endif
constexpr int GS_CAP = DCSR_GS_CAP;  // pairs of a group sorted in LDS
// Count of digit d in tile t: 64-digit blocks, each tile-major.
__host__ __device__ inline int64_t gs_hist_index(int d, int64_t t, int64_t ntiles)
{
    return (int64_t)(d >> 6) * ntiles * 64 + t * 64 + (d & 63);
}
size_t gs_hist_words(int64_t n, int ndig);
hipError_t group_sort_pairs(uint64_t *k0, uint32_t *v0, uint64_t *k1, uint32_t *v1, int64_t n, int lb, int ndig,
                            uint32_t *hist, hipStream_t s);
// Exclusive group offsets (ndig + 1 entries) that group_sort_pairs leaves in
// its hist buffer: pairs of group g sit at [gbase[g], gbase[g + 1]).
inline uint32_t *gs_group_base(uint32_t *hist, int64_t n, int ndig)
{
    const int64_t nb = (n + RADIX_TILE - 1) / RADIX_TILE;
    const int64_t blocks = (ndig + 63) / 64;
    return hist + nb * 64 * blocks + GS_MAX_DIGITS + 64;
}

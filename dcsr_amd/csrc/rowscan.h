// rowscan.h -- one tile of the SpMV's flat row offsets and row map.
//
// foff[r] = len[0] + ... + len[r-1] (foff[R] = total) and, for g <= total /
// RS_GRAN, rowmap[g] = the row r with foff[r] < g*RS_GRAN <= foff[r+1]
// (granule 0 belongs to row 0): any RS_GRAN-aligned range of the flat order
// finds its first row with one load.  A tile is 4096 rows (16 contiguous
// rows per thread, vector loads and stores); tiles take tickets in launch
// order and chain their prefixes by decoupled look-back (lookback.h), so a
// matrix of a few hundred thousand rows resolves in one or two rounds.  The
// tile's granules are filled by the whole workgroup, one granule per thread
// and a search of the tile's offsets in LDS, so that a long row does not
// leave one thread storing its granules one by one.
//
// Used by the SpMV's own scan kernel (spmv.hip) over the row lengths, and by
// the update's move kernel (update.hip), whose extra workgroups scan the
// lengths the update is committing while the entries move.
#pragma once
#include "lookback.h"
#include "wave.h"

constexpr int RS_TB = 256;
constexpr int RS_IPT = 16;
constexpr int RS_TILE = RS_TB * RS_IPT;
constexpr uint32_t RS_GRAN = 64;  // row-map granule (nonzeros)

// Look-back block of a row scan: 4 control words, then one 64-bit status
// word per tile.
static inline int64_t rowscan_tiles(int64_t R) { return (R + RS_TILE - 1) / RS_TILE; }
static inline size_t rowscan_part_bytes(int64_t R) { return (size_t)(4 + 2 * rowscan_tiles(R)) * 4; }

struct RowScanLds {
    uint32_t lds4[4], tk[2], pre;
    uint32_t off[RS_TILE + 1];
};

// One tile (its index is the ticket it draws), by a workgroup of RS_TB
// threads.  lens(r0, full, v) loads the lengths of rows r0 .. r0+15 (full:
// all of them exist; else rows past R read 0); prefix(n) is the exact sum of
// the lengths of rows [0, n) by one whole wave (the look-back fallback).
// ntiles: the workgroups that run this function in the launch (the last one
// to finish resets the control words).
template <class Lens, class Prefix>
__device__ void rowscan_tile(uint32_t R, Lens &&lens, Prefix &&prefix, uint32_t *__restrict__ foff,
                             uint32_t *__restrict__ rowmap, uint32_t *__restrict__ part, unsigned int *fault,
                             int limit, uint32_t ntiles, RowScanLds &L)
{
    unsigned int *ctl = part;
    uint64_t *st = (uint64_t *)(part + 4);
    uint32_t tag;
    const int64_t tile = lb_ticket(ctl, L.tk, &tag);
    const int64_t r0 = tile * RS_TILE + (int64_t)threadIdx.x * RS_IPT;
    const bool full = r0 + RS_IPT <= (int64_t)R;
    uint32_t v[RS_IPT];
    lens(r0, full, v);
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < RS_IPT; ++j) sum += v[j];
    uint32_t tot;
    const uint32_t pre = block_excl_scan_256(sum, L.lds4, &tot);
    if (threadIdx.x == 0) lb_publish(&st[tile], tag, tile == 0 ? LB_INC : LB_AGG, tot);
    if (threadIdx.x < 64) {
        uint32_t p[1] = {0u};
        if (tile > 0) lb_wave_lookback<1>(st, tile, tag, ctl, fault, limit, [&](int) { return prefix(tile * RS_TILE); }, p);
        if (threadIdx.x == 0) {
            L.pre = p[0];
            if (tile > 0) lb_publish(&st[tile], tag, LB_INC, p[0] + tot);
        }
    }
    __syncthreads();
    const int64_t tr0 = tile * RS_TILE;
    const uint32_t n = (uint32_t)((int64_t)R - tr0 < RS_TILE ? (int64_t)R - tr0 : RS_TILE);  // rows in the tile
    uint32_t a = L.pre + pre;
    uint32_t o[RS_IPT];
#pragma unroll
    for (int j = 0; j < RS_IPT; ++j) {
        o[j] = a;
        const uint32_t lj = threadIdx.x * RS_IPT + j;
        if (lj < n) L.off[lj] = a;
        a += v[j];
    }
    if (threadIdx.x == 0) L.off[n] = L.pre + tot;
    __syncthreads();
    {
        // granules g with off[0] < g*RS_GRAN <= off[n]: the tile's row i
        // with off[i] < g*RS_GRAN <= off[i+1] (the last i with off[i] <
        // g*RS_GRAN, found by bisection)
        const uint32_t g0 = L.off[0] / RS_GRAN + 1, g1 = L.off[n] / RS_GRAN;
        for (uint32_t g = g0 + threadIdx.x; g <= g1; g += RS_TB) {
            const uint32_t x = g * RS_GRAN;
            uint32_t l = 0, hgh = n - 1;
            while (l < hgh) {
                const uint32_t mid = (l + hgh + 1) >> 1;
                if (L.off[mid] < x) l = mid;
                else hgh = mid - 1;
            }
            rowmap[g] = (uint32_t)(tr0 + l);
        }
    }
    if (full) {
        uint4 *q = (uint4 *)(foff + r0);
#pragma unroll
        for (int k = 0; k < RS_IPT / 4; ++k) q[k] = make_uint4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
    } else {
#pragma unroll
        for (int j = 0; j < RS_IPT; ++j)
            if (r0 + j < (int64_t)R) foff[r0 + j] = o[j];
    }
    // the thread holding row R-1 stores the total
    if (r0 < (int64_t)R && r0 + RS_IPT >= (int64_t)R) foff[R] = a;
    if (tile == 0 && threadIdx.x == 0) rowmap[0] = 0u;
    __syncthreads();
    if (threadIdx.x == 0) lb_finish_n(ctl, ntiles);
}

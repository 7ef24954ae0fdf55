// update.hip -- in-place insert-or-add of a batch of triples.
//
// Replaces device::UpdateMatrix -> UpdateMatrix_dcsr (include/load.h:258-325,
// include/sparse_update.inl:16-130).  The reference appends a batch row into
// a new segment claimed by an atomicAdd on the pool bump pointer, without
// merging duplicates, adding to existing coordinates or dropping zeros.  This
// engine produces exactly the host path's result (sort batch by row,
// COO->CSR, cusp::add; cusp/system/detail/sequential/elementwise.h:44-131):
//
//   1. pack (row, col) keys (counting the sort's first digit per tile on
//      the way), stable LSD radix sort of (key, value) pairs;
//   2. heads (two passes over 2048-entry tiles): one record per distinct
//      coordinate, summed left-to-right in batch order starting from 0 (the
//      host path's B_row[j] accumulation), one per touched row;
//   3. probe every distinct coordinate in its (column-sorted) row: existing
//      -> new = old + sum, removed when == 0; missing -> inserted when != 0;
//   4. plan per touched row: rows whose structure is unchanged get their
//      values written in place; the others claim a new chunk of the pool's
//      bump region.  Claims are slot offsets from a prefix sum, so no two
//      rows contend on an atomic.  The last workgroup of the plan launch
//      turns every per-tile count into offsets and decides the batch: a bad
//      index or pool exhaustion rejects it whole (no-op), otherwise the bump
//      pointer advances;
//   5. offsets: per-entry exclusive counts (inserts / deletes before each
//      coordinate, claims / moved entries before each row);
//   6. move surviving old entries and write inserted ones at their merged
//      positions, then commit the row table.
//
// Cross-tile prefixes come from per-tile counts scanned by the last
// workgroup to finish (an atomic finish counter); nothing polls.
//
// dcsr_update is synchronous by default: the batch's status comes back with
// the call, from one read of the device state (at the end when the host can
// bound the claims below the free pool, else between the plan and the
// writes).  With DCSR_UPDATE_NOSYNC nothing is read back: the host keeps
// upper bounds of the live entries and of the bump pointer, tightened from
// asynchronous state snapshots, and compacts in stream order when they
// cannot rule out exhaustion.
#include "dcsr_internal.h"
#include "prims.h"
#include "rowscan.h"
#include "segments.h"

namespace {

constexpr int TB = 256;
enum { C_NOP = 0, C_ADD = 1, C_DEL = 2, C_INS = 3 };
enum { M_VALUES = 0, M_REWRITE = 1 };
constexpr int NPLAN = 5;  // per-block plan counters

// Pack the caller's triples into sort keys (row << cbits | col) and value
// bits, one RADIX_TILE tile per workgroup.  Invalid indices become ~0 keys:
// with one key bit above row|col they sort strictly after every valid key,
// and they raise the batch's flag.  With `hist`, the tile's digit counts go
// to the sort's count table on the way: for the radix sort the first digit
// (key bits 0..7, digit-major table), for the row-group sort (lb > 0) the
// row group (key >> lb, ~0 -> ndig - 1; blocked table of gs_hist_index).
__global__ __launch_bounds__(TB) void k_upd_pack(const int32_t *__restrict__ rows, const int32_t *__restrict__ cols,
                                                 const float *__restrict__ vals, int64_t n, int64_t R, int64_t C,
                                                 int cbits, uint64_t *__restrict__ key, uint32_t *__restrict__ val,
                                                 uint32_t *__restrict__ hist, int64_t ntiles, DevState *ds, int lb,
                                                 int ndig)
{
    __shared__ uint32_t cnt[GS_MAX_DIGITS];
    static_assert(TB == 256, "one digit counter per thread");
    const int ncnt = lb > 0 ? ((ndig + 63) & ~63) : 256;
    if (hist) {
        for (int d = threadIdx.x; d < ncnt; d += TB) cnt[d] = 0;
        __syncthreads();
    }
    const int64_t base = (int64_t)blockIdx.x * RADIX_TILE;
    bool bad = false;
    // all loads in flight first (a tail item reads the tile's first triple
    // and is dropped)
    int32_t rr[RADIX_TILE / TB], cc[RADIX_TILE / TB];
    uint32_t vv[RADIX_TILE / TB];
#pragma unroll
    for (int j = 0; j < RADIX_TILE / TB; ++j) {
        const int64_t i = base + j * TB + threadIdx.x;
        const int64_t ci = i < n ? i : base;
        rr[j] = rows[ci];
        cc[j] = cols[ci];
        vv[j] = __float_as_uint(vals[ci]);
    }
#pragma unroll
    for (int j = 0; j < RADIX_TILE / TB; ++j) {
        const int64_t i = base + j * TB + threadIdx.x;
        if (i < n) {
            const int32_t r = rr[j], c = cc[j];
            const bool b = (r < 0) | (r >= R) | (c < 0) | (c >= C);
            const uint64_t k = b ? ~0ull : (((uint64_t)(uint32_t)r << cbits) | (uint32_t)c);
            key[i] = k;
            val[i] = vv[j];
            if (hist) {
                const uint32_t d = lb > 0 ? (b ? (uint32_t)(ndig - 1) : (uint32_t)(k >> lb)) : ((uint32_t)k & 255u);
                atomicAdd(&cnt[d], 1u);
            }
            bad |= b;
        }
    }
    if (__ballot(bad) && lane_id() == 0) atomicOr(&ds->ubad, 1u);
    if (hist) {
        __syncthreads();
        if (lb > 0) {
            for (int d = threadIdx.x; d < ncnt; d += TB) hist[gs_hist_index(d, blockIdx.x, ntiles)] = cnt[d];
        } else {
            hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
        }
    }
}

static inline unsigned pack_tiles(int64_t n) { return (unsigned)((n + RADIX_TILE - 1) / RADIX_TILE); }

// Sum of one 32-bit counter over the 256 threads of a block (fixed order).
__device__ __forceinline__ uint32_t block_sum_u32(uint32_t v, uint32_t *lds4)
{
    const uint32_t w = wave_sum_u32(v);
    if (lane_id() == 0) lds4[threadIdx.x >> 6] = w;
    __syncthreads();
    const uint32_t t = lds4[0] + lds4[1] + lds4[2] + lds4[3];
    __syncthreads();
    return t;
}

__device__ __forceinline__ long long wave_sum_i64(long long v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// ---------------------------------------------------------------------------
// Tile tables.  The per-element passes work on tiles of 2048 entries (8 per
// thread, blocked).  Each tile writes its counts to a table; the last
// workgroup of the launch to finish (an atomic finish counter: nothing ever
// waits) turns the table's columns into exclusive offsets, and a later
// kernel adds them to the in-tile prefixes.
// ---------------------------------------------------------------------------
constexpr int IPT = 8;
constexpr int TILE = TB * IPT;
// Positions of the probe space (touched rows end to end) and of the move
// space (old entries of rewritten rows end to end) owned by one wave.
constexpr uint32_t PR_SPAN = 1024;
constexpr uint32_t MV_SPAN = 512;
// batch tile table: heads / row heads (count, offset), claims / moved (sum, offset)
enum { TT_H = 0, TT_R, TT_HX, TT_RX, TT_C, TT_M, TT_CX, TT_MX, TT_WORDS };
// probe tile table: inserts, adds, deletes; insert / delete offsets; probe
// lengths of the tile's touched rows (sum, offset)
enum { TP_I = 0, TP_A, TP_D, TP_IX, TP_DX, TP_L, TP_LX, TP_WORDS = 8 };
// finish counters (self-resetting)
enum { FC_HEADS = 0, FC_PLAN = 1 };

// True in every thread of the last workgroup of the launch to finish; it then
// sees what every other workgroup wrote before finishing.  Each workgroup
// releases (L2 write-back) before it is counted and only the last one
// acquires (L2 invalidate): a full fence in every workgroup would also
// invalidate the L2 under the workgroups still running.
__device__ __forceinline__ bool last_block(unsigned int *ctr, uint32_t *s_flag)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const bool last = atomicAdd(ctr, 1u) == gridDim.x - 1;
        if (last) atomicExch(ctr, 0u);
        *s_flag = last ? 1u : 0u;
    }
    __syncthreads();
    const bool last = *s_flag != 0u;
    if (last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return last;
}

// One workgroup: exclusive scans of NC columns of tile tables with 8 words
// per row, tab[c][i * 8 + src[c]] -> tab[c][i * 8 + dst[c]] for i < cnt;
// tot[c] receives column c's total.  Every load of a 2048-row chunk (all
// columns) is issued before its first scan step, so a chunk costs one
// memory round trip.
static_assert(TT_WORDS == 8 && TP_WORDS == 8, "tile tables have 8 words per tile");
template <int NC>
__device__ void block_scan_columns(uint32_t *const (&tab)[NC], const int (&src)[NC], const int (&dst)[NC],
                                   int64_t cnt, uint32_t *lds4, uint32_t (&tot)[NC])
{
#pragma unroll
    for (int c = 0; c < NC; ++c) tot[c] = 0;
    for (int64_t base = 0; base < cnt; base += TILE) {
        const int64_t i0 = base + (int64_t)threadIdx.x * IPT;
        uint32_t v[NC][IPT];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
#pragma unroll
            for (int j = 0; j < IPT; ++j) v[c][j] = i0 + j < cnt ? tab[c][(i0 + j) * 8 + src[c]] : 0u;
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            uint32_t sum = 0;
#pragma unroll
            for (int j = 0; j < IPT; ++j) sum += v[c][j];
            uint32_t t;
            uint32_t p = tot[c] + block_excl_scan_256(sum, lds4, &t);
#pragma unroll
            for (int j = 0; j < IPT; ++j) {
                if (i0 + j < cnt) tab[c][(i0 + j) * 8 + dst[c]] = p;
                p += v[c][j];
            }
            tot[c] += t;
        }
    }
}

// The thread's IPT blocked keys (~0 past n) and the key before them (~0
// before item 0: it differs from every valid key and, shifted, from every
// valid row).
__device__ __forceinline__ void load_keys(const uint64_t *__restrict__ key, int64_t n, int64_t i0,
                                          uint64_t (&k)[IPT], uint64_t &prev)
{
    // unconditional loads (a tail item reads the last key and becomes ~0),
    // all in flight together
    prev = key[i0 > 0 ? (i0 - 1 < n ? i0 - 1 : n - 1) : 0];
#pragma unroll
    for (int j = 0; j < IPT; ++j) k[j] = key[i0 + j < n ? i0 + j : n - 1];
    if (!(i0 > 0 && i0 - 1 < n)) prev = ~0ull;
#pragma unroll
    for (int j = 0; j < IPT; ++j)
        if (i0 + j >= n) k[j] = ~0ull;
}

// Heads of the sorted batch: an item is a coordinate head if it is the
// first of its (row, col) run and a row head if the first of its row; items
// past n and invalid keys (~0) are neither.
struct HeadFlags {
    uint32_t h, r;  // bit j: item j of the thread
};
__device__ __forceinline__ HeadFlags head_flags(const uint64_t (&k)[IPT], uint64_t prev, int cbits)
{
    HeadFlags f{0u, 0u};
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const bool valid = k[j] != ~0ull;
        if (valid && prev != k[j]) f.h |= 1u << j;
        if (valid && (prev >> cbits) != (k[j] >> cbits)) f.r |= 1u << j;
        prev = k[j];
    }
    return f;
}

// Probe length of a touched row: its entries, at least 1 (a row without
// entries still owns one position, where its records are probed).
__device__ __forceinline__ uint32_t probe_len(int32_t len) { return len > 1 ? (uint32_t)len : 1u; }

// Sum of the probe lengths of the row heads among the thread's items, and
// their row lengths (one round trip: every row length is loaded at once).
__device__ __forceinline__ uint32_t head_rows_len(const uint64_t (&k)[IPT], uint32_t rflags, int cbits,
                                                  const int32_t *__restrict__ row_len, int32_t (&lj)[IPT])
{
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        lj[j] = 0;
        if ((rflags >> j) & 1u) lj[j] = row_len[(uint32_t)(k[j] >> cbits)];
    }
#pragma unroll
    for (int j = 0; j < IPT; ++j) sum += (rflags >> j) & 1u ? probe_len(lj[j]) : 0u;
    return sum;
}

// Pass 1 of the heads: per-tile counts (coordinates, touched rows, and the
// rows' probe lengths).  With `table` the last workgroup makes them offsets
// and stores nuniq / ntouched / ptotal; without it (batches of at most
// HEADS_DIRECT_TILES tiles) pass 2 sums the counts before each tile itself,
// and this kernel ends at the counts.
constexpr int64_t HEADS_DIRECT_TILES = 4 * TB;
__global__ __launch_bounds__(TB) void k_upd_heads_count(const uint64_t *__restrict__ key, int64_t n, int cbits,
                                                        const int32_t *__restrict__ row_len,
                                                        uint32_t *__restrict__ tt, uint32_t *__restrict__ tp,
                                                        unsigned int *fin, DevState *ds, bool table)
{
    __shared__ uint32_t lds4[4], s_flag;
    const int64_t tile = blockIdx.x;
    uint64_t k[IPT], prev;
    load_keys(key, n, tile * TILE + (int64_t)threadIdx.x * IPT, k, prev);
    const HeadFlags f = head_flags(k, prev, cbits);
    int32_t lj[IPT];
    const uint32_t ls = head_rows_len(k, f.r, cbits, row_len, lj);
    const uint32_t hc = block_sum_u32(__popc(f.h), lds4);
    const uint32_t rc = block_sum_u32(__popc(f.r), lds4);
    const uint32_t lc = block_sum_u32(ls, lds4);
    if (threadIdx.x == 0) {
        tt[tile * TT_WORDS + TT_H] = hc;
        tt[tile * TT_WORDS + TT_R] = rc;
        tp[tile * TP_WORDS + TP_L] = lc;
    }
    if (!table) return;
    if (last_block(&fin[FC_HEADS], &s_flag)) {
        uint32_t *const tabs[3] = {tt, tt, tp};
        uint32_t tot[3];
        block_scan_columns<3>(tabs, {TT_H, TT_R, TP_L}, {TT_HX, TT_RX, TP_LX}, gridDim.x, lds4, tot);
        if (threadIdx.x == 0) {
            ds->nuniq = tot[0];
            ds->ntouched = tot[1];
            ds->ptotal = tot[2];
        }
    }
}

// Pass 2 of the heads: every head writes its compacted record: ucol/usum/
// ut for coordinate u, trow/tubeg/tlpre (probe-length offset), the row's
// pool start, length and capacity (trs, tlen, tcap) and zeroed tins/tdel for touched row t,
// and (pspan) touched row t for every probe span that starts in it; (rowt)
// rowt[row] = rtag << 32 | t, the update's tag beside the row's touched
// index, for the move kernel's scan of the new row lengths (entries of
// earlier updates carry older tags).  Sums run left to right from 0 in batch order, as the host path
// accumulates B_row[j]; a run is summed from the thread's registers and
// only a run that continues past the thread's items reads memory.  A tile's
// records are contiguous, so they are staged in LDS and written out
// coalesced.  The tubeg[ntouched] = nuniq sentinel is stored by workgroup 0
// (table) or by the last tile, which then also stores nuniq / ntouched.
__global__ __launch_bounds__(TB) void k_upd_heads_emit(const uint64_t *__restrict__ key,
                                                       const uint32_t *__restrict__ val, int64_t n, int cbits,
                                                       const uint32_t *__restrict__ tt, DevState *ds, bool table,
                                                       int32_t *__restrict__ ucol, float *__restrict__ usum,
                                                       uint32_t *__restrict__ ut, uint32_t *__restrict__ trow,
                                                       uint32_t *__restrict__ tubeg, const int32_t *__restrict__ row_len,
                                                       const uint32_t *__restrict__ row_start,
                                                       const int32_t *__restrict__ row_cap,
                                                       const uint32_t *__restrict__ tp, uint32_t *__restrict__ tlpre,
                                                       uint32_t *__restrict__ trs, uint32_t *__restrict__ tlen,
                                                       uint32_t *__restrict__ tcap,
                                                       uint32_t *__restrict__ tins, uint32_t *__restrict__ tdel,
                                                       uint32_t *__restrict__ pspan, uint64_t *__restrict__ rowt,
                                                       uint32_t rtag)
{
    __shared__ uint32_t lds3[4][3];
    __shared__ int32_t s_ucol[TILE];
    __shared__ float s_usum[TILE];
    __shared__ uint32_t s_ut[TILE], s_trow[TILE], s_tub[TILE], s_tl[TILE];
    const int64_t tile = blockIdx.x;
    const int64_t i0 = tile * TILE + (int64_t)threadIdx.x * IPT;
    // counts of the tiles before this one (loaded with the keys): each
    // thread sums up to four of them (index clamped to tile 0, masked)
    uint32_t ch[HEADS_DIRECT_TILES / TB], cr[HEADS_DIRECT_TILES / TB], cl[HEADS_DIRECT_TILES / TB];
    uint32_t ubase = 0, tbase = 0, lbase = 0;
    if (table) {
        ubase = tt[tile * TT_WORDS + TT_HX];
        tbase = tt[tile * TT_WORDS + TT_RX];
        lbase = tp[tile * TP_WORDS + TP_LX];
    } else {
#pragma unroll
        for (int q = 0; q < HEADS_DIRECT_TILES / TB; ++q) {
            const int64_t pt = (int64_t)threadIdx.x + q * TB;
            ch[q] = tt[(pt < tile ? pt : 0) * TT_WORDS + TT_H];
            cr[q] = tt[(pt < tile ? pt : 0) * TT_WORDS + TT_R];
            cl[q] = tp[(pt < tile ? pt : 0) * TP_WORDS + TP_L];
        }
    }
    uint64_t k[IPT], prev;
    load_keys(key, n, i0, k, prev);
    uint32_t vb[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) vb[j] = val[i0 + j < n ? i0 + j : n - 1];
    const HeadFlags f = head_flags(k, prev, cbits);
    int32_t lj[IPT];
    uint32_t rsj[IPT], rcj[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const bool rh = (f.r >> j) & 1u;
        rsj[j] = rh ? row_start[(uint32_t)(k[j] >> cbits)] : 0u;
        rcj[j] = rh ? (uint32_t)row_cap[(uint32_t)(k[j] >> cbits)] : 0u;
    }
    const uint32_t ls = head_rows_len(k, f.r, cbits, row_len, lj);
    if (!table) {
        uint32_t sb[3] = {0, 0, 0};
#pragma unroll
        for (int q = 0; q < HEADS_DIRECT_TILES / TB; ++q) {
            const bool in = (int64_t)threadIdx.x + q * TB < tile;
            sb[0] += in ? ch[q] : 0u;
            sb[1] += in ? cr[q] : 0u;
            sb[2] += in ? cl[q] : 0u;
        }
        block_sum_n<3>(sb, lds3);
        ubase = sb[0];
        tbase = sb[1];
        lbase = sb[2];
    }
    uint32_t sc[3] = {(uint32_t)__popc(f.h), (uint32_t)__popc(f.r), ls}, stot[3];
    block_excl_scan_n<3>(sc, lds3, stot);
    const uint32_t htot = stot[0], rtot = stot[1], ltot = stot[2];
    uint32_t lu = sc[0], lt = sc[1], ll = sc[2] + lbase;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        if (!((f.h >> j) & 1u)) continue;
        const bool rh = (f.r >> j) & 1u;
        float s = 0.0f + __uint_as_float(vb[j]);
        bool open = true;
#pragma unroll
        for (int m = j + 1; m < IPT; ++m) {
            if (open && k[m] == k[j]) s = s + __uint_as_float(vb[m]);
            else open = false;
        }
        if (open)
            for (int64_t m = i0 + IPT; m < n && key[m] == k[j]; ++m) s = s + __uint_as_float(val[m]);
        const uint32_t r = (uint32_t)(k[j] >> cbits);
        s_ucol[lu] = (int32_t)(k[j] & ((1ull << cbits) - 1ull));
        s_usum[lu] = s;
        // a coordinate that is not a row head belongs to the last row head
        // before it (possibly in an earlier tile)
        s_ut[lu] = tbase + lt - (rh ? 0u : 1u);
        if (rh) {
            s_trow[lt] = r;
            s_tub[lt] = ubase + lu;
            s_tl[lt] = ll;
            if (pspan)
                for (uint32_t k = (ll + PR_SPAN - 1u) / PR_SPAN; k * PR_SPAN < ll + probe_len(lj[j]); ++k)
                    pspan[k] = tbase + lt;
            ll += probe_len(lj[j]);
            trs[tbase + lt] = rsj[j];
            tlen[tbase + lt] = (uint32_t)lj[j];
            tcap[tbase + lt] = rcj[j];
            if (rowt) rowt[r] = ((uint64_t)rtag << 32) | (uint64_t)(tbase + lt);
            ++lt;
        }
        ++lu;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < htot; i += TB) {
        ucol[ubase + i] = s_ucol[i];
        usum[ubase + i] = s_usum[i];
        ut[ubase + i] = s_ut[i];
    }
    for (uint32_t i = threadIdx.x; i < rtot; i += TB) {
        trow[tbase + i] = s_trow[i];
        tubeg[tbase + i] = s_tub[i];
        tlpre[tbase + i] = s_tl[i];
        tins[tbase + i] = 0;
        tdel[tbase + i] = 0;
    }
    if (threadIdx.x == 0) {
        if (table) {
            if (tile == 0) {
                tubeg[ds->ntouched] = ds->nuniq;
                tlpre[ds->ntouched] = ds->ptotal;
            }
        } else if (tile == gridDim.x - 1) {
            ds->nuniq = ubase + htot;
            ds->ntouched = tbase + rtot;
            ds->ptotal = lbase + ltot;
            tubeg[tbase + rtot] = ubase + htot;
            tlpre[tbase + rtot] = lbase + ltot;
        }
    }
}

// Probe every distinct coordinate in its (column-sorted) row: existing ->
// old + sum (delete when it cancels to 0.0), missing -> insert when the sum
// is not 0.0.  The touched rows, each of probe length max(entries, 1), lie
// end to end in a flat probe space (offsets tlpre, from the heads); each
// wave owns PR_SPAN positions of it.  It stages their columns in LDS with
// every load in flight at once, then each coordinate of the rows it covers
// finds its lower bound there.  A row cut by a span's edge gives each span
// the coordinates whose lower bound lies in it, told apart by comparing
// with the columns at the cut.  So a row is read once, coalesced, and the
// work per wave does not depend on the row-length distribution.  Per-row
// insert / delete counts are summed in LDS and added once per row and span;
// the per-tile code counts are made by the plan.
constexpr uint32_t PR_PER = PR_SPAN / 64;  // staged columns per lane
constexpr int PR_RU = 4;                   // record rounds of 64 in flight
struct ProbeLds {
    int32_t col[PR_SPAN];
    uint32_t pmap[PR_SPAN / 32];        // start map of the span's positions
    uint32_t rmap[64 * PR_RU / 32];     // start map of the rows' coordinates (few)
    uint32_t a[65];    // probe offset of window row j (~0 past the last row)
    uint32_t ub[65];   // first record of window row j (~0 past the last row)
    uint32_t rs[64];   // pool start of row j
    uint32_t len[64];  // entries of row j
    uint32_t ni[64], nd[64];
};

// Code and value of record u, whose lower bound in its row (pool start rs)
// is pos; `match`: the entry there holds the record's column.
__device__ __forceinline__ void probe_finish(uint32_t u, uint32_t pos, bool match, uint32_t rs, float s,
                                             const float *__restrict__ pval, float *__restrict__ uval,
                                             uint32_t *__restrict__ upos, uint32_t *__restrict__ ucode,
                                             uint32_t *ni, uint32_t *nd)
{
    uint32_t code;
    float nv;
    if (match) {
        nv = pval[(uint64_t)rs + pos] + s;
        code = (nv != 0.0f) ? C_ADD : C_DEL;
    } else {
        nv = s;
        code = (s != 0.0f) ? C_INS : C_NOP;
    }
    uval[u] = nv;
    upos[u] = pos;
    ucode[u] = code;
    if (code == C_INS) atomicAdd(ni, 1u);
    else if (code == C_DEL) atomicAdd(nd, 1u);
}

// First index in [lo, hi) whose a[] is > key (UPPER) or >= key (!UPPER),
// for a[] non-decreasing on [lo, hi) (hi when there is none).  By the whole
// wave, 16 loads per lane: one memory round trip when hi - lo <= 1024, one
// more per 1024-fold narrowing.  All lanes must call it; every lane gets the
// result.
template <bool UPPER, class T>
__device__ __forceinline__ uint32_t wave_bound(const T *__restrict__ a, uint32_t lo, uint32_t hi, T key)
{
    const uint32_t lane = (uint32_t)lane_id();
    while (hi - lo > 1024u) {
        const uint32_t step = (hi - lo + 1023u) / 1024u;
        uint32_t c = 0;
#pragma unroll
        for (uint32_t i = 0; i < 16u; ++i) {
            const uint32_t idx = lo + (lane * 16u + i) * step;
            c += (idx < hi && (UPPER ? a[idx] <= key : a[idx] < key)) ? 1u : 0u;
        }
        const uint32_t k = wave_sum_u32(c);  // probes below the bound: a prefix
        if (k == 0) return lo;
        const uint32_t nh = lo + k * step;
        lo = lo + (k - 1u) * step + 1u;
        hi = nh < hi ? nh : hi;
    }
    uint32_t c = 0;
#pragma unroll
    for (uint32_t i = 0; i < 16u; ++i) {
        const uint32_t idx = lo + lane * 16u + i;
        c += (idx < hi && (UPPER ? a[idx] <= key : a[idx] < key)) ? 1u : 0u;
    }
    return lo + wave_sum_u32(c);
}

__global__ __launch_bounds__(TB) void k_upd_probe(const DevState *__restrict__ ds,
                                                  const uint32_t *__restrict__ row_start,
                                                  const int32_t *__restrict__ row_len,
                                                  const int32_t *__restrict__ pcol, const float *__restrict__ pval,
                                                  const uint32_t *__restrict__ tubeg, const uint32_t *__restrict__ tlpre,
                                                  const uint32_t *__restrict__ trs, const uint32_t *__restrict__ tlen,
                                                  const int32_t *__restrict__ ucol, const float *__restrict__ usum,
                                                  float *__restrict__ uval, uint32_t *__restrict__ upos,
                                                  uint32_t *__restrict__ ucode, uint32_t *__restrict__ tins,
                                                  uint32_t *__restrict__ tdel, const uint32_t *__restrict__ pspan)
{
    __shared__ ProbeLds sh[TB / 64];
    ProbeLds &S = sh[threadIdx.x >> 6];
    const uint32_t lane = (uint32_t)lane_id();
    const uint32_t nt = ds->ntouched;
    const uint32_t total = ds->ptotal;
    const uint64_t nwaves = (uint64_t)gridDim.x * (TB / 64);
    for (uint64_t wave = (uint64_t)blockIdx.x * (TB / 64) + (threadIdx.x >> 6); wave * PR_SPAN < total;
         wave += nwaves) {
        const uint32_t e0 = (uint32_t)(wave * PR_SPAN);
        const uint32_t e1 = total - e0 < PR_SPAN ? total : e0 + PR_SPAN;
        // every probe length is >= 1, so exactly one row holds e0
        uint32_t j = pspan ? pspan[wave] : wave_find_last_le(tlpre, nt, e0);
        uint32_t e = e0;
        while (e < e1) {
            // rows j .. j+63 (nt >= 1: the span is non-empty)
            const uint32_t t = j + lane;
            const uint32_t tq = t <= nt ? t : nt;
            const uint32_t tc = t < nt ? t : nt - 1;
            const uint32_t t64 = j + 64u <= nt ? j + 64u : nt;
            const uint32_t a_t = tlpre[tq], a_64 = tlpre[t64];
            const uint32_t ub = tubeg[tq], ub_64 = tubeg[t64];
            const uint32_t rs = trs[tc];
            const uint32_t L = tlen[tc];
            S.a[lane] = t <= nt ? a_t : 0xffffffffu;
            S.ub[lane] = t <= nt ? ub : 0xffffffffu;
            if (lane == 0) {
                S.a[64] = j + 64u <= nt ? a_64 : 0xffffffffu;
                S.ub[64] = j + 64u <= nt ? ub_64 : 0xffffffffu;
            }
            S.rs[lane] = rs;
            S.len[lane] = t < nt ? L : 0u;
            S.ni[lane] = 0;
            S.nd[lane] = 0;
            __builtin_amdgcn_wave_barrier();
            // positions [e, wend): at least one (row j holds e)
            const uint32_t wend = S.a[64] < e1 ? S.a[64] : e1;
            const uint32_t cnt = wend - e;
            const int jA = find_seg64(S.a, e), jB = find_seg64(S.a, wend - 1u);
            const uint32_t kA = e - S.a[jA];
            // the coordinates of rows jA .. jB; when they are few, they are
            // loaded with the columns and the cut rows' ones are told apart
            // by column, else located by two searches
            const uint32_t R0 = S.ub[jA], R1 = S.ub[jB + 1];
            const bool few = R1 - R0 <= 64u * PR_RU;
            // the rows' first positions (and first coordinates) as start
            // maps: a position's row is a popcount away, no search
            if (lane < PR_SPAN / 32) S.pmap[lane] = 0u;
            if (lane < 2 * PR_RU) S.rmap[lane] = 0u;
            __builtin_amdgcn_wave_barrier();
            if (t < nt) {
                if (a_t >= e && a_t < wend) atomicOr(&S.pmap[(a_t - e) >> 5], 1u << ((a_t - e) & 31u));
                if (few && ub >= R0 && ub < R1) atomicOr(&S.rmap[(ub - R0) >> 5], 1u << ((ub - R0) & 31u));
            }
            __builtin_amdgcn_wave_barrier();
            const StartMap pm = start_map_load(S.pmap, PR_SPAN / 64);
            const StartMap rm = start_map_load(S.rmap, PR_RU);
            const int pbase = jA - (kA == 0u ? 1 : 0);
            int32_t rc[PR_RU];
            float rsv[PR_RU];
            // the column before the cut of row jA (a cut row has >= 2 entries)
            int32_t before = 0;
            {
                if (kA > 0u) before = pcol[(uint64_t)S.rs[jA] + kA - 1u];
#pragma unroll
                for (int q = 0; q < PR_RU; ++q) {
                    const uint32_t u = R0 + (uint32_t)q * 64u + lane;
                    rc[q] = 0;
                    rsv[q] = 0.0f;
                    if (few && u < R1) {
                        rc[q] = ucol[u];
                        rsv[q] = usum[u];
                    }
                }
                int32_t cv[PR_PER];
#pragma unroll
                for (uint32_t q = 0; q < PR_PER; ++q) {
                    const uint32_t p = q * 64u + lane;
                    cv[q] = 0x7fffffff;
                    const int jr = pbase + (int)start_map_count(pm, (int)q);
                    if (p < cnt) {
                        const uint32_t k = e + p - S.a[jr];
                        if (k < S.len[jr]) cv[q] = pcol[(uint64_t)S.rs[jr] + k];
                    }
                }
#pragma unroll
                for (uint32_t q = 0; q < PR_PER; ++q) {
                    const uint32_t p = q * 64u + lane;
                    if (p < cnt) S.col[p] = cv[q];
                }
            }
            __builtin_amdgcn_wave_barrier();
            // a coordinate of the span: its lower bound among the row's
            // staged entries [k0, k1), at p0
            auto resolve = [&](uint32_t u, int jr, int32_t c, float sv) {
                const uint32_t ar = S.a[jr];
                const uint32_t k0 = ar < e ? e - ar : 0u;
                const uint32_t lend = ar + S.len[jr];
                const uint32_t k1 = (lend < wend ? lend : wend) - ar;
                const uint32_t p0 = ar + k0 - e;
                uint32_t lo = 0, hi = k1 > k0 ? k1 - k0 : 0u;
                const uint32_t n = hi;
                while (lo < hi) {
                    const uint32_t m = (lo + hi) >> 1;
                    if (S.col[p0 + m] < c) lo = m + 1; else hi = m;
                }
                const bool match = lo < n && S.col[p0 + lo] == c;
                probe_finish(u, k0 + lo, match, S.rs[jr], sv, pval, uval, upos, ucode, &S.ni[jr], &S.nd[jr]);
            };
            // of the row cut by e, the coordinates above its column before the
            // cut; of the row cut by wend, those up to its last staged column
            // (a row cut at both ends: both)
            const bool cutB = S.a[jB + 1] > wend;
            const int32_t lastc = cutB ? S.col[cnt - 1u] : 0;
            if (few) {
                // in three stages over all rounds, so that the searches and
                // then the value loads of every round overlap: lower bounds
                // (a fixed number of halving steps), old values, results
                uint32_t pos[PR_RU], at[PR_RU];
                int jq[PR_RU];
                bool act[PR_RU], mt[PR_RU];
#pragma unroll
                for (int q = 0; q < PR_RU; ++q) {
                    const uint32_t u = R0 + (uint32_t)q * 64u + lane;
                    const int jr = jA - 1 + (int)start_map_count(rm, q);
                    act[q] = u < R1 && !(jr == jA && kA > 0u && rc[q] <= before) &&
                             !(jr == jB && cutB && rc[q] > lastc);
                    jq[q] = jr;
                    const uint32_t ar = S.a[jr];
                    const uint32_t k0 = ar < e ? e - ar : 0u;
                    const uint32_t lend = ar + S.len[jr];
                    const uint32_t k1 = (lend < wend ? lend : wend) - ar;
                    const uint32_t p0 = ar + k0 - e;
                    uint32_t b = p0, len = act[q] && k1 > k0 ? k1 - k0 : 0u;
#pragma unroll
                    for (int it = 0; it < 12; ++it) {  // 2^12 > PR_SPAN
                        const uint32_t half = len >> 1;
                        const uint32_t ix = b + half < PR_SPAN ? b + half : PR_SPAN - 1u;
                        const bool go = len > 0u && S.col[ix] < rc[q];
                        b = go ? b + half + 1u : b;
                        len = go ? len - half - 1u : half;
                    }
                    pos[q] = k0 + (b - p0);
                    mt[q] = act[q] && b - p0 < (k1 > k0 ? k1 - k0 : 0u) && S.col[b < PR_SPAN ? b : PR_SPAN - 1u] == rc[q];
                    at[q] = S.rs[jr] + pos[q];
                }
                float hv[PR_RU];
#pragma unroll
                for (int q = 0; q < PR_RU; ++q) hv[q] = mt[q] ? pval[at[q]] : 0.0f;
#pragma unroll
                for (int q = 0; q < PR_RU; ++q) {
                    if (!act[q]) continue;
                    const uint32_t u = R0 + (uint32_t)q * 64u + lane;
                    uint32_t code;
                    float nv;
                    if (mt[q]) {
                        nv = hv[q] + rsv[q];
                        code = (nv != 0.0f) ? C_ADD : C_DEL;
                    } else {
                        nv = rsv[q];
                        code = (rsv[q] != 0.0f) ? C_INS : C_NOP;
                    }
                    uval[u] = nv;
                    upos[u] = pos[q];
                    ucode[u] = code;
                    if (code == C_INS) atomicAdd(&S.ni[jq[q]], 1u);
                    else if (code == C_DEL) atomicAdd(&S.nd[jq[q]], 1u);
                }
            } else {
                uint32_t rA = R0, rB = R1;
                if (kA > 0u) rA = wave_bound<true, int32_t>(ucol, rA, S.ub[jA + 1], before);
                if (cutB) rB = wave_bound<true, int32_t>(ucol, S.ub[jB] > rA ? S.ub[jB] : rA, rB, lastc);
                for (uint32_t ur = rA; ur < rB; ur += 64u * PR_RU) {
                    int32_t c[PR_RU];
                    float sv[PR_RU];
#pragma unroll
                    for (int q = 0; q < PR_RU; ++q) {
                        const uint32_t u = ur + (uint32_t)q * 64u + lane;
                        c[q] = 0;
                        sv[q] = 0.0f;
                        if (u < rB) {
                            c[q] = ucol[u];
                            sv[q] = usum[u];
                        }
                    }
#pragma unroll
                    for (int q = 0; q < PR_RU; ++q) {
                        const uint32_t u = ur + (uint32_t)q * 64u + lane;
                        if (u < rB) resolve(u, find_seg64(S.ub, u), c[q], sv[q]);
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
            if (t < nt) {
                if (S.ni[lane]) atomicAdd(&tins[t], S.ni[lane]);
                if (S.nd[lane]) atomicAdd(&tdel[t], S.nd[lane]);
            }
            __builtin_amdgcn_wave_barrier();
            e = wend;
            j += 64;
        }
    }
}

// The batch's totals and decision, by one workgroup once every plan and
// probe tile table is final: per-tile plan counters (pplan) and probe adds
// are summed over all ntiles tiles, the claims / moved / inserts / deletes
// totals come in.  The batch goes ahead only if every index was valid and
// bump + claims fits the pool; only then does the bump pointer advance.
__device__ void plan_totals_decide(DevState *ds, unsigned long long pool_cap, const long long *__restrict__ pplan,
                                   const uint32_t *__restrict__ tp, int64_t ntiles, uint32_t claim_total,
                                   uint32_t moved_total, uint32_t ins_total, uint32_t del_total,
                                   long long (*lds)[8])
{
    long long acc[NPLAN + 1] = {0, 0, 0, 0, 0, 0};
#pragma unroll 4
    for (int64_t b = threadIdx.x; b < ntiles; b += TB) {
#pragma unroll
        for (int i = 0; i < NPLAN; ++i) acc[i] += pplan[b * NPLAN + i];
        acc[NPLAN] += tp[b * TP_WORDS + TP_A];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NPLAN + 1; ++i) {
        const long long w = wave_sum_i64(acc[i]);
        if (lane_id() == 0) lds[threadIdx.x >> 6][i] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long tot[NPLAN + 1];
#pragma unroll
        for (int i = 0; i < NPLAN + 1; ++i) tot[i] = (lds[0][i] + lds[1][i]) + (lds[2][i] + lds[3][i]);
        ds->claim_total = claim_total;
        ds->nrewrite_old = moved_total;
        ds->n_ins = ins_total;
        ds->n_add = tot[NPLAN];
        ds->n_del = del_total;
        ds->d_nnz = tot[0];
        ds->d_garbage = tot[1];
        ds->d_hubs = tot[2];
        ds->d_hub_chunks = tot[3];
        ds->hub_touch = (unsigned)tot[4];
        const unsigned long long need = claim_total;
        const unsigned long long base = ds->bump;  // claims start at the current bump
        ds->need = need;
        ds->bump_saved = base;
        const unsigned bad = ds->ubad;
        ds->ubad = 0;  // the next batch's pack starts from a clear flag
        ds->ubad_last = bad;
        if (bad || base + need > pool_cap) {
            ds->go = 0;
        } else {
            ds->go = 1;
            ds->bump = base + need;
        }
    }
}

// Per touched row: new length, mode (values in place / rewrite), chunk
// claim and old entries to move, per-tile sums to the batch tile table and per-tile counters to
// pplan; per tile of coordinates, the probe's insert / add / delete counts
// to the probe table.  fresh: the row lengths and capacities the heads pass
// recorded (tlen, tcap) still hold, so every load is one coalesced round
// trip; otherwise (a re-plan after the pool was compacted or grown) they
// and the row's pool start are re-read from the row table and trs is
// rewritten.  The last workgroup to finish then decides the batch: it makes the
// claim / moved / insert / delete columns exclusive offsets, totals every
// counter, rejects the batch (no-op) on a bad index or when the claims do
// not fit the pool, and otherwise advances the bump pointer.
__global__ __launch_bounds__(TB) void k_upd_plan(DevState *ds, unsigned long long pool_cap,
                                                 const uint32_t *__restrict__ trow, const uint32_t *__restrict__ tins,
                                                 const uint32_t *__restrict__ tdel, const uint32_t *__restrict__ ucode,
                                                 const int32_t *__restrict__ row_len,
                                                 const int32_t *__restrict__ row_cap,
                                                 const uint32_t *__restrict__ row_start, uint32_t *__restrict__ trs,
                                                 uint32_t *__restrict__ tnewl,
                                                 uint32_t *__restrict__ tmode, uint32_t *__restrict__ tmoved,
                                                 uint32_t *__restrict__ tt,
                                                 uint32_t *__restrict__ tp, long long *__restrict__ pplan,
                                                 unsigned int *fin, bool table, bool fresh,
                                                 const uint32_t *__restrict__ tlen, const uint32_t *__restrict__ tcap)
{
    __shared__ long long lds[4][8];
    __shared__ uint32_t lds4[4], lds44[4][4], s_flag;
    const int64_t tile = blockIdx.x;
    const int64_t t0 = tile * TILE + (int64_t)threadIdx.x * IPT;
    const uint32_t nt = ds->ntouched, nu = ds->nuniq;
    uint32_t csum = 0, msum = 0;
    long long v[NPLAN] = {0, 0, 0, 0, 0};
    // loads in two phases (row ids, counts and the tile's codes, then row
    // lengths), so each phase is one memory round trip for all eight rows
    uint32_t rr[IPT], ti[IPT], td[IPT], rl[IPT], rc[IPT], cd[IPT];
    uint32_t rsv[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        rr[j] = ti[j] = td[j] = rl[j] = rc[j] = rsv[j] = 0;
        cd[j] = t0 + j < nu ? ucode[t0 + j] : (uint32_t)C_NOP;
        if (t0 + j < nt) {
            ti[j] = tins[t0 + j];
            td[j] = tdel[t0 + j];
            if (fresh) {
                rl[j] = tlen[t0 + j];
                rc[j] = tcap[t0 + j];
            } else {
                rr[j] = trow[t0 + j];
            }
        }
    }
    if (!fresh) {
#pragma unroll
        for (int j = 0; j < IPT; ++j) {
            if (t0 + j < nt) {
                rl[j] = (uint32_t)row_len[rr[j]];
                rc[j] = (uint32_t)row_cap[rr[j]];
                rsv[j] = row_start[rr[j]];
            }
        }
    }
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const int64_t t = t0 + j;
        if (t >= nt) continue;
        const uint32_t L = rl[j];
        const uint32_t cap = rc[j];
        const uint32_t ni = ti[j], nd = td[j];
        const uint32_t nl = L + ni - nd;
        const uint32_t mode = (ni | nd) ? M_REWRITE : M_VALUES;
        if (mode == M_REWRITE) {
            csum += chunk_round(nl);
            msum += L;
            const int hub_now = nl > DCSR_HUB_LEN, hub_was = L > DCSR_HUB_LEN;
            v[0] += (long long)nl - (long long)L;
            v[1] += (long long)cap;
            v[2] += hub_now - hub_was;
            v[3] += (hub_now ? ((long long)nl + DCSR_HUB_CHUNK - 1) / DCSR_HUB_CHUNK : 0) -
                    (hub_was ? ((long long)L + DCSR_HUB_CHUNK - 1) / DCSR_HUB_CHUNK : 0);
            v[4] += hub_now | hub_was;
        }
        tnewl[t] = nl;
        tmode[t] = mode;
        if (!fresh) trs[t] = rsv[j];
        tmoved[t] = mode == M_REWRITE ? L : 0u;
    }
    // claims, moved, inserts | adds << 16 (each at most TILE), deletes
    uint32_t sums[4] = {csum, msum, 0u, 0u};
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        sums[2] += (cd[j] == C_INS ? 1u : 0u) + (cd[j] == C_ADD ? 0x10000u : 0u);
        sums[3] += cd[j] == C_DEL ? 1u : 0u;
    }
    block_sum_n<4>(sums, lds44);
    csum = sums[0];
    msum = sums[1];
    const uint32_t cia = sums[2], cdl = sums[3];
#pragma unroll
    for (int i = 0; i < NPLAN; ++i) {
        const long long w = wave_sum_i64(v[i]);
        if (lane_id() == 0) lds[threadIdx.x >> 6][i] = w;
    }
    __syncthreads();
    if (threadIdx.x < NPLAN) {
        const int i = threadIdx.x;
        pplan[tile * NPLAN + i] = ((lds[0][i] + lds[1][i]) + (lds[2][i] + lds[3][i]));
    }
    if (threadIdx.x == 0) {
        tt[tile * TT_WORDS + TT_C] = csum;
        tt[tile * TT_WORDS + TT_M] = msum;
        tp[tile * TP_WORDS + TP_I] = cia & 0xffffu;
        tp[tile * TP_WORDS + TP_A] = cia >> 16;
        tp[tile * TP_WORDS + TP_D] = cdl;
    }
    // without a table scan (at most HEADS_DIRECT_TILES tiles) the offsets
    // pass sums the counts before each tile and decides the batch
    if (!table || !last_block(&fin[FC_PLAN], &s_flag)) return;

    const int64_t ntiles = gridDim.x;
    uint32_t *const tabs[4] = {tt, tt, tp, tp};
    uint32_t ctot[4];
    block_scan_columns<4>(tabs, {TT_C, TT_M, TP_I, TP_D}, {TT_CX, TT_MX, TP_IX, TP_DX}, ntiles, lds4, ctot);
    const uint32_t claim_total = ctot[0], moved_total = ctot[1], ins_total = ctot[2], del_total = ctot[3];
    plan_totals_decide(ds, pool_cap, pplan, tp, ntiles, claim_total, moved_total, ins_total, del_total, lds);
}

// Exclusive offsets per entry: upre[2u], upre[2u+1] = inserts / deletes
// before coordinate u (u <= nuniq; entry nuniq holds the totals) and
// tpre[2t], tpre[2t+1] = chunk slots claimed / old entries moved before
// touched row t (t <= ntouched), from the in-tile prefixes plus the tile
// offsets; (mspan) touched row t for every move span that starts in it.
__global__ __launch_bounds__(TB) void k_upd_offsets(DevState *__restrict__ ds,
                                                    const uint32_t *__restrict__ ucode,
                                                    const uint32_t *__restrict__ tmode,
                                                    const uint32_t *__restrict__ tnewl,
                                                    const uint32_t *__restrict__ tmoved,
                                                    const uint32_t *__restrict__ tt, const uint32_t *__restrict__ tp,
                                                    uint32_t *__restrict__ upre, uint32_t *__restrict__ tpre,
                                                    const long long *__restrict__ pplan, unsigned long long pool_cap,
                                                    bool table, uint32_t *__restrict__ mspan)
{
    __shared__ uint32_t lds44[4][4];
    __shared__ long long lds[4][8];
    const int64_t tile = blockIdx.x;
    const int64_t e0 = tile * TILE + (int64_t)threadIdx.x * IPT;
    // tile bases: scanned by the plan's last workgroup (table), or summed
    // here from the counts of the tiles before this one (loaded with the
    // entries: up to four tiles per thread, index clamped to tile 0, masked)
    constexpr int QT = HEADS_DIRECT_TILES / TB;
    uint32_t qc[QT], qm[QT], qi[QT], qd[QT];
    if (!table) {
#pragma unroll
        for (int q = 0; q < QT; ++q) {
            const int64_t pt = (int64_t)threadIdx.x + q * TB;
            const int64_t ix = pt < tile ? pt : 0;
            qc[q] = tt[ix * TT_WORDS + TT_C];
            qm[q] = tt[ix * TT_WORDS + TT_M];
            qi[q] = tp[ix * TP_WORDS + TP_I];
            qd[q] = tp[ix * TP_WORDS + TP_D];
        }
    }
    const uint32_t nu = ds->nuniq, nt = ds->ntouched;
    uint32_t ins[IPT], del[IPT], cl[IPT], mv[IPT];
    uint32_t code[IPT], md[IPT], nl[IPT];
    // every load first (one memory round trip), then the counts
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const int64_t e = e0 + j;
        code[j] = C_NOP;
        md[j] = M_VALUES;
        nl[j] = mv[j] = 0;
        if (e < nu) code[j] = ucode[e];
        if (e < nt) {
            md[j] = tmode[e];
            nl[j] = tnewl[e];
            mv[j] = tmoved[e];
        }
    }
    uint32_t si = 0, sd = 0, sc = 0, sm = 0;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        ins[j] = code[j] == C_INS;
        del[j] = code[j] == C_DEL;
        cl[j] = md[j] == M_REWRITE ? chunk_round(nl[j]) : 0u;
        si += ins[j];
        sd += del[j];
        sc += cl[j];
        sm += mv[j];
    }
    uint32_t bi, bd, bc, bm;
    if (table) {
        bi = tp[tile * TP_WORDS + TP_IX];
        bd = tp[tile * TP_WORDS + TP_DX];
        bc = tt[tile * TT_WORDS + TT_CX];
        bm = tt[tile * TT_WORDS + TT_MX];
    } else {
        uint32_t a[4] = {0, 0, 0, 0};
#pragma unroll
        for (int q = 0; q < QT; ++q) {
            const bool in = (int64_t)threadIdx.x + q * TB < tile;
            a[0] += in ? qi[q] : 0u;
            a[1] += in ? qd[q] : 0u;
            a[2] += in ? qc[q] : 0u;
            a[3] += in ? qm[q] : 0u;
        }
        block_sum_n<4>(a, lds44);
        bi = a[0];
        bd = a[1];
        bc = a[2];
        bm = a[3];
    }
    uint32_t sv[4] = {si, sd, sc, sm}, stot[4];
    block_excl_scan_n<4>(sv, lds44, stot);
    const uint32_t ti = stot[0], td = stot[1], tc = stot[2], tm = stot[3];
    si = sv[0] + bi;
    sd = sv[1] + bd;
    sc = sv[2] + bc;
    sm = sv[3] + bm;
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
        const int64_t e = e0 + j;
        if (e <= nu) {
            upre[2 * e] = si;
            upre[2 * e + 1] = sd;
        }
        if (e <= nt) {
            tpre[2 * e] = sc;
            tpre[2 * e + 1] = sm;
            if (mspan && e < nt)
                for (uint32_t k = (sm + MV_SPAN - 1u) / MV_SPAN; k * MV_SPAN < sm + mv[j]; ++k) mspan[k] = (uint32_t)e;
        }
        si += ins[j];
        sd += del[j];
        sc += cl[j];
        sm += mv[j];
    }
    // the last tile knows every total: its bases plus its own counts
    if (!table && tile == gridDim.x - 1)
        plan_totals_decide(ds, pool_cap, pplan, tp, gridDim.x, bc + tc, bm + tm, bi + ti, bd + td, lds);
}

// Move surviving old entries of rewritten rows to their merged positions
// (with their old values: k_upd_write stores the new ones of added-to
// entries).
// tpre[2t] = claim offset of row t (relative to bump_saved), tpre[2t+1] =
// offset of row t's old entries in the flat move space (rows written in
// place hold none).  Each wave owns MV_SPAN entries of that space, taken in
// chunks of at most 64 rows.  The loads of a chunk's entries are issued
// first; while they are in flight, the update records of the chunk's rows
// mark the entries they shift (an insert +1 on the entry it goes before, a
// delete a tombstone on its entry and -1 on the row's next one) and a
// per-row scan of the marks, in 64-wide rounds, gives
// every entry its shift.  So the copy is a coalesced stream with no
// per-entry search.  A row cut by the chunk's start begins from the shift
// of its records before the cut, located by a search on their positions.
constexpr uint32_t MV_PER = MV_SPAN / 64;  // entries per lane
constexpr int MV_RU = 4;                   // record rounds of 64 in flight
// The SpMV's flat row scan (rowscan.h) run by the first `ntiles`
// workgroups of the move kernel over the row lengths the update commits:
// row r's new length is tnewl[t] when rowt[r] carries this update's tag
// (touched row t) and the batch goes ahead, else row_len[r].  ntiles = 0:
// no scan (the SpMV scans on its own).
struct RowScanArgs {
    const int32_t *row_len;
    const uint64_t *rowt;
    const uint32_t *tnewl;
    uint32_t *foff, *rowmap, *part;
    unsigned int *fault;
    uint32_t R, rtag, ntiles;
    int limit;
};

__device__ void move_rowscan(const DevState *__restrict__ ds, const RowScanArgs &a, RowScanLds &L)
{
    const bool go = ds->go != 0;
    const uint32_t R = a.R;
    auto lens = [&](int64_t r0, bool full, uint32_t(&v)[RS_IPT]) {
        uint64_t x[RS_IPT];
        uint32_t l[RS_IPT];
        if (full) {
            const int4 *p = (const int4 *)(a.row_len + r0);
            const ulonglong2 *q = (const ulonglong2 *)(a.rowt + r0);
#pragma unroll
            for (int k = 0; k < RS_IPT / 4; ++k) {
                const int4 t = p[k];
                l[4 * k] = (uint32_t)t.x;
                l[4 * k + 1] = (uint32_t)t.y;
                l[4 * k + 2] = (uint32_t)t.z;
                l[4 * k + 3] = (uint32_t)t.w;
            }
#pragma unroll
            for (int k = 0; k < RS_IPT / 2; ++k) {
                const ulonglong2 t = q[k];
                x[2 * k] = t.x;
                x[2 * k + 1] = t.y;
            }
        } else {
#pragma unroll
            for (int j = 0; j < RS_IPT; ++j) {
                const bool in = r0 + j < (int64_t)R;
                l[j] = in ? (uint32_t)a.row_len[r0 + j] : 0u;
                x[j] = in ? a.rowt[r0 + j] : 0ull;
            }
        }
#pragma unroll
        for (int j = 0; j < RS_IPT; ++j)
            v[j] = go && (uint32_t)(x[j] >> 32) == a.rtag ? a.tnewl[(uint32_t)x[j]] : l[j];
    };
    // fallback: the new lengths of every earlier tile's rows
    auto prefix = [&](int64_t n) {
        uint32_t s = 0;
        for (int64_t i = lane_id(); i < n; i += 64) {
            const uint64_t x = a.rowt[i];
            s += go && (uint32_t)(x >> 32) == a.rtag ? a.tnewl[(uint32_t)x] : (uint32_t)a.row_len[i];
        }
        return wave_sum_u32(s);
    };
    rowscan_tile(R, lens, prefix, a.foff, a.rowmap, a.part, a.fault, a.limit, a.ntiles, L);
}

struct MoveLds {
    int32_t mk[MV_SPAN];             // marks
    uint32_t tomb[MV_SPAN / 32];     // deleted entries
    uint32_t pmap[MV_SPAN / 32];     // start map of the chunk's rows with entries
    uint32_t rmap[64 * MV_RU / 32];  // start map of the rows' records (few)
    uint32_t tab[64];                // window row of the k-th start in pmap
    uint32_t off[65];      // move-space start of window row j (~0 past the last row)
    uint32_t ub[65];       // first record of window row j (~0 past the last row)
    uint32_t src[64];      // old pool start of row j
    uint32_t dst[64];      // claim offset of row j
};

__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(4))) void k_upd_move(const DevState *__restrict__ ds, const uint32_t *__restrict__ tpre,
                                                 const uint32_t *__restrict__ trs,
                                                 const uint32_t *__restrict__ tubeg,
                                                 const uint32_t *__restrict__ ucode, const uint32_t *__restrict__ upos,
                                                 const uint32_t *__restrict__ upre, int32_t *__restrict__ pcol,
                                                 float *__restrict__ pval, const uint32_t *__restrict__ mspan,
                                                 RowScanArgs rsa)
{
    __shared__ union {
        MoveLds mv[TB / 64];
        RowScanLds rs;
    } sh;
    // the first rsa.ntiles workgroups scan the new row lengths for the SpMV
    // (whether or not the batch goes ahead); the others move entries
    if (blockIdx.x < rsa.ntiles) {
        move_rowscan(ds, rsa, sh.rs);
        return;
    }
    if (!ds->go) return;
    MoveLds &S = sh.mv[threadIdx.x >> 6];
    const uint32_t total = ds->nrewrite_old;
    const uint32_t nt = ds->ntouched;
    const uint64_t base = ds->bump_saved;
    const uint32_t lane = (uint32_t)lane_id();
    const uint64_t nwaves = (uint64_t)(gridDim.x - rsa.ntiles) * (TB / 64);
    for (uint64_t wave = (uint64_t)(blockIdx.x - rsa.ntiles) * (TB / 64) + (threadIdx.x >> 6); wave * MV_SPAN < total;
         wave += nwaves) {
        const uint32_t e0 = (uint32_t)(wave * MV_SPAN);
        const uint32_t e1 = total - e0 < MV_SPAN ? total : e0 + MV_SPAN;
        uint32_t j = mspan ? mspan[wave] : wave_find_last_le<2>(tpre + 1, nt, e0);
        uint32_t e = e0;
        while (e < e1) {
            // rows j .. j+63 (nt >= 1 here: the range is non-empty); lanes
            // past the last row keep ~0
            const uint32_t t = j + lane;
            const uint32_t tq = t <= nt ? t : nt;     // tables of nt + 1 entries
            const uint32_t tc = t < nt ? t : nt - 1;  // tables of nt entries
            const uint32_t t64 = j + 64u <= nt ? j + 64u : nt;
            const uint32_t off_t = tpre[2 * (uint64_t)tq + 1];
            const uint32_t off_64 = tpre[2 * (uint64_t)t64 + 1];
            const uint32_t ub = tubeg[tq];
            const uint32_t ub_64 = tubeg[t64];
            const uint32_t dst = tpre[2 * (uint64_t)tc];
            const uint32_t src = trs[tc];
            S.off[lane] = t <= nt ? off_t : 0xffffffffu;
            S.ub[lane] = t <= nt ? ub : 0xffffffffu;
            if (lane == 0) {
                S.off[64] = j + 64u <= nt ? off_64 : 0xffffffffu;
                S.ub[64] = j + 64u <= nt ? ub_64 : 0xffffffffu;
            }
            S.src[lane] = src;
            S.dst[lane] = dst;
            __builtin_amdgcn_wave_barrier();
            const uint32_t wend = S.off[64] < e1 ? S.off[64] : e1;
            const uint32_t cnt = wend - e;
            if (cnt == 0u) {  // 64 rows without old entries to move
                __builtin_amdgcn_wave_barrier();
                j += 64;
                continue;
            }
            // old entries of window row j
            auto old_len = [&](int jr) { return S.off[jr + 1] - S.off[jr]; };
            const int jA = find_seg64(S.off, e), jB = find_seg64(S.off, wend - 1u);
            const uint32_t kA = e - S.off[jA];
            const uint32_t R0 = S.ub[jA], R1 = S.ub[jB + 1];
            const bool few = R1 - R0 <= 64u * MV_RU;
            // start maps: of the chunk's rows with entries (their ranks to
            // rows in tab; rows without entries share a start) and of the
            // rows' records (every touched row has one), so that an entry's
            // or a record's row is a popcount away, no search
            if (lane < MV_SPAN / 32) {
                S.pmap[lane] = 0u;
                S.tomb[lane] = 0u;
            }
            if (lane < 2 * MV_RU) S.rmap[lane] = 0u;
            __builtin_amdgcn_wave_barrier();
            const bool starts_here = t < nt && off_t >= e && off_t < wend && old_len((int)lane) > 0u;
            if (starts_here) atomicOr(&S.pmap[(off_t - e) >> 5], 1u << ((off_t - e) & 31u));
            if (few && t < nt && ub >= R0 && ub < R1) atomicOr(&S.rmap[(ub - R0) >> 5], 1u << ((ub - R0) & 31u));
            __builtin_amdgcn_wave_barrier();
            const StartMap pm = start_map_load(S.pmap, MV_SPAN / 64);
            const StartMap rm = start_map_load(S.rmap, MV_RU);
            const uint32_t rank = start_map_rank(pm, starts_here ? off_t - e : 0u);  // every lane takes part
            if (starts_here) S.tab[rank] = lane;
            __builtin_amdgcn_wave_barrier();
            // the chunk's entries: row (8 bits each, four to a word) and
            // the loads (bit q of `real`: position q holds an entry)
            int32_t c[MV_PER];
            float v[MV_PER];
            uint32_t jpk[MV_PER / 4];
            uint32_t real = 0;
#pragma unroll
            for (uint32_t q = 0; q < MV_PER; ++q) {
                const uint32_t p = q * 64u + lane;
                c[q] = 0;
                v[q] = 0.0f;
                uint32_t jr = 64u;  // past the chunk: above every row
                const uint32_t nst = start_map_count(pm, (int)q);
                if (p < cnt) {
                    jr = nst == 0u ? (uint32_t)jA : S.tab[nst - 1u];
                    const uint32_t k = e + p - S.off[jr];
                    if (k < old_len((int)jr)) {
                        real |= 1u << q;
                        const uint64_t sp = (uint64_t)S.src[jr] + k;
                        c[q] = pcol[sp];
                        v[q] = pval[sp];
                    }
                }
                if ((q & 3u) == 0u) jpk[q >> 2] = 0u;
                jpk[q >> 2] |= jr << (8u * (q & 3u));
            }
            uint32_t kr[MV_RU], cr[MV_RU];
#pragma unroll
            for (int q = 0; q < MV_RU; ++q) {
                const uint32_t u = R0 + (uint32_t)q * 64u + lane;
                kr[q] = cr[q] = 0;
                if (few && u < R1) {
                    kr[q] = upos[u];
                    cr[q] = ucode[u];
                }
            }
            for (uint32_t p = lane; p < cnt; p += 64u) S.mk[p] = 0;
            __builtin_amdgcn_wave_barrier();
            // a record of window row jr whose mark falls in [e, wend)
            auto mark = [&](uint32_t u, int jr, uint32_t k, uint32_t code) {
                const uint32_t rlen = old_len(jr);
                const uint32_t pa = S.off[jr] + k;
                // rows without old entries to move (written in place, or
                // new) have rlen 0; inserts past a row's end shift nothing
                if (k < rlen && pa >= e && pa < wend) {
                    const uint32_t p = pa - e;
                    if (code == C_INS) {
                        atomicAdd(&S.mk[p], 1);
                    } else if (code == C_DEL) {
                        atomicOr(&S.tomb[p >> 5], 1u << (p & 31u));
                        if (k + 1u < rlen && p + 1u < cnt) atomicAdd(&S.mk[p + 1u], -1);
                    }
                }
            };
            // the cut row jA starts from the shift of its records before the cut
            int32_t bshift = 0;
            if (few) {
                int32_t bs = 0;
#pragma unroll
                for (int q = 0; q < MV_RU; ++q) {
                    const uint32_t u = R0 + (uint32_t)q * 64u + lane;
                    const int jr = jA - 1 + (int)start_map_count(rm, q);
                    if (u >= R1) continue;
                    if (jr == jA && kr[q] < kA) bs += (cr[q] == C_INS ? 1 : 0) - (cr[q] == C_DEL ? 1 : 0);
                    mark(u, jr, kr[q], cr[q]);
                }
                bshift = (int32_t)wave_sum_u32((uint32_t)bs);
            } else {
                uint32_t rA = R0;
                if (kA > 0u) {
                    rA = wave_bound<false, uint32_t>(upos, R0, S.ub[jA + 1], kA);
                    bshift = (int32_t)((upre[2 * (uint64_t)rA] - upre[2 * (uint64_t)R0]) -
                                       (upre[2 * (uint64_t)rA + 1] - upre[2 * (uint64_t)R0 + 1]));
                }
                uint32_t rB = R1;
                if (S.off[jB + 1] > wend)
                    rB = wave_bound<false, uint32_t>(upos, S.ub[jB] > rA ? S.ub[jB] : rA, rB, wend - S.off[jB]);
                for (uint32_t ur = rA; ur < rB; ur += 64u * MV_RU) {
                    uint32_t kk[MV_RU], cc[MV_RU];
#pragma unroll
                    for (int q = 0; q < MV_RU; ++q) {
                        const uint32_t u = ur + (uint32_t)q * 64u + lane;
                        kk[q] = cc[q] = 0;
                        if (u < rB) {
                            kk[q] = upos[u];
                            cc[q] = ucode[u];
                        }
                    }
#pragma unroll
                    for (int q = 0; q < MV_RU; ++q) {
                        const uint32_t u = ur + (uint32_t)q * 64u + lane;
                        if (u < rB) mark(u, find_seg64(S.ub, u), kk[q], cc[q]);
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
            // per-row inclusive scan of the marks in rounds of 64 entries,
            // carrying the last lane's sum into the next round while the row
            // continues; the cut row jA starts from bshift
            int32_t carry = bshift;
            uint32_t carry_row = (uint32_t)jA;
            uint32_t dq[MV_PER];  // destination: the row's claim + offset in the row + shift
#pragma unroll
            for (uint32_t q = 0; q < MV_PER; ++q) {
                const uint32_t p = q * 64u + lane;
                const uint32_t jr = (jpk[q >> 2] >> (8u * (q & 3u))) & 0xffu;
                const bool in = p < cnt;
                int32_t m = in ? S.mk[p] : 0;
                if (lane == 0 && jr == carry_row) m += carry;
                const int32_t inc = (int32_t)wave_seg_incl_scan_u32((uint32_t)m, jr);
                dq[q] = in ? S.dst[jr] + (e + p - S.off[jr]) + (uint32_t)inc : 0u;
                carry = (int32_t)wave_read_lane_u32((uint32_t)inc, 63);
                carry_row = wave_read_lane_u32(jr, 63);
            }
#pragma unroll
            for (uint32_t q = 0; q < MV_PER; ++q) {
                const uint32_t p = q * 64u + lane;
                if (!((real >> q) & 1u) || ((S.tomb[p >> 5] >> (p & 31u)) & 1u)) continue;
                pcol[base + dq[q]] = c[q];
                pval[base + dq[q]] = v[q];
            }
            __builtin_amdgcn_wave_barrier();
            e = wend;
            j += 64;
        }
    }
}

// A NOSYNC update's state snapshot into coherent host memory: the values,
// then the sequence number with a release at system scope, so a host that
// reads seq k then reads values of update k or later.
__device__ __forceinline__ void snap_publish(HostSnap *p, unsigned long long nnz, unsigned long long bump, uint64_t seq)
{
    __hip_atomic_store(&p->nnz, (uint64_t)nnz, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&p->bump, (uint64_t)bump, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&p->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Inserted entries, in-place value updates, row-table commit.  A rejected
// batch writes nothing; with `report` (a caller that does not read the
// decision back) its cause becomes a sticky status bit.  The per-coordinate
// loads (in bounds for every u < n) are issued before the decision is read,
// then the per-row ones: two round trips before the writes.
constexpr int WR_PER = 4;  // coordinates per thread of the write pass
__global__ __launch_bounds__(TB) void k_upd_write(DevState *ds, uint32_t report, int64_t n,
                                                  const uint32_t *__restrict__ ut,
                                                  const uint32_t *__restrict__ trow, const uint32_t *__restrict__ tubeg,
                                                  const uint32_t *__restrict__ tpre,
                                                  const uint32_t *__restrict__ tnewl,
                                                  const uint32_t *__restrict__ tmode, const int32_t *__restrict__ ucol,
                                                  const float *__restrict__ uval, const uint32_t *__restrict__ upos,
                                                  const uint32_t *__restrict__ ucode,
                                                  const uint32_t *__restrict__ upre, int32_t *__restrict__ pcol,
                                                  float *__restrict__ pval, uint32_t *__restrict__ row_start,
                                                  int32_t *__restrict__ row_len, int32_t *__restrict__ row_cap,
                                                  HostSnap *snap, uint64_t seq)
{
    // WR_PER coordinates per thread, TB apart, so that each of the pass's
    // three rounds of loads (the records, their rows' constants, the row
    // starts or prefixes) has every coordinate's loads in flight at once
    const int64_t u0 = (int64_t)blockIdx.x * (TB * WR_PER) + threadIdx.x;
    uint32_t t[WR_PER], code[WR_PER], pos[WR_PER], ui[WR_PER], ud[WR_PER];
    int32_t col[WR_PER];
    float v[WR_PER];
#pragma unroll
    for (int k = 0; k < WR_PER; ++k) {
        const int64_t u = u0 + (int64_t)k * TB;
        t[k] = code[k] = pos[k] = ui[k] = ud[k] = 0;
        code[k] = C_NOP;
        col[k] = 0;
        v[k] = 0.0f;
        if (u < n) {
            t[k] = ut[u];
            code[k] = ucode[u];
            pos[k] = upos[u];
            v[k] = uval[u];
            col[k] = ucol[u];
            ui[k] = upre[2 * u];
            ud[k] = upre[2 * u + 1];
        }
    }
    if (!ds->go) {
        if (report && u0 == 0) ds->status |= ds->ubad_last ? ST_BAD_INDEX : ST_POOL_EXHAUSTED;
        if (snap && u0 == 0) snap_publish(snap, ds->nnz, ds->bump, seq);
        return;
    }
    const uint32_t nu = ds->nuniq;
    if (u0 == 0) {
        const long long nnz = ds->nnz + ds->d_nnz;
        ds->nnz = nnz;
        ds->garbage += ds->d_garbage;
        if (snap) snap_publish(snap, (unsigned long long)nnz, ds->bump, seq);
    }
    if (u0 >= nu) return;
    const uint64_t bsaved = ds->bump_saved;
    bool in[WR_PER];
    uint32_t mode[WR_PER], r[WR_PER], ub[WR_PER], tp[WR_PER], nl[WR_PER];
#pragma unroll
    for (int k = 0; k < WR_PER; ++k) {
        in[k] = u0 + (int64_t)k * TB < (int64_t)nu;
        const uint32_t tk = in[k] ? t[k] : 0u;
        mode[k] = tmode[tk];
        r[k] = trow[tk];
        ub[k] = tubeg[tk];
        tp[k] = tpre[2 * tk];
        nl[k] = tnewl[tk];
    }
    // values in place need the row's pool start, a rewritten row's entries
    // the inserts / deletes before the row
    uint32_t a0[WR_PER], a1[WR_PER];
#pragma unroll
    for (int k = 0; k < WR_PER; ++k) {
        a0[k] = a1[k] = 0;
        if (!in[k]) continue;
        if (mode[k] == M_VALUES) {
            if (code[k] == C_ADD) a0[k] = row_start[r[k]];
        } else {
            a0[k] = upre[2 * ub[k]];
            a1[k] = upre[2 * ub[k] + 1];
        }
    }
#pragma unroll
    for (int k = 0; k < WR_PER; ++k) {
        if (!in[k]) continue;
        if (mode[k] == M_VALUES) {
            if (code[k] == C_ADD) pval[(uint64_t)a0[k] + pos[k]] = v[k];
            continue;
        }
        const uint64_t dest = bsaved + tp[k];
        if (code[k] == C_INS || code[k] == C_ADD) {
            // the inserts / deletes before it in the row shift it (an
            // added-to entry was moved there with its old value)
            const int64_t shift = (int64_t)(ui[k] - a0[k]) - (int64_t)(ud[k] - a1[k]);
            const uint64_t d = dest + (uint64_t)((int64_t)pos[k] + shift);
            if (code[k] == C_INS) pcol[d] = col[k];
            pval[d] = v[k];
        }
        if (u0 + (int64_t)k * TB == (int64_t)ub[k]) {
            row_start[r[k]] = (uint32_t)dest;
            row_len[r[k]] = (int32_t)nl[k];
            row_cap[r[k]] = (int32_t)chunk_round(nl[k]);
        }
    }
}

__global__ void k_fill_u64(uint64_t *p, int64_t n, uint64_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

struct UpdBufs {
    uint64_t *k0, *k1;
    uint32_t *v0, *v1, *hist, *part;
    int32_t *ucol;
    float *usum, *uval;
    uint32_t *ut, *upos, *ucode, *upre;
    uint32_t *trow, *tubeg, *tlpre, *trs, *tlen, *tcap, *tins, *tdel, *tnewl, *tmode, *tmoved, *tpre;
    uint32_t *tt, *tp, *fin;
    // first touched row of every probe / move span (null: searched for)
    uint32_t *pspan, *mspan;
    // tagged row -> touched-row map for the move kernel's row scan (null: no
    // scan there)
    uint64_t *rowt;
    long long *pplan;
    // sort: row groups (lb = gshift + cbits > 0, ndig = groups + 1) or, for
    // batches past GS_MAX_N triples, the LSD radix sort (lb = 0)
    int lb, ndig;
};

__global__ void k_upd_validate_begin(DevState *ds) { ds->ubad = 0; }


// Tiles of the per-entry kernels (entries 0..n: index n carries totals).
static inline unsigned ptiles(int64_t n) { return (unsigned)((n + 1 + TILE - 1) / TILE); }

// Row-group sort plan of a batch of n triples: groups of 2^gshift rows, as
// few rows per group as keeps the group count within GS_MAX_DIGITS - 1 and
// the mean group near GS_TARGET pairs (groups are sorted in LDS up to
// GS_CAP pairs; power-law batches drawn by degree peak near 3x the mean).
constexpr int64_t GS_MAX_N = 1 << 21;
#ifndef DCSR_GS_TARGET
#define DCSR_GS_TARGET 512
#endif
constexpr int64_t GS_TARGET = DCSR_GS_TARGET;
void sort_plan(dcsr_handle h, int64_t n, int &lb, int &ndig)
{
    lb = 0;
    ndig = 0;
    if (n > GS_MAX_N || n < 1) return;
    int64_t want = n / GS_TARGET;
    if (want < 1) want = 1;
    if (want > GS_MAX_DIGITS - 1) want = GS_MAX_DIGITS - 1;
    int gshift = 0;
    while (((h->rows + (int64_t(1) << gshift) - 1) >> gshift) > want) ++gshift;
    const int64_t groups = (h->rows + (int64_t(1) << gshift) - 1) >> gshift;
    lb = gshift + h->cbits;
    ndig = (int)groups + 1;
}

// Where an update's scratch comes from: the handle's grow-only workspace
// (base == nullptr), or a fixed device region carved in order at 256-byte
// alignment (the allocation-free entry point).  With `measure`, nothing is
// handed out and `used` only counts the bytes the region must hold.
struct BufSrc {
    dcsr_handle h;
    uint8_t *base;
    size_t cap, used;
    bool measure;
    hipStream_t s;  // orders the zero fill of grown control-word slots
};

dcsr_status take(BufSrc &src, WsSlot slot, size_t bytes, void **out)
{
    if (!src.base && !src.measure) {
        if (slot == WS_SCAN || slot == WS_TOLDS) return ws_get_zeroed(src.h, slot, bytes, out, src.s);
        return ws_get(src.h, slot, bytes, out);
    }
    const size_t at = (src.used + 255) & ~size_t(255);
    src.used = at + bytes;
    if (src.measure) {
        *out = nullptr;
        return DCSR_OK;
    }
    if (src.used > src.cap) return DCSR_ERR_CAPACITY;
    *out = src.base + at;
    return DCSR_OK;
}

dcsr_status upd_bufs_src(BufSrc &src, int64_t n, UpdBufs &b)
{
    dcsr_handle h = src.h;
    const size_t n4 = (size_t)(n + 2) * 4, n8 = (size_t)(n + 2) * 8;
    const size_t nt = ptiles(n);
    sort_plan(h, n, b.lb, b.ndig);
    size_t hw = radix_hist_words(n);
    if (b.lb > 0 && gs_hist_words(n, b.ndig) > hw) hw = gs_hist_words(n, b.ndig);
    DCSR_TRY(take(src, WS_KEY_A, n8, (void **)&b.k0));
    DCSR_TRY(take(src, WS_KEY_B, n8, (void **)&b.k1));
    DCSR_TRY(take(src, WS_VAL_A, n4, (void **)&b.v0));
    DCSR_TRY(take(src, WS_VAL_B, n4, (void **)&b.v1));
    DCSR_TRY(take(src, WS_HIST, hw * 4, (void **)&b.hist));
    DCSR_TRY(take(src, WS_SCAN, radix_part_words(n) * 4, (void **)&b.part));
    DCSR_TRY(take(src, WS_UCOL, n4, (void **)&b.ucol));
    DCSR_TRY(take(src, WS_UPOS, n4, (void **)&b.upos));
    DCSR_TRY(take(src, WS_UVAL, n4, (void **)&b.uval));
    DCSR_TRY(take(src, WS_UT, n4, (void **)&b.ut));
    DCSR_TRY(take(src, WS_UCODE, n4, (void **)&b.ucode));
    DCSR_TRY(take(src, WS_UPRE, n8, (void **)&b.upre));
    DCSR_TRY(take(src, WS_TROW, n4, (void **)&b.trow));
    DCSR_TRY(take(src, WS_TUBEG, n4 + 4, (void **)&b.tubeg));
    DCSR_TRY(take(src, WS_UROW, n4 + 4, (void **)&b.tlpre));
    DCSR_TRY(take(src, WS_TRS, n4, (void **)&b.trs));
    DCSR_TRY(take(src, WS_TLEN, n4, (void **)&b.tlen));
    DCSR_TRY(take(src, WS_TCAP, n4, (void **)&b.tcap));
    DCSR_TRY(take(src, WS_TINS, n4, (void **)&b.tins));
    DCSR_TRY(take(src, WS_TDEL, n4, (void **)&b.tdel));
    DCSR_TRY(take(src, WS_TNEWL, n4, (void **)&b.tnewl));
    DCSR_TRY(take(src, WS_TMODE, n4, (void **)&b.tmode));
    DCSR_TRY(take(src, WS_TMOVED, n4, (void **)&b.tmoved));
    DCSR_TRY(take(src, WS_TPAIRPRE, n8, (void **)&b.tpre));
    DCSR_TRY(take(src, WS_TMP2, n4, (void **)&b.usum));
    DCSR_TRY(take(src, WS_TDEST, nt * TT_WORDS * 4, (void **)&b.tt));
    DCSR_TRY(take(src, WS_PPROBE, nt * TP_WORDS * 4, (void **)&b.tp));
    DCSR_TRY(take(src, WS_PPLAN, nt * NPLAN * 8, (void **)&b.pplan));
    // finish counters: zero-filled on allocation, reset by their last user
    DCSR_TRY(take(src, WS_TOLDS, 64, (void **)&b.fin));
    // span tables grow with the matrix, so a fixed region goes without them
    b.pspan = b.mspan = nullptr;
    b.rowt = nullptr;
    if (!src.base && !src.measure) {
        const int64_t live = live_bound(h);
        DCSR_TRY(take(src, WS_PSPAN, (size_t)((live + n) / PR_SPAN + 2) * 4, (void **)&b.pspan));
        DCSR_TRY(take(src, WS_MSPAN, (size_t)(live / MV_SPAN + 2) * 4, (void **)&b.mspan));
    }
    return DCSR_OK;
}

dcsr_status upd_bufs(dcsr_handle h, int64_t n, UpdBufs &b, hipStream_t s)
{
    BufSrc src{h, nullptr, 0, 0, false, s};
    DCSR_TRY(upd_bufs_src(src, n, b));
    // zeroed when first allocated: no entry carries a tag before the first
    // update's
    if (h->rows > 0) DCSR_TRY(ws_get_zeroed(h, WS_ROWT, (size_t)h->rows * 8, (void **)&b.rowt, s));
    return DCSR_OK;
}

// Tile tables scanned by a last workgroup (table) only for batches of more
// than HEADS_DIRECT_TILES tiles; smaller ones sum them where they are used.
static inline bool use_tables(int64_t n) { return ptiles(n) > HEADS_DIRECT_TILES; }

static void launch_offsets(dcsr_handle h, UpdBufs &b, int64_t n, hipStream_t s)
{
    k_upd_offsets<<<ptiles(n), TB, 0, s>>>(h->ds, b.ucode, b.tmode, b.tnewl, b.tmoved, b.tt, b.tp, b.upre, b.tpre,
                                           b.pplan, (unsigned long long)h->pool_cap, use_tables(n), b.mspan);
}

// Plan and decision.  With tile tables the plan's last workgroup decides;
// without them the offsets pass does (its last tile), so it runs here.
// fresh: straight after the heads pass (false for a re-plan after the pool
// was compacted or grown).
dcsr_status plan_and_check(dcsr_handle h, UpdBufs &b, int64_t n, hipStream_t s, bool fresh)
{
    const bool table = use_tables(n);
    k_upd_plan<<<ptiles(n), TB, 0, s>>>(h->ds, (unsigned long long)h->pool_cap, b.trow, b.tins, b.tdel, b.ucode, h->row_len,
                                        h->row_cap, h->row_start, b.trs, b.tnewl, b.tmode, b.tmoved, b.tt, b.tp,
                                        b.pplan, b.fin, table, fresh, b.tlen, b.tcap);
    if (!table) launch_offsets(h, b, n, s);
    return hipGetLastError() == hipSuccess ? DCSR_OK : DCSR_ERR_HIP;
}

// Probe launch: one wave per PR_SPAN positions of the probe space (at most
// the live entries plus one per touched row), the grid capped at what the
// device holds at once (waves stride over the spans).
constexpr int64_t PR_BLOCKS = 1024;
static inline unsigned probe_blocks(dcsr_handle h, int64_t n)
{
    const int64_t spans = (live_bound(h) + n + PR_SPAN - 1) / PR_SPAN;
    const int64_t b = (spans + 3) / 4;
    return (unsigned)(b < 1 ? 1 : (b > PR_BLOCKS ? PR_BLOCKS : b));
}

// Sort, heads, probe, plan (+ decision), offsets of a packed batch: the
// launches shared by every update path (the matrix is only read).  They
// read and write workspace only, never the caller's triples.
dcsr_status sort_plan_launch(dcsr_handle h, int64_t n, UpdBufs &b, hipStream_t s)
{
    const int nbits = h->rbits + h->cbits + 1;
    int which = 0;
    if (b.lb > 0)
        DCSR_HIP(group_sort_pairs(b.k0, b.v0, b.k1, b.v1, n, b.lb, b.ndig, b.hist, s));
    else
        DCSR_HIP(radix_sort_pairs(b.k0, b.v0, b.k1, b.v1, n, nbits, b.hist, b.part, &which, s, true));
    const uint64_t *key = which ? b.k1 : b.k0;
    const uint32_t *val = which ? b.v1 : b.v0;
    const unsigned ht = (unsigned)((n + TILE - 1) / TILE);
    const bool table = ht > HEADS_DIRECT_TILES;
    if (b.rowt && ++h->upd_tag == 0u) {
        // tags wrapped: clear the map, restart at 1
        k_fill_u64<<<grid_for(h->rows, TB), TB, 0, s>>>(b.rowt, h->rows, 0ull);
        h->upd_tag = 1u;
    }
    k_upd_heads_count<<<ht, TB, 0, s>>>(key, n, h->cbits, h->row_len, b.tt, b.tp, b.fin, h->ds, table);
    k_upd_heads_emit<<<ht, TB, 0, s>>>(key, val, n, h->cbits, b.tt, h->ds, table, b.ucol, b.usum, b.ut, b.trow,
                                       b.tubeg, h->row_len, h->row_start, h->row_cap, b.tp, b.tlpre, b.trs, b.tlen,
                                       b.tcap, b.tins, b.tdel, b.pspan, b.rowt, h->upd_tag);
    k_upd_probe<<<probe_blocks(h, n), TB, 0, s>>>(h->ds, h->row_start, h->row_len, h->pool_col, h->pool_val,
                                                  b.tubeg, b.tlpre, b.trs, b.tlen, b.ucol, b.usum, b.uval, b.upos,
                                                  b.ucode, b.tins, b.tdel, b.pspan);
    DCSR_HIP(hipGetLastError());
    DCSR_TRY(plan_and_check(h, b, n, s, true));
    if (use_tables(n)) launch_offsets(h, b, n, s);
    DCSR_HIP(hipGetLastError());
    return DCSR_OK;
}

// Pack (validate + key) the caller's triples into (k0, v0), counting the
// first digit of the batch's sort on the way.
void launch_pack(dcsr_handle h, const int32_t *rows, const int32_t *cols, const float *vals, int64_t n,
                 const UpdBufs &b, hipStream_t s)
{
    k_upd_pack<<<pack_tiles(n), TB, 0, s>>>(rows, cols, vals, n, h->rows, h->cols, h->cbits, b.k0, b.v0, b.hist,
                                            pack_tiles(n), h->ds, b.lb, b.ndig);
}

// Pack (validate + key) the caller's triples, then sort .. offsets.
dcsr_status prepare_launch(dcsr_handle h, const int32_t *rows, const int32_t *cols, const float *vals, int64_t n,
                           UpdBufs &b, hipStream_t s)
{
    launch_pack(h, rows, cols, vals, n, b, s);
    DCSR_HIP(hipGetLastError());
    return sort_plan_launch(h, n, b, s);
}

// Move + write launches.  move_bound bounds the old entries to move: the
// move kernel strides over them with at most MV_BLOCKS workgroups (waves
// past the device's count exit at once).
constexpr int64_t MV_BLOCKS = 1280;
dcsr_status commit_launch(dcsr_handle h, const UpdBufs &b, int64_t n, int64_t move_bound, bool report,
                          hipStream_t s, uint64_t seq = 0)
{
    const int64_t waves = (move_bound + MV_SPAN - 1) / MV_SPAN;
    int64_t blocks = (waves + 3) / 4;
    if (blocks > MV_BLOCKS) blocks = MV_BLOCKS;
    if (blocks < 1) blocks = 1;
    // the SpMV's row scan rides along in the move launch (its own first
    // workgroups) when the handle's workspace has the tagged row map
    RowScanArgs rsa{};
    const bool scan = b.rowt != nullptr && h->rows > 0;
    if (scan) {
        FlatBufs fb;
        DCSR_TRY(flat_buffers(h, s, &fb));
        rsa = RowScanArgs{h->row_len, b.rowt, b.tnewl, fb.foff, fb.rowmap, fb.part, fault_word(h),
                          (uint32_t)h->rows, h->upd_tag, (uint32_t)rowscan_tiles(h->rows), lb_spin_limit()};
    }
    k_upd_move<<<(unsigned)(blocks + rsa.ntiles), TB, 0, s>>>(h->ds, b.tpre, b.trs, b.tubeg, b.ucode, b.upos, b.upre,
                                                              h->pool_col, h->pool_val, b.mspan, rsa);
    k_upd_write<<<grid_for(n, TB * WR_PER), TB, 0, s>>>(h->ds, report ? 1u : 0u, n, b.ut, b.trow, b.tubeg, b.tpre, b.tnewl, b.tmode, b.ucol, b.uval,
                                                b.upos, b.ucode, b.upre, h->pool_col, h->pool_val, h->row_start,
                                                h->row_len, h->row_cap, seq ? h->dsnap : nullptr, seq);
    DCSR_HIP(hipGetLastError());
    if (scan) flat_rebuilt(h);
    else h->flat_dirty = true;
    return DCSR_OK;
}

// Sort through commit of a packed batch of n triples (the commit kernels
// skip themselves when the plan rejected the batch; `report`: see k_upd_write).
// seq: a NOSYNC update's snapshot sequence number (0: no snapshot).
dcsr_status pipeline_launch(dcsr_handle h, int64_t n, UpdBufs &b, bool report, hipStream_t s, uint64_t seq = 0)
{
    DCSR_TRY(sort_plan_launch(h, n, b, s));
    return commit_launch(h, b, n, h->pool_cap, report, s, seq);
}


}  // namespace

// Phase 1 (reads the matrix, writes only workspace): validation, sort,
// merge plan and pool check.  *go (host) receives whether the batch may
// commit; on pool exhaustion *need receives the claimed slot count.
dcsr_status update_prepare(dcsr_handle h, const int32_t *rows, const int32_t *cols, const float *vals, int64_t n,
                           hipStream_t s, bool *go, int64_t *need)
{
    *go = false;
    *need = 0;
    if (!h->canonical) {
        // validate first: a rejected batch must not canonicalize the matrix
        if (n > 0) {
            uint64_t *k0;
            uint32_t *v0;
            DCSR_TRY(ws_get(h, WS_KEY_A, (size_t)n * 8, (void **)&k0));
            DCSR_TRY(ws_get(h, WS_VAL_A, (size_t)n * 4, (void **)&v0));
            k_upd_validate_begin<<<1, 1, 0, s>>>(h->ds);
            k_upd_pack<<<pack_tiles(n), TB, 0, s>>>(rows, cols, vals, n, h->rows, h->cols, h->cbits, k0, v0, nullptr,
                                                    pack_tiles(n), h->ds, 0, 0);
            DCSR_HIP(hipGetLastError());
            DCSR_TRY(pull_state(h, s));
            if (h->hs->ubad) return DCSR_ERR_BAD_INDEX;
        }
        if (!h->state_known) DCSR_TRY(pull_state(h, s));
        DCSR_TRY(canonicalize(h, s));
    }
    if (n == 0) {
        h->last_ins = h->last_add = h->last_del = h->last_touched = 0;
        return DCSR_OK;
    }
    UpdBufs b;
    DCSR_TRY(upd_bufs(h, n, b, s));
    DCSR_TRY(prepare_launch(h, rows, cols, vals, n, b, s));
    DCSR_TRY(pull_state(h, s));
    if (h->hs->ubad_last) return DCSR_ERR_BAD_INDEX;
    *go = h->hs->go != 0;
    *need = (int64_t)h->hs->need;
    return DCSR_OK;
}

// After pool growth/compaction between prepare and commit: plan again (row
// contents are unchanged, only their locations moved; the check takes the
// claims from the current bump pointer).
dcsr_status update_replan(dcsr_handle h, int64_t n, hipStream_t s, bool *go)
{
    UpdBufs b;
    DCSR_TRY(upd_bufs(h, n, b, s));
    DCSR_TRY(plan_and_check(h, b, n, s, false));
    DCSR_TRY(pull_state(h, s));
    *go = h->hs->go != 0;
    return DCSR_OK;
}

namespace {
// Per-call counters of a committed batch (host state read after prepare or
// after commit: both carry them).
void note_batch(dcsr_handle h, const DevState &st)
{
    h->last_ins = st.n_ins;
    h->last_add = st.n_add;
    h->last_del = st.n_del;
    h->last_touched = st.ntouched;
    if (st.hub_touch) h->hubs_dirty = true;
}

// The claims of a batch of n triples are at most: every touched row (at
// most min(n, rows)) rewritten with all of its live entries, every triple
// inserted, each chunk rounded up.
int64_t claims_bound(dcsr_handle h, int64_t n)
{
    const int64_t touched = n < h->rows ? n : h->rows;
    return live_bound(h) + n + (int64_t)(DCSR_CHUNK_ALIGN - 1u) * touched;
}

// When the claim bound fits the free pool (from the host's bound of the
// bump pointer), the device's pool check cannot fail.
bool claims_fit(dcsr_handle h, int64_t n) { return bump_bound(h) + claims_bound(h, n) <= h->pool_cap; }
}  // namespace

// Phase 2 (writes the matrix): moves, inserts, in-place values, row-table
// commit.  Launches only; the host mirrors are set from the prepare check.
dcsr_status update_commit(dcsr_handle h, int64_t n, hipStream_t s, bool nosync)
{
    UpdBufs b;
    DCSR_TRY(upd_bufs(h, n, b, s));
    DCSR_TRY(commit_launch(h, b, n, nosync ? h->pool_cap : (int64_t)h->hs->nrewrite_old, nosync, s));
    if (nosync) {
        state_unknown(h, n, claims_bound(h, n));
        return DCSR_OK;
    }
    const DevState &st = *h->hs;
    h->nnz_h = st.nnz + st.d_nnz;
    h->bump_h = (int64_t)st.bump;
    h->garbage_h = st.garbage + st.d_garbage;
    note_batch(h, st);
    state_exact(h);
    return DCSR_OK;
}

// NOSYNC: no state read-back.  The host keeps upper bounds of the live
// entries and the bump pointer (tightened from asynchronous state snapshots
// once they land).  When they cannot rule out pool exhaustion and growth is
// allowed, the pool is first compacted in stream order (no read-back) or,
// when even a compacted pool might not hold the claims, the state is read
// and the pool grown: the batch is not lost to exhaustion.
// Synchronous update: the batch's status is known on return.
static dcsr_status update_sync(dcsr_handle h, const int32_t *rows, const int32_t *cols, const float *vals, int64_t n,
                               uint32_t flags, hipStream_t s)
{
    bool go = false;
    int64_t need = 0;
    if (n > 0 && h->canonical && claims_fit(h, n)) {
        // prepare and commit back to back, one state read at the end (it
        // then holds the committed counters: write folds the deltas in)
        UpdBufs b;
        DCSR_TRY(upd_bufs(h, n, b, s));
        launch_pack(h, rows, cols, vals, n, b, s);
        DCSR_HIP(hipGetLastError());
        DCSR_TRY(pipeline_launch(h, n, b, false, s));
        DCSR_TRY(pull_state(h, s));
        if (h->hs->ubad_last) return DCSR_ERR_BAD_INDEX;
        if (h->hs->go) {
            note_batch(h, *h->hs);
            return DCSR_OK;
        }
        // not committed (the bound rules this out): grow and re-plan below
        need = (int64_t)h->hs->need;
    } else {
        DCSR_TRY(update_prepare(h, rows, cols, vals, n, s, &go, &need));
        if (n == 0) return DCSR_OK;
    }
    if (!go) {
        if (!h->cfg.allow_grow || (flags & DCSR_UPDATE_NOGROW)) return DCSR_ERR_POOL_EXHAUSTED;
        DCSR_TRY(pool_reserve(h, need + need / 8 + 64, s, nullptr));
        DCSR_TRY(update_replan(h, n, s, &go));
        if (!go) return DCSR_ERR_POOL_EXHAUSTED;
    }
    return update_commit(h, n, s, false);
}

static dcsr_status update_nosync(dcsr_handle h, const int32_t *rows, const int32_t *cols, const float *vals, int64_t n,
                                 uint32_t flags, hipStream_t s)
{
    // the first update of a matrix loaded with duplicates canonicalizes it:
    // done synchronously, once
    if (!h->canonical) return update_sync(h, rows, cols, vals, n, flags, s);
    if (n == 0) return DCSR_OK;
    refresh_bounds(h);
    if (!claims_fit(h, n) && h->cfg.allow_grow && !(flags & DCSR_UPDATE_NOGROW)) {
        // compaction (or growth, when even a compacted pool might not hold
        // the claims) in stream order, from the host's bounds alone: no
        // read-back, so the host never waits for the GPU here
        const int64_t need = claims_bound(h, n);
        const int64_t live = live_bound(h);
        if (live + need <= h->pool_cap) {
            DCSR_TRY(compact_in_order(h, s));
        } else {
            const int64_t want = live + 2 * need;
            DCSR_TRY(grow_in_order(h, want > 2 * h->pool_cap ? want : 2 * h->pool_cap, s));
        }
    }
    const int64_t claimed = claims_bound(h, n);
    UpdBufs b;
    DCSR_TRY(upd_bufs(h, n, b, s));
    launch_pack(h, rows, cols, vals, n, b, s);
    DCSR_HIP(hipGetLastError());
    const uint64_t seq = snapshot_next(h);
    DCSR_TRY(pipeline_launch(h, n, b, true, s, seq));
    state_unknown(h, n, claimed);
    snapshot_note(h, seq);
    return DCSR_OK;
}

namespace {
__global__ void k_zero_u32(uint32_t *p, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}
}  // namespace

// Bytes of scratch an update of n triples on h carves from a fixed region.
size_t update_scratch_bytes(dcsr_handle h, int64_t n)
{
    BufSrc src{h, nullptr, 0, 0, true, nullptr};
    UpdBufs b;
    if (upd_bufs_src(src, n, b) != DCSR_OK) return ~size_t(0);
    return src.used;
}

// An update whose scratch is carved from the fixed device region [base,
// base + cap) instead of the handle's workspace: nothing is allocated, the
// pool is never grown or compacted (a batch whose claims do not fit is
// rejected whole: DCSR_ERR_POOL_EXHAUSTED, or with NOSYNC the sticky bit),
// and the matrix must be canonical (canonicalizing allocates).  The region's
// finish counters are cleared first (a fill kernel: safe under capture).
dcsr_status update_fixed_scratch(dcsr_handle h, const int32_t *rows, const int32_t *cols, const float *vals,
                                 int64_t n, uint32_t flags, hipStream_t s, uint8_t *base, size_t cap)
{
    if (!h->canonical) return DCSR_ERR_INVALID;
    h->last_stream = s;
    if (n == 0) return DCSR_OK;
    BufSrc src{h, base, cap, 0, false, s};
    UpdBufs b;
    DCSR_TRY(upd_bufs_src(src, n, b));
    k_zero_u32<<<1, 64, 0, s>>>(b.fin, 16);
    const bool nosync = (flags & DCSR_UPDATE_NOSYNC) != 0;
    if (nosync) refresh_bounds(h);
    launch_pack(h, rows, cols, vals, n, b, s);
    DCSR_HIP(hipGetLastError());
    const uint64_t seq = nosync ? snapshot_next(h) : 0;
    const int64_t claimed = claims_bound(h, n);
    DCSR_TRY(pipeline_launch(h, n, b, nosync, s, seq));
    if (nosync) {
        state_unknown(h, n, claimed);
        snapshot_note(h, seq);
        return DCSR_OK;
    }
    DCSR_TRY(pull_state(h, s));
    if (h->hs->ubad_last) return DCSR_ERR_BAD_INDEX;
    if (!h->hs->go) return DCSR_ERR_POOL_EXHAUSTED;
    note_batch(h, *h->hs);
    return DCSR_OK;
}

dcsr_status dcsr_update(dcsr_handle h, const int32_t *rows, const int32_t *cols, const float *vals, int64_t n,
                        uint32_t flags, dcsr_stream_t stream)
{
    if (!h || n < 0 || (n > 0 && (!rows || !cols || !vals))) return DCSR_ERR_INVALID;
    if (n >= (int64_t)0x7ffffff0ll) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    h->last_stream = s;
    if (flags & DCSR_UPDATE_NOSYNC) return update_nosync(h, rows, cols, vals, n, flags, s);
    return update_sync(h, rows, cols, vals, n, flags, s);
}

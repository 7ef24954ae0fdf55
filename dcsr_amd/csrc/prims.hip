// prims.hip -- device-wide scan and stable LSD radix sort for gfx950.
//
// Replace the thrust/cub primitives the reference calls on its hot path
// (thrust::exclusive_scan include/load.h:418, thrust::sort_by_key
// include/load.h:602-603, cusp::indices_to_offsets, cub::BlockRadixSort
// include/spmm.inl:489-490) with wave64 code: DPP scans and ballot-based
// digit ranking (8 ballots per item give the lanes sharing a digit).
// Scans are single-pass (one launch; cross-tile prefixes by decoupled
// look-back, lookback.h).  A radix pass is three launches: per-tile digit
// counts, a scan of the count table, stable scatter.  The table scan runs
// one workgroup per digit over that digit's tile counts (all loads of a
// digit issued at once); the scatter adds the exclusive prefix of the 256
// digit totals, which it scans itself.  (A onesweep pass that resolves each
// digit's offset by look-back walks hundreds of tiles serially per digit
// here, where every look-back poll is a cross-XCD coherent load; the table
// scan is faster.)
#include "lookback.h"
#include "prims.h"
#include "wave.h"

namespace {

__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_IPT = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_IPT;

// Single-pass exclusive scan of NC interleaved components: tile = ticket,
// local blocked scan, decoupled look-back per component, write.
// part: [ctl 4 words][status: tiles x NC words of 64 bits].
template <int NC>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan1(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                        int64_t n, uint32_t *__restrict__ part,
                                                        uint32_t *__restrict__ total, unsigned int *fault, int limit)
{
    __shared__ uint32_t tile_s[NC][SCAN_TILE];
    __shared__ uint32_t lds4[4];
    __shared__ uint32_t s_ticket[2];
    __shared__ uint32_t s_agg[NC];
    __shared__ uint32_t s_pre[NC];
    unsigned int *ctl = part;
    uint64_t *st = (uint64_t *)(part + 4);
    uint32_t tag;
    const int64_t tile = lb_ticket(ctl, s_ticket, &tag);
    const int64_t ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    const int64_t base = tile * SCAN_TILE;
    // all loads in flight first (a tail item reads the tile's first element
    // and counts as 0)
    uint32_t ld[NC][SCAN_IPT];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int i = 0; i < SCAN_IPT; ++i) {
            const int64_t idx = base + i * SCAN_THREADS + threadIdx.x;
            ld[c][i] = in[(idx < n ? idx : base) * NC + c];
        }
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int i = 0; i < SCAN_IPT; ++i) {
            const int64_t idx = base + i * SCAN_THREADS + threadIdx.x;
            tile_s[c][i * SCAN_THREADS + threadIdx.x] = idx < n ? ld[c][i] : 0u;
        }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        uint32_t loc[SCAN_IPT];
        uint32_t sum = 0;
#pragma unroll
        for (int i = 0; i < SCAN_IPT; ++i) {
            loc[i] = sum;
            sum += tile_s[c][threadIdx.x * SCAN_IPT + i];
        }
        uint32_t tot;
        const uint32_t pre = block_excl_scan_256(sum, lds4, &tot);
#pragma unroll
        for (int i = 0; i < SCAN_IPT; ++i) tile_s[c][threadIdx.x * SCAN_IPT + i] = pre + loc[i];
        if (threadIdx.x == 0) s_agg[c] = tot;
    }
    __syncthreads();
    // fallback: the inputs of every earlier tile (never overwritten: the
    // scan is out of place)
    lb_tile_prefix<NC>(s_agg, s_pre, st, tile, tag, ctl, fault, limit,
                       [&](int c) { return lb_wave_sum_prefix(in + c, base, NC); });
    if (threadIdx.x < NC && tile == ntiles - 1 && total) total[threadIdx.x] = s_pre[threadIdx.x] + s_agg[threadIdx.x];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const uint32_t p = s_pre[c];
#pragma unroll
        for (int i = 0; i < SCAN_IPT; ++i) {
            const int64_t idx = base + i * SCAN_THREADS + threadIdx.x;
            if (idx < n) out[idx * NC + c] = p + tile_s[c][i * SCAN_THREADS + threadIdx.x];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) lb_finish(ctl);
}

// ---------------------------------------------------------------------------
// radix sort
// ---------------------------------------------------------------------------
constexpr int RS_THREADS = 256;
constexpr int RS_WAVES = RS_THREADS / 64;
constexpr int RS_IPT = 8;                        // items per lane
constexpr int RS_WAVE_ITEMS = RS_IPT * 64;       // contiguous items per wave
constexpr int RS_TILE = RS_WAVES * RS_WAVE_ITEMS; // 2048
static_assert(RS_TILE == RADIX_TILE, "count-table tiles are shared with callers");

__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint64_t *__restrict__ keys, int64_t n, int shift,
                                                        uint32_t *__restrict__ hist, int64_t nblocks)
{
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    // all loads in flight before the first count (a tail item reads the
    // tile's first key and is not counted)
    uint64_t k[RS_TILE / RS_THREADS];
#pragma unroll
    for (int j = 0; j < RS_TILE / RS_THREADS; ++j) {
        const int64_t idx = base + j * RS_THREADS + threadIdx.x;
        k[j] = keys[idx < n ? idx : base];
    }
#pragma unroll
    for (int j = 0; j < RS_TILE / RS_THREADS; ++j)
        if (base + j * RS_THREADS + threadIdx.x < n) atomicAdd(&cnt[(uint32_t)(k[j] >> shift) & 255u], 1u);
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive scan of each digit's row of the count table (digit-major:
// hist[d * nb + tile]), in place, one workgroup per digit; dtot[d] gets the
// digit's total.  The row is taken RS_SCAN_CHUNKS * 2048 tiles at a time,
// every load of such a group in flight before its first scan step (one
// group covers batches of up to 16M keys).
constexpr int RS_SCAN_CHUNKS = 4;
__global__ __launch_bounds__(RS_THREADS) void k_rs_scan_digits(uint32_t *__restrict__ hist, int64_t nb,
                                                               uint32_t *__restrict__ dtot)
{
    __shared__ uint32_t lds4[4];
    uint32_t *row = hist + (int64_t)blockIdx.x * nb;
    uint32_t carry = 0;
    for (int64_t g0 = 0; g0 < nb; g0 += (int64_t)RS_SCAN_CHUNKS * RS_TILE) {
        uint32_t v[RS_SCAN_CHUNKS][RS_IPT];
#pragma unroll
        for (int c = 0; c < RS_SCAN_CHUNKS; ++c) {
            const int64_t i0 = g0 + (int64_t)c * RS_TILE + (int64_t)threadIdx.x * RS_IPT;
#pragma unroll
            for (int j = 0; j < RS_IPT; ++j) v[c][j] = i0 + j < nb ? row[i0 + j] : 0u;
        }
#pragma unroll
        for (int c = 0; c < RS_SCAN_CHUNKS; ++c) {
            if (g0 + (int64_t)c * RS_TILE >= nb) break;
            const int64_t i0 = g0 + (int64_t)c * RS_TILE + (int64_t)threadIdx.x * RS_IPT;
            uint32_t s = 0;
#pragma unroll
            for (int j = 0; j < RS_IPT; ++j) s += v[c][j];
            uint32_t tot;
            uint32_t p = carry + block_excl_scan_256(s, lds4, &tot);
#pragma unroll
            for (int j = 0; j < RS_IPT; ++j) {
                if (i0 + j < nb) row[i0 + j] = p;
                p += v[c][j];
            }
            carry += tot;
        }
    }
    if (threadIdx.x == 0) dtot[blockIdx.x] = carry;
}

// Stable scatter of one tile.  offs[d * nb + tile] is the tile's offset
// within digit d; the scatter turns the digit totals dtot into digit bases.
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint64_t *__restrict__ kin,
                                                           const uint32_t *__restrict__ vin,
                                                           uint64_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                           int64_t n, int shift,
                                                           const uint32_t *__restrict__ offs, int64_t nblocks,
                                                           const uint32_t *__restrict__ dtot)
{
    __shared__ uint64_t skey[RS_TILE];
    __shared__ uint32_t sval[RS_TILE];
    __shared__ uint32_t wcnt[RS_WAVES][256];
    __shared__ uint32_t loff[256];
    __shared__ uint32_t goff[256];
    __shared__ uint32_t lds4[4];

    const int w = threadIdx.x >> 6;
    const int lane = lane_id();
    // the tile's offsets do not depend on the ranking: load them first
    const uint32_t go = offs[(int64_t)threadIdx.x * nblocks + blockIdx.x];
    const uint32_t dt = dtot[threadIdx.x];
    for (int i = threadIdx.x; i < RS_WAVES * 256; i += RS_THREADS) (&wcnt[0][0])[i] = 0;
    __syncthreads();

    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    const int64_t wbase = base + (int64_t)w * RS_WAVE_ITEMS;
    uint64_t key[RS_IPT];
    uint32_t val[RS_IPT];
    uint32_t rnk[RS_IPT];
    const uint64_t lt = lanemask_lt();
    // every load issued before the ranking starts: out-of-range items read
    // the tile's first pair, then become ~0 keys (last digit, ranked last)
#pragma unroll
    for (int j = 0; j < RS_IPT; ++j) {
        const int64_t idx = wbase + j * 64 + lane;
        key[j] = kin[idx < n ? idx : base];
        val[j] = vin[idx < n ? idx : base];
    }
#pragma unroll
    for (int j = 0; j < RS_IPT; ++j) {
        if (wbase + j * 64 + lane >= n) key[j] = ~0ull;
        uint32_t d = (uint32_t)(key[j] >> shift) & 255u;
        uint64_t peers = wave_match_digit8(d, ~0ull);
        int leader = __ffsll((long long)peers) - 1;
        // every lane of the digit reads its count before the leader adds to
        // it (a wave's LDS accesses complete in order): no broadcast needed
        const uint32_t b = wcnt[w][d];
        if (lane == leader) wcnt[w][d] = b + (uint32_t)__popcll(peers);
        rnk[j] = b + (uint32_t)__popcll(peers & lt);
    }
    __syncthreads();
    {
        const int d = threadIdx.x;  // one digit per thread
        uint32_t c[RS_WAVES];
        uint32_t s = 0;
#pragma unroll
        for (int i = 0; i < RS_WAVES; ++i) {
            c[i] = wcnt[i][d];
            wcnt[i][d] = s;
            s += c[i];
        }
        uint32_t tot;
        loff[d] = block_excl_scan_256(s, lds4, &tot);
        goff[d] = go + block_excl_scan_256(dt, lds4, &tot);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RS_IPT; ++j) {
        uint32_t d = (uint32_t)(key[j] >> shift) & 255u;
        uint32_t p = loff[d] + wcnt[w][d] + rnk[j];
        skey[p] = key[j];
        sval[p] = val[j];
    }
    __syncthreads();
    const int64_t valid = (n - base) < RS_TILE ? (n - base) : RS_TILE;
    for (int p = threadIdx.x; p < valid; p += RS_THREADS) {
        uint64_t k = skey[p];
        uint32_t d = (uint32_t)(k >> shift) & 255u;
        int64_t o = (int64_t)goff[d] + (p - (int64_t)loff[d]);
        kout[o] = k;
        vout[o] = sval[p];
    }
}

}  // namespace

size_t scan_part_words(int64_t n, int nc)
{
    const int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    return (size_t)(4 + 2 * (tiles < 1 ? 1 : tiles) * nc);
}

hipError_t scan_exclusive_u32(const uint32_t *in, uint32_t *out, int64_t n, int nc, uint32_t *part,
                              uint32_t *total, hipStream_t s, unsigned int *fault)
{
    if (in == out) return hipErrorInvalidValue;  // the look-back fallback re-reads inputs
    if (n <= 0) {
        if (total) k_fill_u32<<<1, 64, 0, s>>>(total, nc, 0u);
        return hipGetLastError();
    }
    const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nc == 1)
        k_scan1<1><<<(unsigned)nb, SCAN_THREADS, 0, s>>>(in, out, n, part, total, fault, lb_spin_limit());
    else if (nc == 2)
        k_scan1<2><<<(unsigned)nb, SCAN_THREADS, 0, s>>>(in, out, n, part, total, fault, lb_spin_limit());
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// Count table (256 x tiles) plus the 256 digit totals.
size_t radix_hist_words(int64_t n)
{
    const int64_t nb = (n + RS_TILE - 1) / RS_TILE;
    return (size_t)(nb < 1 ? 1 : nb) * 256 + 256;
}

size_t radix_part_words(int64_t) { return 4; }

// Per pass: per-tile digit counts (digit-major, tile-minor), a per-digit
// scan of that table, stable scatter.
hipError_t radix_sort_pairs(uint64_t *k0, uint32_t *v0, uint64_t *k1, uint32_t *v1, int64_t n, int nbits,
                            uint32_t *hist, uint32_t *, int *which, hipStream_t s, bool counted)
{
    *which = 0;
    if (n <= 1) return hipSuccess;
    const int64_t nb = (n + RS_TILE - 1) / RS_TILE;
    uint32_t *dtot = hist + nb * 256;
    uint64_t *ki = k0, *ko = k1;
    uint32_t *vi = v0, *vo = v1;
    int cur = 0;
    for (int shift = 0; shift < nbits; shift += 8) {
        if (!(counted && shift == 0)) k_rs_hist<<<(unsigned)nb, RS_THREADS, 0, s>>>(ki, n, shift, hist, nb);
        k_rs_scan_digits<<<256, RS_THREADS, 0, s>>>(hist, nb, dtot);
        k_rs_scatter<<<(unsigned)nb, RS_THREADS, 0, s>>>(ki, vi, ko, vo, n, shift, hist, nb, dtot);
        uint64_t *tk = ki; ki = ko; ko = tk;
        uint32_t *tv = vi; vi = vo; vo = tv;
        cur ^= 1;
    }
    *which = cur;
    return hipGetLastError();
}

static int g_lb_limit = LB_SPIN_LIMIT;
int lb_spin_limit() { return g_lb_limit; }

extern "C" void dcsr_set_scan_wait_limit(int32_t polls) { g_lb_limit = polls < 0 ? LB_SPIN_LIMIT : polls; }

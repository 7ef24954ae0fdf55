// gsort.hip -- row-group sort of an update batch's (key, value) pairs.
//
// Keys are (row << cbits) | col; invalid ones are ~0.  The rows are split
// into G groups of 2^gshift consecutive rows, and a key's digit is its group
// (key >> lb with lb = gshift + cbits; ~0 goes to digit G, after every
// group).  The caller's pack pass counts the digits per 2048-pair tile into
// a blocked table; then
//
//   1. k_gs_scan: each digit's tile counts become exclusive offsets (one
//      workgroup per 64 digits, one lane per digit, every load in flight
//      before the first add);
//   2. k_gs_scatter: one stable scatter places each group's pairs
//      contiguously, in arrival order (digits ranked by wave ballots, staged
//      in LDS so runs are written together);
//   3. k_gs_sort: one workgroup per group sorts the group's pairs by their
//      low lb key bits (all of its keys share the bits above), a stable LSD
//      sort with 8- or 9-bit digits (whichever takes fewer passes); groups of
//      up to GS_CAP pairs sort in LDS (one 32-bit word per pair where the
//      local key's unranked bits and the pair's position fit one, else a
//      key and a position array), larger ones out of global memory with the
//      same ranking.
//
// The result is exactly the stable sort of the batch by key, which the LSD
// radix sort (prims.hip) gives in one count/scan/scatter round per 8 key
// bits: here the pairs cross HBM in one scatter and one group pass.  Stable
// ranks keep equal keys in arrival order, which the insert-or-add sums rely
// on.
#include "prims.h"
#include "wave.h"

namespace {

constexpr int GT = 256;              // threads of the scan-free kernels
constexpr int GIPT = 8;              // pairs per lane in a scatter tile
constexpr int GTILE = GT * GIPT;     // 2048 == RADIX_TILE
static_assert(GTILE == RADIX_TILE, "pack counts tiles of the scatter's size");
constexpr int GS_SCAN_T = 1024;      // k_gs_scan: 16 waves per 64 digits
constexpr int GS_SCAN_Q = 32;        // tiles per wave per round

// Lanes holding the same digit of `bits` bits (every lane active).
__device__ __forceinline__ uint64_t wave_match_bits(uint32_t d, int bits)
{
    uint64_t peers = ~0ull;
    for (int b = 0; b < bits; ++b) {
        const uint64_t m = __ballot((d >> b) & 1u);
        peers &= ((d >> b) & 1u) ? m : ~m;
    }
    return peers;
}

__device__ __forceinline__ uint32_t group_digit(uint64_t k, int lb, int ndig)
{
    return k == ~0ull ? (uint32_t)(ndig - 1) : (uint32_t)(k >> lb);
}

// Exclusive offsets of each digit's tile counts, in place; dtot[d] receives
// the digit's total.  Block b covers digits [64b, 64b + 64): lane = digit.
// A round covers NW * GS_SCAN_Q consecutive tiles, GS_SCAN_Q per wave in
// wave order; the wave's counts stay in registers between the sum and the
// write of the prefixes.
__global__ __launch_bounds__(GS_SCAN_T) void k_gs_scan(uint32_t *__restrict__ hist, int64_t nb,
                                                       uint32_t *__restrict__ dtot)
{
    constexpr int NW = GS_SCAN_T / 64;
    __shared__ uint32_t wsum[NW][64];
    const int w = threadIdx.x >> 6, lane = lane_id();
    uint32_t *col = hist + (int64_t)blockIdx.x * nb * 64 + lane;
    uint32_t carry_in = 0;  // the digit's count in earlier rounds
    for (int64_t r0 = 0; r0 < nb; r0 += (int64_t)NW * GS_SCAN_Q) {
        const int64_t t0 = r0 + (int64_t)w * GS_SCAN_Q;
        const int64_t t1 = t0 + GS_SCAN_Q < nb ? t0 + GS_SCAN_Q : nb;
        uint32_t v[GS_SCAN_Q];
        uint32_t s = 0;
#pragma unroll
        for (int q = 0; q < GS_SCAN_Q; ++q) {
            const int64_t t = t0 + q;
            v[q] = t < t1 ? col[t * 64] : 0u;
        }
#pragma unroll
        for (int q = 0; q < GS_SCAN_Q; ++q) s += v[q];
        wsum[w][lane] = s;
        __syncthreads();
        uint32_t before = 0, tot = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            const uint32_t x = wsum[i][lane];
            before += i < w ? x : 0u;
            tot += x;
        }
        __syncthreads();
        uint32_t p = carry_in + before;
#pragma unroll
        for (int q = 0; q < GS_SCAN_Q; ++q) {
            const int64_t t = t0 + q;
            if (t < t1) col[t * 64] = p;
            p += v[q];
        }
        carry_in += tot;
    }
    if (w == 0) dtot[(int64_t)blockIdx.x * 64 + lane] = carry_in;
}

// Stable scatter of one 2048-pair tile into the group order.  gbase_out
// (block 0) receives the exclusive offsets of the ndig digits plus the total.
__global__ __launch_bounds__(GT) void k_gs_scatter(const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                   uint64_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                   int64_t n, int lb, int ndig, int dbits,
                                                   const uint32_t *__restrict__ hist, int64_t nb,
                                                   const uint32_t *__restrict__ dtot, uint32_t *__restrict__ gbase_out)
{
    __shared__ uint32_t wcnt[GT / 64][GS_MAX_DIGITS];
    __shared__ uint32_t lofs[GS_MAX_DIGITS];
    __shared__ uint32_t gofs[GS_MAX_DIGITS];
    __shared__ uint64_t skey[GTILE];
    __shared__ uint32_t sval[GTILE];
    __shared__ uint32_t lds4[4];
    constexpr int DPT = GS_MAX_DIGITS / GT;  // digits per thread in the offset phase
    const int w = threadIdx.x >> 6, lane = lane_id();
    const int64_t tile = blockIdx.x;
    const int64_t base = tile * GTILE;
    const int64_t wbase = base + (int64_t)w * (GIPT * 64);
    // every load first: the tile's offsets and the digit totals of this
    // thread's digits (dpt contiguous digits per thread, spread over every
    // thread), then the pairs
    uint32_t ho[DPT], dt[DPT];
    const int dpt = (ndig + GT - 1) / GT;
    const int d0 = (int)threadIdx.x * dpt;
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
        const int d = d0 + q;
        const bool in = q < dpt && d < ndig;
        ho[q] = in ? hist[gs_hist_index(d, tile, nb)] : 0u;
        dt[q] = in ? dtot[d] : 0u;
    }
    uint64_t key[GIPT];
    uint32_t val[GIPT];
#pragma unroll
    for (int j = 0; j < GIPT; ++j) {
        const int64_t idx = wbase + j * 64 + lane;
        key[j] = kin[idx < n ? idx : base];
        val[j] = vin[idx < n ? idx : base];
    }
    for (int d = threadIdx.x; d < ndig; d += GT)
#pragma unroll
        for (int i = 0; i < GT / 64; ++i) wcnt[i][d] = 0;
    __syncthreads();
    // rank: items past n take the last digit and rank after every real one
    const uint64_t lt = lanemask_lt();
    uint32_t rnk[GIPT], dg[GIPT];
#pragma unroll
    for (int j = 0; j < GIPT; ++j) {
        const bool real = wbase + j * 64 + lane < n;
        const uint32_t d = real ? group_digit(key[j], lb, ndig) : (uint32_t)(ndig - 1);
        const uint64_t peers = wave_match_bits(d, dbits);
        const int leader = __ffsll((long long)peers) - 1;
        // every lane of the digit reads its count before the leader adds to
        // it (a wave's LDS accesses complete in order): no broadcast needed
        const uint32_t b = wcnt[w][d];
        if (lane == leader) wcnt[w][d] = b + (uint32_t)__popcll(peers);
        rnk[j] = b + (uint32_t)__popcll(peers & lt);
        dg[j] = d;
    }
    __syncthreads();
    {
        uint32_t cs[DPT], gs[DPT];
        uint32_t csum = 0, gsum = 0;
#pragma unroll
        for (int q = 0; q < DPT; ++q) {
            const int d = d0 + q;
            uint32_t c = 0;
            if (q < dpt && d < ndig) {
#pragma unroll
                for (int i = 0; i < GT / 64; ++i) {
                    const uint32_t x = wcnt[i][d];
                    wcnt[i][d] = c;
                    c += x;
                }
            }
            cs[q] = csum;
            csum += c;
            gs[q] = gsum;
            gsum += dt[q];
        }
        uint32_t tot;
        const uint32_t lp = block_excl_scan_256(csum, lds4, &tot);
        const uint32_t gp = block_excl_scan_256(gsum, lds4, &tot);
#pragma unroll
        for (int q = 0; q < DPT; ++q) {
            const int d = d0 + q;
            if (q < dpt && d < ndig) {
                lofs[d] = lp + cs[q];
                gofs[d] = gp + gs[q] + ho[q];
                if (tile == 0) gbase_out[d] = gp + gs[q];
            }
        }
        if (tile == 0 && threadIdx.x == 0) gbase_out[ndig] = tot;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < GIPT; ++j) {
        const uint32_t p = lofs[dg[j]] + wcnt[w][dg[j]] + rnk[j];
        skey[p] = key[j];
        sval[p] = val[j];
    }
    __syncthreads();
    const int64_t valid = (n - base) < GTILE ? (n - base) : GTILE;
#pragma unroll
    for (int j = 0; j < GIPT; ++j) {
        const int p = j * GT + (int)threadIdx.x;
        if (p < valid) {
            const uint64_t k = skey[p];
            const uint32_t d = group_digit(k, lb, ndig);
            const int64_t o = (int64_t)gofs[d] + (p - (int64_t)lofs[d]);
            kout[o] = k;
            vout[o] = sval[p];
        }
    }
}

// ---------------------------------------------------------------------------
// Per-group sort.  Ranking of one chunk of up to 4 * 64 * J items in their
// current order (wave w holds positions [w * 64J, (w + 1) * 64J), item j of
// a lane at w * 64J + 64 j + lane): rnk[j] = stable rank of item j among the
// chunk's items of its digit in earlier waves excluded; afterwards
// wcnt[w][d] = items of digit d in waves before w, loff[d] = items of the
// chunk with a smaller digit, cnt_out[d] (if not null) = items of digit d.
// ---------------------------------------------------------------------------
constexpr int SJ = GS_CAP / GT;  // items per lane in a full LDS group

template <int J, int DB = 8, typename C = uint32_t>
__device__ __forceinline__ void rank_chunk(const uint32_t (&dg)[J], const bool (&real)[J], int jmax,
                                           uint32_t (&rnk)[J], C (*wcnt)[1 << DB], uint32_t *loff,
                                           uint32_t *lds4)
{
    constexpr int ND = 1 << DB;     // digits
    constexpr int DPT = ND / GT;    // digits per thread in the offset scan
    static_assert(DPT >= 1 && ND % GT == 0, "at least one digit per thread");
    const int w = threadIdx.x >> 6, lane = lane_id();
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int i = 0; i < ND / 64; ++i) wcnt[w][lane + 64 * i] = 0;
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int j = 0; j < J; ++j) {
        rnk[j] = 0;
        if (j >= jmax) continue;  // uniform over the workgroup
        const uint64_t act = __ballot(real[j]);
        if (!real[j]) continue;
        uint64_t peers = act;
#pragma unroll
        for (int b = 0; b < DB; ++b) {
            const uint64_t m = __ballot((dg[j] >> b) & 1u);
            peers &= ((dg[j] >> b) & 1u) ? m : ~m;
        }
        const int leader = __ffsll((long long)peers) - 1;
        // every lane of the digit reads its count before the leader adds to
        // it (a wave's LDS accesses complete in order): no broadcast needed
        const uint32_t c = wcnt[w][dg[j]];
        if (lane == leader) wcnt[w][dg[j]] = (C)(c + (uint32_t)__popcll(peers));
        rnk[j] = c + (uint32_t)__popcll(peers & lt);
    }
    __syncthreads();
    {
        // digits threadIdx.x * DPT .. + DPT - 1 (contiguous: the scan stays
        // in digit order)
        uint32_t sd[DPT], st = 0;
#pragma unroll
        for (int k = 0; k < DPT; ++k) {
            const int d = threadIdx.x * DPT + k;
            uint32_t s = 0;
#pragma unroll
            for (int i = 0; i < GT / 64; ++i) {
                const uint32_t x = wcnt[i][d];
                wcnt[i][d] = (C)s;
                s += x;
            }
            sd[k] = s;
            st += s;
        }
        uint32_t tot;
        uint32_t pre = block_excl_scan_256(st, lds4, &tot);
#pragma unroll
        for (int k = 0; k < DPT; ++k) {
            loff[threadIdx.x * DPT + k] = pre;
            pre += sd[k];
        }
    }
    __syncthreads();
}

// LDS sort of cnt <= GS_CAP pairs of group g: keys (local bits) and the
// pairs' indices ping-pong between two LDS arrays; digits of DB bits (9 when
// that takes fewer passes over the lb local bits than 8).
template <typename K, int DB>
struct GsLds {
    K key[2][GS_CAP];
    uint16_t idx[2][GS_CAP];
    uint32_t wcnt[GT / 64][1 << DB];
    uint32_t loff[1 << DB];
    uint32_t lds4[4];
};

template <typename K, int DB>
__device__ void group_sort_lds(GsLds<K, DB> &S, const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                               uint64_t *__restrict__ kout, uint32_t *__restrict__ vout, int64_t s0, int cnt, int lb,
                               uint64_t high)
{
    const int w = threadIdx.x >> 6, lane = lane_id();
    const uint64_t lmask = lb >= 64 ? ~0ull : ((1ull << lb) - 1ull);
    // every key load in flight before the first store (indices clamped into
    // the group, extra items dropped)
    {
        uint64_t kl[SJ];
#pragma unroll
        for (int j = 0; j < SJ; ++j) {
            const int i = j * GT + threadIdx.x;
            kl[j] = kin[s0 + (i < cnt ? i : cnt - 1)];
        }
#pragma unroll
        for (int j = 0; j < SJ; ++j) {
            const int i = j * GT + threadIdx.x;
            if (i < cnt) {
                S.key[0][i] = (K)(kl[j] & lmask);
                S.idx[0][i] = (uint16_t)i;
            }
        }
    }
    __syncthreads();
    // positions per wave: a multiple of 64 covering cnt / 4
    const int jmax = (cnt + GT - 1) / GT;
    const int pw = jmax * 64;
    int cur = 0;
    for (int sh = 0; sh < lb; sh += DB) {
        // only the ranks stay in registers; keys are read again from LDS for
        // the scatter
        uint32_t dg[SJ], rnk[SJ];
        bool real[SJ];
#pragma unroll
        for (int j = 0; j < SJ; ++j) {
            const int p = w * pw + j * 64 + lane;
            real[j] = j < jmax && p < cnt;
            dg[j] = real[j] ? (uint32_t)(S.key[cur][p] >> sh) & ((1u << DB) - 1u) : 0u;
        }
        rank_chunk<SJ, DB>(dg, real, jmax, rnk, S.wcnt, S.loff, S.lds4);
#pragma unroll
        for (int j = 0; j < SJ; ++j) {
            if (!real[j]) continue;
            const int p = w * pw + j * 64 + lane;
            const uint32_t q = S.loff[dg[j]] + S.wcnt[w][dg[j]] + rnk[j];
            S.key[cur ^ 1][q] = S.key[cur][p];
            S.idx[cur ^ 1][q] = S.idx[cur][p];
        }
        __syncthreads();
        cur ^= 1;
    }
    // the value gather: every load in flight before the first store
    uint32_t vv[SJ];
#pragma unroll
    for (int j = 0; j < SJ; ++j) {
        const int i = j * GT + threadIdx.x;
        vv[j] = vin[s0 + (i < cnt ? S.idx[cur][i] : 0)];
    }
#pragma unroll
    for (int j = 0; j < SJ; ++j) {
        const int i = j * GT + threadIdx.x;
        if (i < cnt) {
            kout[s0 + i] = high | (uint64_t)S.key[cur][i];
            vout[s0 + i] = vv[j];
        }
    }
}

// LDS sort of cnt <= GS_CAP pairs of a group whose local keys have at most
// GS_PK_BITS + DB bits: one 32-bit word per pair and pass, the key bits not
// yet ranked above the pair's position in the group (GS_PK_POS bits).  The
// first pass ranks the digits straight from the loaded keys; the last one
// leaves the positions in sorted order, and each pair's key and value are
// gathered from the group's input by position.  Half the LDS of the
// two-array sort (16-bit per-wave digit counts too): five workgroups fit a
// CU, so the bench-sized batch of ~1,270 groups runs in one round.
constexpr int GS_PK_POS = 12;                 // position bits (GS_CAP <= 4096)
constexpr int GS_PK_BITS = 32 - GS_PK_POS;    // key bits a word carries
static_assert(GS_CAP <= (1 << GS_PK_POS), "a position fits its field");
static_assert(GS_CAP <= 65535, "16-bit digit counts");

template <int DB>
struct GsPacked {
    uint32_t w[2][GS_CAP];
    uint16_t wcnt[GT / 64][1 << DB];
    uint32_t loff[1 << DB];
    uint32_t lds4[4];
};

template <int DB>
__device__ void group_sort_packed(GsPacked<DB> &S, const uint64_t *__restrict__ kin,
                                  const uint32_t *__restrict__ vin, uint64_t *__restrict__ kout,
                                  uint32_t *__restrict__ vout, int64_t s0, int cnt, int lb)
{
    constexpr uint32_t DM = (1u << DB) - 1u;
    constexpr uint32_t PM = (1u << GS_PK_POS) - 1u;
    const int w = threadIdx.x >> 6, lane = lane_id();
    const uint64_t lmask = (1ull << lb) - 1ull;
    // positions per wave: a multiple of 64 covering cnt / 4; item j of a
    // lane sits at w * pw + 64 j + lane in every pass
    const int jmax = (cnt + GT - 1) / GT;
    const int pw = jmax * 64;
    uint32_t dg[SJ], rnk[SJ], wd[SJ];
    bool real[SJ];
    {
        uint64_t kl[SJ];
#pragma unroll
        for (int j = 0; j < SJ; ++j) {
            const int p = w * pw + j * 64 + lane;
            real[j] = j < jmax && p < cnt;
            kl[j] = real[j] ? kin[s0 + p] : 0ull;
        }
#pragma unroll
        for (int j = 0; j < SJ; ++j) {
            const uint64_t lk = kl[j] & lmask;
            dg[j] = (uint32_t)lk & DM;
            wd[j] = ((uint32_t)(lk >> DB) << GS_PK_POS) | (uint32_t)(w * pw + j * 64 + lane);
        }
    }
    int cur = 0;
    for (int sh = 0;; sh += DB) {
        rank_chunk<SJ, DB, uint16_t>(dg, real, jmax, rnk, S.wcnt, S.loff, S.lds4);
#pragma unroll
        for (int j = 0; j < SJ; ++j) {
            if (!real[j]) continue;
            S.w[cur][S.loff[dg[j]] + S.wcnt[w][dg[j]] + rnk[j]] = wd[j];
        }
        __syncthreads();
        if (sh + DB >= lb) break;
        // the next digit: the lowest key bits the words still carry
#pragma unroll
        for (int j = 0; j < SJ; ++j) {
            const int p = w * pw + j * 64 + lane;
            wd[j] = real[j] ? S.w[cur][p] : 0u;
            dg[j] = (wd[j] >> GS_PK_POS) & DM;
            wd[j] = (((wd[j] >> GS_PK_POS) >> DB) << GS_PK_POS) | (wd[j] & PM);
        }
        cur ^= 1;
    }
    // S.w[cur]: the positions in sorted order; the key and value gathers of
    // every item in flight before the first store
    uint64_t kk[SJ];
    uint32_t vv[SJ];
#pragma unroll
    for (int j = 0; j < SJ; ++j) {
        const int i = j * GT + threadIdx.x;
        const uint32_t src = i < cnt ? (S.w[cur][i] & PM) : 0u;
        kk[j] = kin[s0 + src];
        vv[j] = vin[s0 + src];
    }
#pragma unroll
    for (int j = 0; j < SJ; ++j) {
        const int i = j * GT + threadIdx.x;
        if (i < cnt) {
            kout[s0 + i] = kk[j];
            vout[s0 + i] = vv[j];
        }
    }
}

// A group larger than GS_CAP: stable LSD passes out of global memory by this
// workgroup alone, chunk after chunk (each chunk ranked as above, each
// digit's running offset carried in LDS).  Ping-pongs between the group's
// ranges of (a_k, a_v) (input) and (b_k, b_v) and ends in (b_k, b_v).
struct GsBig {
    uint32_t wcnt[GT / 64][256];
    uint32_t loff[256];
    uint32_t run[256];
    uint32_t lds4[4];
};

__device__ void group_sort_big(GsBig &S, uint64_t *__restrict__ a_k, uint32_t *__restrict__ a_v,
                               uint64_t *__restrict__ b_k, uint32_t *__restrict__ b_v, int64_t s0, int64_t cnt,
                               int lb)
{
    constexpr int BJ = 2;                 // items per lane per chunk
    constexpr int CH = GT * BJ;           // 512 pairs per chunk
    const int w = threadIdx.x >> 6, lane = lane_id();
    uint64_t *sk = a_k, *dk = b_k;
    uint32_t *sv = a_v, *dv = b_v;
    bool in_b = false;
    for (int sh = 0; sh < lb; sh += 8) {
        // digit counts of the whole group
        S.run[threadIdx.x] = 0;
        __syncthreads();
        for (int64_t i = threadIdx.x; i < cnt; i += GT) atomicAdd(&S.run[(uint32_t)(sk[s0 + i] >> sh) & 255u], 1u);
        __syncthreads();
        {
            uint32_t tot;
            const uint32_t c = S.run[threadIdx.x];
            __syncthreads();
            S.run[threadIdx.x] = block_excl_scan_256(c, S.lds4, &tot);
        }
        __syncthreads();
        for (int64_t c0 = 0; c0 < cnt; c0 += CH) {
            const int64_t left = cnt - c0;
            const int jmax = left >= CH ? BJ : (int)((left + GT - 1) / GT);
            const int pw = jmax * 64;
            // only digits and ranks stay in registers; the pairs are read
            // again for the scatter (L2 hits)
            uint32_t dg[BJ], rnk[BJ];
            bool real[BJ];
#pragma unroll
            for (int j = 0; j < BJ; ++j) {
                const int64_t p = (int64_t)w * pw + j * 64 + lane;
                real[j] = j < jmax && p < left;
                dg[j] = real[j] ? (uint32_t)(sk[s0 + c0 + p] >> sh) & 255u : 0u;
            }
            rank_chunk<BJ>(dg, real, jmax, rnk, S.wcnt, S.loff, S.lds4);
#pragma unroll
            for (int j = 0; j < BJ; ++j) {
                if (!real[j]) continue;
                const int64_t p = (int64_t)w * pw + j * 64 + lane;
                const int64_t o = (int64_t)S.run[dg[j]] + S.wcnt[w][dg[j]] + rnk[j];
                dk[s0 + o] = sk[s0 + c0 + p];
                dv[s0 + o] = sv[s0 + c0 + p];
            }
            __syncthreads();
            // the chunk's count of digit d: loff[d + 1] - loff[d]
            {
                const int d = threadIdx.x;
                const uint32_t c = (d == 255 ? (uint32_t)(left < CH ? left : CH) : S.loff[d + 1]) - S.loff[d];
                S.run[d] += c;
            }
            __syncthreads();
        }
        uint64_t *tk = sk; sk = dk; dk = tk;
        uint32_t *tv = sv; sv = dv; dv = tv;
        in_b = !in_b;
    }
    if (!in_b) {
        for (int64_t i = threadIdx.x; i < cnt; i += GT) {
            b_k[s0 + i] = a_k[s0 + i];
            b_v[s0 + i] = a_v[s0 + i];
        }
    }
}

// One workgroup per digit (group); the last digit holds the invalid keys,
// copied as they are.  In: (kin, vin) in group order; out: (kout, vout)
// sorted.  gbase: ndig + 1 exclusive group offsets.
template <typename K, int DB>
__global__ __launch_bounds__(GT) void k_gs_sort(uint64_t *__restrict__ kin, uint32_t *__restrict__ vin,
                                                uint64_t *__restrict__ kout, uint32_t *__restrict__ vout, int lb,
                                                int ndig, const uint32_t *__restrict__ gbase)
{
    __shared__ union {
        GsLds<K, DB> lds;
        GsBig big;
    } S;
    const int g = blockIdx.x;
    const int64_t s0 = gbase[g];
    const int64_t cnt = (int64_t)gbase[g + 1] - s0;
    if (cnt <= 0) return;
    if (g == ndig - 1) {
        for (int64_t i = threadIdx.x; i < cnt; i += GT) {
            kout[s0 + i] = kin[s0 + i];
            vout[s0 + i] = vin[s0 + i];
        }
        return;
    }
    if (cnt <= GS_CAP) {
        group_sort_lds<K, DB>(S.lds, kin, vin, kout, vout, s0, (int)cnt, lb, (uint64_t)g << lb);
    } else {
        group_sort_big(S.big, kin, vin, kout, vout, s0, cnt, lb);
    }
}

// The same with the one-word LDS sort (lb <= GS_PK_BITS + DB).
template <int DB>
__global__ __launch_bounds__(GT) __attribute__((amdgpu_waves_per_eu(5))) void k_gs_sort_packed(
    uint64_t *__restrict__ kin, uint32_t *__restrict__ vin, uint64_t *__restrict__ kout, uint32_t *__restrict__ vout,
    int lb, int ndig, const uint32_t *__restrict__ gbase)
{
    __shared__ union {
        GsPacked<DB> pk;
        GsBig big;
    } S;
    const int g = blockIdx.x;
    const int64_t s0 = gbase[g];
    const int64_t cnt = (int64_t)gbase[g + 1] - s0;
    if (cnt <= 0) return;
    if (g == ndig - 1) {
        for (int64_t i = threadIdx.x; i < cnt; i += GT) {
            kout[s0 + i] = kin[s0 + i];
            vout[s0 + i] = vin[s0 + i];
        }
        return;
    }
    if (cnt <= GS_CAP) {
        group_sort_packed<DB>(S.pk, kin, vin, kout, vout, s0, (int)cnt, lb);
    } else {
        group_sort_big(S.big, kin, vin, kout, vout, s0, cnt, lb);
    }
}

}  // namespace

size_t gs_hist_words(int64_t n, int ndig)
{
    const int64_t nb = (n + GTILE - 1) / GTILE;
    const int64_t blocks = (ndig + 63) / 64;
    return (size_t)((nb < 1 ? 1 : nb) * 64 * blocks) + 2 * (GS_MAX_DIGITS + 64);
}

hipError_t group_sort_pairs(uint64_t *k0, uint32_t *v0, uint64_t *k1, uint32_t *v1, int64_t n, int lb, int ndig,
                            uint32_t *hist, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (ndig < 1 || ndig > GS_MAX_DIGITS || lb < 1 || lb > 62) return hipErrorInvalidValue;
    const int64_t nb = (n + GTILE - 1) / GTILE;
    const int blocks = (ndig + 63) / 64;
    uint32_t *dtot = hist + nb * 64 * blocks;
    uint32_t *gbase = dtot + GS_MAX_DIGITS + 64;
    int dbits = 1;
    while ((1 << dbits) < ndig) ++dbits;
    k_gs_scan<<<blocks, GS_SCAN_T, 0, s>>>(hist, nb, dtot);
    k_gs_scatter<<<(unsigned)nb, GT, 0, s>>>(k0, v0, k1, v1, n, lb, ndig, dbits, hist, nb, dtot, gbase);
    // 9-bit digits where they save a pass over the lb local bits
    const bool nine = (lb + 8) / 9 < (lb + 7) / 8;
    if (lb <= GS_PK_BITS + (nine ? 9 : 8)) {
        if (nine) k_gs_sort_packed<9><<<ndig, GT, 0, s>>>(k1, v1, k0, v0, lb, ndig, gbase);
        else k_gs_sort_packed<8><<<ndig, GT, 0, s>>>(k1, v1, k0, v0, lb, ndig, gbase);
    } else if (lb <= 32) {
        // lb in 29..32 here, where 9-bit digits take as many passes as 8-bit
        k_gs_sort<uint32_t, 8><<<ndig, GT, 0, s>>>(k1, v1, k0, v0, lb, ndig, gbase);
    } else {
        if (nine) k_gs_sort<uint64_t, 9><<<ndig, GT, 0, s>>>(k1, v1, k0, v0, lb, ndig, gbase);
        else k_gs_sort<uint64_t, 8><<<ndig, GT, 0, s>>>(k1, v1, k0, v0, lb, ndig, gbase);
    }
    return hipGetLastError();
}

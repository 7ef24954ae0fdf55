// spgemm.hip -- C = A * B for DCSR operands.
//
// Replaces device::spmm(const dcsr& A, const dcsr& B, dcsr& C, streams)
// (include/spmm.h:30-391) and its binned ESC kernels (include/spmm.inl:361-668,
// 951-1061).  Result semantics are those of the host path
// (cusp/system/detail/sequential/multiply/csr_spgemm.h:75-150): the structure
// holds every (i, k) with a partial product whose FP32 sum, accumulated from
// 0 in A-row order then B-row order, is not exactly 0.
//
// Rows are bucketed by their partial-product count P_i:
//   P_i <=  128  a 16-lane team per row, four rows per wave, the same LDS
//                ESC as below with small per-row buffers (many rows in
//                flight per CU: a row is a chain of dependent loads);
//   P_i <= 1024  one wave per row, LDS ESC: the row's partials are expanded
//                with their generation index (A-entry order, then B-row
//                order), bitonic-sorted in LDS by (column, generation), and
//                each run of equal columns is summed left-to-right;
//   P_i <= 4096  the same with one 256-thread workgroup per row;
//   P_i >  4096  (power-law hub rows, which the reference skips) a global
//                ESC: expand, stable LSD radix sort by (row, column), reduce.
// Output rows come out column-sorted with exact zeros removed, and the
// accumulation order is fixed, so C is deterministic.
#include "dcsr_internal.h"
#include "prims.h"
#include "segments.h"

namespace {

constexpr int TB = 256;
constexpr uint32_t TEAM_MAX = 128;    // partials per row for the team kernel
constexpr int TEAM = 16;              // lanes per row in the team kernel
static_assert(TEAM_MAX <= 8 * TEAM, "team rows: the rank sort holds at most 8 keys per lane");
constexpr int TEAM_GB = 7;            // generation bits of the team kernel's 32-bit keys
// the team kernel's keys are 32-bit: B's columns must stay below this
constexpr int64_t TEAM_MAX_COLS = (1ll << (32 - TEAM_GB)) - 1;
constexpr int WAVE_GB = 10;           // generation bits of the wave kernel's keys
// the wave kernel's keys are 32-bit up to this many columns of B, else 64-bit
constexpr int64_t WAVE_MAX_COLS32 = (1ll << (32 - WAVE_GB)) - 1;
constexpr uint32_t WAVE_MAX = 1024;   // partials per row for the wave kernel
constexpr uint32_t BLOCK_MAX = 4096;  // partials per row for the workgroup kernel
constexpr int SG_EX = 8;              // partials per lane whose loads are in flight together
// row classes (list index): wave, workgroup, hub, team
// team classes CL_TEAM + i: rows of at most 2 (i + 1) * TEAM partials, sorted
// with 2 (i + 1) keys per lane
enum { CL_WAVE = 0, CL_BLOCK = 1, CL_HUB = 2, CL_TEAM = 3, NTEAMCL = 4, NCLS = CL_TEAM + NTEAMCL };

struct SgState {
    unsigned long long total;  // sum of P_i (64-bit)
    unsigned int count[NCLS];  // rows per class
    unsigned int pad;
};

__device__ __forceinline__ uint32_t next_pow2(uint32_t v)
{
    v--;
    v |= v >> 1;
    v |= v >> 2;
    v |= v >> 4;
    v |= v >> 8;
    v |= v >> 16;
    return v + 1;
}

// P_i = sum over the entries of A row i of len(B row col).  One wave per 64
// consecutive rows; entries folded per row with a DPP segmented scan.
__global__ __launch_bounds__(TB) void k_sg_count(uint32_t R, const uint32_t *__restrict__ a_start,
                                                 const int32_t *__restrict__ a_len, const int32_t *__restrict__ a_col,
                                                 const int32_t *__restrict__ b_len, uint32_t *__restrict__ P)
{
    __shared__ uint32_t s_ex[TB / 64][64];
    __shared__ uint32_t s_st[TB / 64][64];
    __shared__ uint32_t s_acc[TB / 64][64];
    const int w = threadIdx.x >> 6;
    const int lane = lane_id();
    const uint32_t r = (blockIdx.x * (TB / 64) + w) * 64u + lane;
    uint32_t st = 0, L = 0;
    if (r < R) {
        st = a_start[r];
        L = (uint32_t)a_len[r];
    }
    const uint32_t inc = wave_incl_scan_u32(L);
    const uint32_t tot = wave_read_lane_u32(inc, 63);
    s_ex[w][lane] = inc - L;
    s_st[w][lane] = st;
    s_acc[w][lane] = 0;
    __builtin_amdgcn_wave_barrier();
    // four 64-entry chunks per round: their A-column loads, then their B-row
    // length loads, each issued together (two round trips per 256 entries)
    constexpr int CU = 4;
    for (uint32_t b = 0; b < tot; b += 64 * CU) {
        uint32_t j[CU], p[CU];
        int32_t c[CU];
#pragma unroll
        for (int u = 0; u < CU; ++u) {
            const uint32_t e = b + (uint32_t)u * 64u + lane;
            j[u] = 64;
            c[u] = -1;
            if (e < tot) {
                j[u] = (uint32_t)find_seg64(s_ex[w], e);
                c[u] = a_col[(uint64_t)s_st[w][j[u]] + (e - s_ex[w][j[u]])];
            }
        }
#pragma unroll
        for (int u = 0; u < CU; ++u) p[u] = c[u] >= 0 ? (uint32_t)b_len[c[u]] : 0u;
#pragma unroll
        for (int u = 0; u < CU; ++u) {
            const uint32_t q = wave_seg_incl_scan_u32(p[u], j[u]);
            const uint32_t jn = __shfl_down(j[u], 1);
            if (j[u] < 64 && (lane == 63 || jn != j[u])) s_acc[w][j[u]] += q;
            __builtin_amdgcn_wave_barrier();
        }
    }
    if (r < R) P[r] = s_acc[w][lane];
}

// Row classes (list appends aggregated per workgroup: one global atomic per
// class per workgroup of SG_CR * TB rows) and per-block 64-bit P sums.  Each
// thread classifies SG_CR rows; with 256 rows per workgroup the class
// counters took ~60K contended atomics per product (94 us on 2M rows).
constexpr int SG_CR = 8;
__global__ __launch_bounds__(TB) void k_sg_classify(uint32_t R, const uint32_t *__restrict__ P,
                                                    uint32_t *__restrict__ lists, uint32_t list_stride,
                                                    SgState *st, uint32_t *__restrict__ cnt,
                                                    unsigned long long *__restrict__ bsum, uint32_t team_ok)
{
    __shared__ unsigned long long ws[TB / 64];
    __shared__ uint32_t s_cnt[NCLS], s_base[NCLS];
    if (threadIdx.x < NCLS) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t r0 = blockIdx.x * (TB * SG_CR) + threadIdx.x;
    uint32_t p[SG_CR], slot[SG_CR];
    int cls[SG_CR];
#pragma unroll
    for (int q = 0; q < SG_CR; ++q) {
        const uint32_t r = r0 + (uint32_t)q * TB;
        p[q] = r < R ? P[r] : 0u;
    }
    unsigned long long v = 0;
#pragma unroll
    for (int q = 0; q < SG_CR; ++q) {
        const uint32_t r = r0 + (uint32_t)q * TB;
        cls[q] = -1;
        slot[q] = 0;
        if (r < R) {
            // team rows in classes by rank-sort width (2, 4, 6 or 8 keys per
            // lane), so the four teams of a wave run the same sort over similar
            // row sizes
            cls[q] = p[q] == 0                         ? -1
                     : (p[q] <= TEAM_MAX && team_ok)   ? CL_TEAM + (int)((p[q] - 1) / (2 * TEAM))
                     : p[q] <= WAVE_MAX                ? CL_WAVE
                     : p[q] <= BLOCK_MAX               ? CL_BLOCK
                                                       : CL_HUB;
            if (p[q] == 0) cnt[r] = 0;
        }
        v += p[q];
        // slot of this row inside its workgroup's share of its class list
#pragma unroll
        for (int c = 0; c < NCLS; ++c) {
            const uint64_t m = __ballot(cls[q] == c);
            if (!m) continue;
            const int leader = __ffsll((long long)m) - 1;
            uint32_t base = 0;
            if (lane_id() == leader) base = atomicAdd(&s_cnt[c], (uint32_t)__popcll(m));
            base = __shfl(base, leader);
            if (cls[q] == c) slot[q] = base + (uint32_t)__popcll(m & lanemask_lt());
        }
    }
    __syncthreads();
    if (threadIdx.x < NCLS && s_cnt[threadIdx.x])
        s_base[threadIdx.x] = atomicAdd(&st->count[threadIdx.x], s_cnt[threadIdx.x]);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SG_CR; ++q)
        if (cls[q] >= 0) lists[(uint64_t)cls[q] * list_stride + s_base[cls[q]] + slot[q]] = r0 + (uint32_t)q * TB;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    if (lane_id() == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) bsum[blockIdx.x] = (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

__global__ void k_sg_state_zero(SgState *st)
{
    st->total = 0;
    for (int c = 0; c < NCLS; ++c) st->count[c] = 0;
}

__global__ __launch_bounds__(1024) void k_sg_total(SgState *st, const unsigned long long *__restrict__ bsum, int nb)
{
    __shared__ unsigned long long red[16];
    unsigned long long v = 0;
    for (int b = threadIdx.x; b < nb; b += 1024) v += bsum[b];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    if (lane_id() == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int i = 0; i < 16; ++i) t += red[i];
        st->total = t;
    }
}

// ---------------------------------------------------------------------------
// Shared row routine: expand, sort, reduce.  NT threads cooperate (64 for the
// wave kernel, 256 for the workgroup kernel); GB = generation-index bits.
// ---------------------------------------------------------------------------
// A team is NT threads: 16 (a quarter wave), 64 (a wave) or 256 (the
// workgroup).  Teams inside one wave never exchange data with each other.
template <int NT>
__device__ __forceinline__ void team_sync()
{
    if (NT <= 64) __builtin_amdgcn_wave_barrier();
    else __syncthreads();
}

// Exclusive scan of one uint32 per team thread.
template <int NT>
__device__ __forceinline__ uint32_t team_excl_scan(uint32_t v, uint32_t *lds_w, uint32_t *total)
{
    if (NT < 64) {
        // within the team's aligned NT-lane slice of the wave
        const int t = lane_id() % NT;
        uint32_t inc = v;
#pragma unroll
        for (int d = 1; d < NT; d <<= 1) {
            const uint32_t u = __shfl_up(inc, d, NT);
            if (t >= d) inc += u;
        }
        *total = __shfl(inc, NT - 1, NT);
        return inc - v;
    }
    const uint32_t inc = wave_incl_scan_u32(v);
    if (NT == 64) {
        *total = wave_read_lane_u32(inc, 63);
        return inc - v;
    }
    const int w = (threadIdx.x % NT) >> 6;
    if (lane_id() == 63) lds_w[w] = inc;
    __syncthreads();
    uint32_t carry = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) {
        const uint32_t t = lds_w[i];
        if (i < w) carry += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return carry + inc - v;
}

// Largest j in [0, NT) with s[j] <= e.
template <int NT>
__device__ __forceinline__ uint32_t find_seg_team(const uint32_t *s, uint32_t e)
{
    uint32_t lo = 0, hi = NT;
#pragma unroll
    for (int it = 0; (1 << it) < NT; ++it) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s[mid] <= e) lo = mid; else hi = mid;
    }
    return lo;
}

// K: the sort key type, (column << GB) | generation.  32-bit keys halve the
// LDS per row (more rows in flight); they need columns below 2^(32-GB) - 1
// so that no key reaches the ~0 padding.
// What a row needs before its expansion: its partial count, output offset
// and A row.  Loaded one row ahead by the kernels' row loops (the loads of
// the next row are in flight while the current row is processed).
struct SgRowInfo {
    uint32_t row, P, ast, la;
    uint64_t out_base;
};

__device__ __forceinline__ SgRowInfo sg_row_info(uint32_t row, const uint32_t *__restrict__ P,
                                                 const uint32_t *__restrict__ Poff,
                                                 const uint32_t *__restrict__ a_start,
                                                 const int32_t *__restrict__ a_len)
{
    SgRowInfo r;
    r.row = row;
    r.P = P[row];
    r.out_base = Poff[row];
    r.ast = a_start[row];
    r.la = (uint32_t)a_len[row];
    return r;
}

// Rank sort of a team's P staged (key, val) pairs in LDS, for teams narrower
// than a wave.  The keys are distinct, so a key's sorted position is the
// number of smaller keys: lane t holds keys t, t + NT, ... in registers and
// compares them with every key of the row, read back one at a time (the
// team's lanes all read the same word: one broadcast per key), then every
// pair is written to its position.  Costs P * KPL compares per lane and no
// LDS traffic beyond the broadcasts, against a bitonic network's
// log^2-many LDS passes.
template <int NT, int KPL, typename K>
__device__ __forceinline__ void team_rank_sort(K *key, float *val, uint32_t P)
{
    const uint32_t t = threadIdx.x % NT;
    K mk[KPL];
    float mv[KPL];
    uint32_t rk[KPL];
#pragma unroll
    for (int r = 0; r < KPL; ++r) {
        const uint32_t i = t + (uint32_t)r * NT;
        mk[r] = i < P ? key[i] : (K)~0ull;
        mv[r] = i < P ? val[i] : 0.0f;
        rk[r] = 0;
    }
    uint32_t j = 0;
    for (; j + 4 <= P; j += 4) {
        K k0, k1, k2, k3;
        if constexpr (sizeof(K) == 4) {
            // four keys per read (the team's key row is 16-byte aligned)
            const uint4 q = *reinterpret_cast<const uint4 *>(key + j);
            k0 = q.x;
            k1 = q.y;
            k2 = q.z;
            k3 = q.w;
        } else {
            k0 = key[j];
            k1 = key[j + 1];
            k2 = key[j + 2];
            k3 = key[j + 3];
        }
#pragma unroll
        for (int r = 0; r < KPL; ++r)
            rk[r] += (uint32_t)(k0 < mk[r]) + (uint32_t)(k1 < mk[r]) + (uint32_t)(k2 < mk[r]) +
                     (uint32_t)(k3 < mk[r]);
    }
    for (; j < P; ++j) {
        const K kj = key[j];
#pragma unroll
        for (int r = 0; r < KPL; ++r) rk[r] += (uint32_t)(kj < mk[r]);
    }
    team_sync<NT>();  // every read of the row is done before the scatter
#pragma unroll
    for (int r = 0; r < KPL; ++r) {
        if (t + (uint32_t)r * NT < P) {
            key[rk[r]] = mk[r];
            val[rk[r]] = mv[r];
        }
    }
    team_sync<NT>();
}

// KPL (teams only): keys per lane of the rank sort, fixed by the row's class
// (P <= KPL * NT); 0 picks it per row.
template <int NT, int GB, typename K, int KPL = 0>
__device__ void sg_row(const SgRowInfo &ri, const int32_t *__restrict__ a_col, const float *__restrict__ a_val,
                       const uint32_t *__restrict__ b_start, const int32_t *__restrict__ b_len,
                       const int32_t *__restrict__ b_col, const float *__restrict__ b_val,
                       int32_t *__restrict__ tcol, float *__restrict__ tval, uint32_t *__restrict__ cnt, K *key,
                       float *val, uint32_t *s_ex, uint32_t *s_sb, float *s_va, uint32_t *lds_w)
{
    const uint32_t t = threadIdx.x % NT;
    const uint32_t row = ri.row, P = ri.P, ast = ri.ast, la = ri.la;
    const uint64_t out_base = ri.out_base;
    // 1. expand in generation order
    uint32_t gbase = 0;
    for (uint32_t a0 = 0; a0 < la; a0 += NT) {
        const uint32_t a = a0 + t;
        uint32_t lb = 0, sb = 0;
        float va = 0.0f;
        if (a < la) {
            const int32_t j = a_col[(uint64_t)ast + a];
            va = a_val[(uint64_t)ast + a];
            lb = (uint32_t)b_len[j];
            sb = b_start[j];
        }
        uint32_t tot;
        const uint32_t ex = team_excl_scan<NT>(lb, lds_w, &tot);
        s_ex[t] = ex;
        s_sb[t] = sb;
        s_va[t] = va;
        team_sync<NT>();
        // up to SG_EX partials per lane per round, every load of a round
        // issued before the first wait (a lane past tot reads entry 0 of B
        // and writes nothing)
        constexpr int EX = KPL ? KPL : SG_EX;
        for (uint32_t e0 = 0; e0 < tot; e0 += EX * NT) {
            int32_t bc[EX];
            float bv[EX], av[EX];
#pragma unroll
            for (int i = 0; i < EX; ++i) {
                const uint32_t pe = e0 + (uint32_t)i * NT + t;
                const uint32_t q = find_seg_team<NT>(s_ex, pe < tot ? pe : 0u);
                const uint64_t bp = pe < tot ? (uint64_t)s_sb[q] + (pe - s_ex[q]) : 0u;
                bc[i] = b_col[bp];
                bv[i] = b_val[bp];
                av[i] = s_va[q];
            }
#pragma unroll
            for (int i = 0; i < EX; ++i) {
                const uint32_t pe = e0 + (uint32_t)i * NT + t;
                if (pe < tot) {
                    const uint32_t g = gbase + pe;
                    key[g] = (K)(((uint64_t)(uint32_t)bc[i] << GB) | g);
                    val[g] = av[i] * bv[i];
                }
            }
        }
        gbase += tot;
        team_sync<NT>();
    }
    // 2. sort by (column, generation): rank sort for quarter-wave teams and
    // for wave rows of up to 256 partials, bitonic network over LDS for the
    // others (padding sorts last)
    if constexpr (NT < 64 && KPL != 0) {
        team_rank_sort<NT, KPL, K>(key, val, P);
    } else if constexpr (NT < 64) {
        if (P <= 4 * NT) team_rank_sort<NT, 4, K>(key, val, P);
        else team_rank_sort<NT, 8, K>(key, val, P);
    } else if (NT == 64 && P <= 4 * NT) {
        // wave rows of up to 256 partials: the rank sort's P * 4 compares per
        // lane cost less than the bitonic network's LDS passes (k_sg_wave
        // 808 -> 538 us on the 2M-row Poisson(8) product); up to 512 with 8
        // keys per lane measured the same as the network
        team_rank_sort<NT, 4, K>(key, val, P);
    } else {
        const uint32_t n2 = next_pow2(P < 2 * NT ? 2 * NT : P);
        for (uint32_t p = P + t; p < n2; p += NT) key[p] = (K)~0ull;
        team_sync<NT>();
        for (uint32_t k = 2; k <= n2; k <<= 1) {
            for (uint32_t js = k >> 1; js > 0; js >>= 1) {
                for (uint32_t c = t; c < n2 / 2; c += NT) {
                    const uint32_t i1 = (c / js) * 2 * js + (c % js);
                    const uint32_t i2 = i1 + js;
                    const bool asc = (i1 & k) == 0;
                    const K ka = key[i1], kb = key[i2];
                    if ((ka > kb) == asc) {
                        key[i1] = kb;
                        key[i2] = ka;
                        const float va = val[i1];
                        val[i1] = val[i2];
                        val[i2] = va;
                    }
                }
                team_sync<NT>();
            }
        }
    }
    // 3. runs of equal columns, summed left to right from 0; drop exact zeros
    uint32_t outc = 0;
    for (uint32_t p0 = 0; p0 < P; p0 += NT) {
        const uint32_t p = p0 + t;
        bool emit = false;
        float s = 0.0f;
        uint32_t c = 0;
        if (p < P) {
            c = (uint32_t)(key[p] >> GB);
            const bool head = p == 0 || (uint32_t)(key[p - 1] >> GB) != c;
            if (head) {
                s = 0.0f + val[p];
                for (uint32_t q = p + 1; q < P && (uint32_t)(key[q] >> GB) == c; ++q) s = s + val[q];
                emit = s != 0.0f;
            }
        }
        uint32_t tot;
        const uint32_t pos = team_excl_scan<NT>(emit ? 1u : 0u, lds_w, &tot);
        if (emit) {
            tcol[out_base + outc + pos] = (int32_t)c;
            tval[out_base + outc + pos] = s;
        }
        outc += tot;
    }
    if (t == 0) cnt[row] = outc;
    team_sync<NT>();
}

// Team loop over a class list: team `first`, stride `step`.  The next row's
// id and info are loaded one and two iterations ahead.
template <typename F>
__device__ __forceinline__ void sg_rows(const uint32_t *__restrict__ list, uint32_t n, uint32_t first, uint32_t step,
                                        const uint32_t *__restrict__ P, const uint32_t *__restrict__ Poff,
                                        const uint32_t *__restrict__ a_start, const int32_t *__restrict__ a_len,
                                        F &&body)
{
    if (first >= n) return;
    SgRowInfo cur = sg_row_info(list[first], P, Poff, a_start, a_len);
    uint32_t next_id = first + step < n ? list[first + step] : 0u;
    for (uint32_t i = first; i < n; i += step) {
        SgRowInfo nxt{};
        if (i + step < n) nxt = sg_row_info(next_id, P, Poff, a_start, a_len);
        const uint32_t nn_id = i + 2 * step < n ? list[i + 2 * step] : 0u;
        body(cur);
        cur = nxt;
        next_id = nn_id;
    }
}

// Rows of at most WAVE_MAX partials: one wave per row.  K: 32-bit sort keys
// (half the LDS per row, more rows in flight) when B's columns allow.
template <typename K>
__global__ __launch_bounds__(TB) void k_sg_wave(const uint32_t *__restrict__ list, const SgState *__restrict__ st,
                                                const uint32_t *__restrict__ P, const uint32_t *__restrict__ Poff,
                                                const uint32_t *__restrict__ a_start, const int32_t *__restrict__ a_len,
                                                const int32_t *__restrict__ a_col, const float *__restrict__ a_val,
                                                const uint32_t *__restrict__ b_start, const int32_t *__restrict__ b_len,
                                                const int32_t *__restrict__ b_col, const float *__restrict__ b_val,
                                                int32_t *__restrict__ tcol, float *__restrict__ tval,
                                                uint32_t *__restrict__ cnt)
{
    __shared__ K key[TB / 64][WAVE_MAX];
    __shared__ float val[TB / 64][WAVE_MAX];
    __shared__ uint32_t s_ex[TB / 64][64], s_sb[TB / 64][64];
    __shared__ float s_va[TB / 64][64];
    const int w = threadIdx.x >> 6;
    const uint32_t n = st->count[CL_WAVE];
    sg_rows(list, n, blockIdx.x * (TB / 64) + w, gridDim.x * (TB / 64), P, Poff, a_start, a_len,
            [&](const SgRowInfo &ri) {
                sg_row<64, WAVE_GB, K>(ri, a_col, a_val, b_start, b_len, b_col, b_val, tcol, tval, cnt, key[w],
                                       val[w], s_ex[w], s_sb[w], s_va[w], nullptr);
            });
}

// Rows of at most KPL * TEAM partials (class CL_TEAM + KPL / 2 - 1): 16-lane
// teams, four rows per wave.
template <int KPL>
__global__ __launch_bounds__(TB) void k_sg_team(const uint32_t *__restrict__ list, const SgState *__restrict__ st,
                                                const uint32_t *__restrict__ P, const uint32_t *__restrict__ Poff,
                                                const uint32_t *__restrict__ a_start, const int32_t *__restrict__ a_len,
                                                const int32_t *__restrict__ a_col, const float *__restrict__ a_val,
                                                const uint32_t *__restrict__ b_start, const int32_t *__restrict__ b_len,
                                                const int32_t *__restrict__ b_col, const float *__restrict__ b_val,
                                                int32_t *__restrict__ tcol, float *__restrict__ tval,
                                                uint32_t *__restrict__ cnt)
{
    constexpr int NTEAM = TB / TEAM;
    // per-team rows padded so the four teams of a wave start on different
    // LDS banks; the pitch keeps every key row 16-byte aligned
    constexpr uint32_t PITCH = TEAM_MAX + 20;
    static_assert(PITCH % 4 == 0, "16-byte aligned key rows");
    __shared__ __align__(16) uint32_t key[NTEAM][PITCH];
    __shared__ float val[NTEAM][PITCH];
    __shared__ uint32_t s_ex[NTEAM][TEAM], s_sb[NTEAM][TEAM];
    __shared__ float s_va[NTEAM][TEAM];
    const int tm = threadIdx.x / TEAM;
    static_assert(KPL * TEAM <= TEAM_MAX, "a team row fits its LDS row");
    const uint32_t n = st->count[CL_TEAM + KPL / 2 - 1];
    sg_rows(list, n, blockIdx.x * NTEAM + tm, gridDim.x * NTEAM, P, Poff, a_start, a_len,
            [&](const SgRowInfo &ri) {
                sg_row<TEAM, TEAM_GB, uint32_t, KPL>(ri, a_col, a_val, b_start, b_len, b_col, b_val, tcol, tval,
                                                     cnt, key[tm], val[tm], s_ex[tm], s_sb[tm], s_va[tm], nullptr);
            });
}

__global__ __launch_bounds__(TB) void k_sg_block(const uint32_t *__restrict__ list, const SgState *__restrict__ st,
                                                 const uint32_t *__restrict__ P, const uint32_t *__restrict__ Poff,
                                                 const uint32_t *__restrict__ a_start, const int32_t *__restrict__ a_len,
                                                 const int32_t *__restrict__ a_col, const float *__restrict__ a_val,
                                                 const uint32_t *__restrict__ b_start, const int32_t *__restrict__ b_len,
                                                 const int32_t *__restrict__ b_col, const float *__restrict__ b_val,
                                                 int32_t *__restrict__ tcol, float *__restrict__ tval,
                                                 uint32_t *__restrict__ cnt)
{
    __shared__ uint64_t key[BLOCK_MAX];
    __shared__ float val[BLOCK_MAX];
    __shared__ uint32_t s_ex[TB], s_sb[TB];
    __shared__ float s_va[TB];
    __shared__ uint32_t lds_w[TB / 64];
    const uint32_t n = st->count[CL_BLOCK];
    sg_rows(list, n, blockIdx.x, gridDim.x, P, Poff, a_start, a_len, [&](const SgRowInfo &ri) {
        sg_row<TB, 12, uint64_t>(ri, a_col, a_val, b_start, b_len, b_col, b_val, tcol, tval, cnt, key, val, s_ex, s_sb,
                                 s_va, lds_w);
    });
}

// ---------------------------------------------------------------------------
// global path for hub rows (P_i > BLOCK_MAX)
// ---------------------------------------------------------------------------
// pre[2h] = len(A row), pre[2h+1] = P for hub row h (index hn: zeros)
__global__ void k_sg_hprep(const uint32_t *__restrict__ hlist, uint32_t hn, const int32_t *__restrict__ a_len,
                           const uint32_t *__restrict__ P, uint32_t *__restrict__ pre)
{
    const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h > hn) return;
    pre[2 * h] = h < hn ? (uint32_t)a_len[hlist[h]] : 0u;
    pre[2 * h + 1] = h < hn ? P[hlist[h]] : 0u;
}

struct EA {
    uint32_t h;
    uint32_t bst;
    float va;
};

// One record per A entry of every hub row (flat over hpre[2h]).
__global__ __launch_bounds__(TB) void k_sg_hentries(const uint32_t *__restrict__ hlist, uint32_t hn,
                                                    const uint32_t *__restrict__ hpre, uint32_t total,
                                                    const uint32_t *__restrict__ a_start,
                                                    const int32_t *__restrict__ a_col, const float *__restrict__ a_val,
                                                    const uint32_t *__restrict__ b_start,
                                                    const int32_t *__restrict__ b_len, EA *__restrict__ ea,
                                                    uint32_t *__restrict__ blen)
{
    __shared__ uint32_t s_off[TB / 64][65];
    const int w = threadIdx.x >> 6;
    const uint64_t wave = (uint64_t)blockIdx.x * (TB / 64) + w;
    const uint32_t per = 256;
    const uint64_t e0 = wave * per;
    if (e0 >= total) return;
    const uint64_t e1 = e0 + per < total ? e0 + per : total;
    wave_visit<2>(hpre, hn, (uint32_t)e0, (uint32_t)e1, s_off[w], [&](bool valid, uint32_t h, uint32_t k, uint32_t e) {
        if (!valid) return;
        const uint64_t ap = (uint64_t)a_start[hlist[h]] + k;
        const int32_t j = a_col[ap];
        EA r;
        r.h = h;
        r.bst = b_start[j];
        r.va = a_val[ap];
        ea[e] = r;
        blen[e] = (uint32_t)b_len[j];
    });
}

// Partials of hub rows in generation order: key = (h, col), value a*b.
__global__ __launch_bounds__(TB) void k_sg_hexpand(const uint32_t *__restrict__ apos, uint32_t nent, uint32_t total,
                                                   const EA *__restrict__ ea, const int32_t *__restrict__ b_col,
                                                   const float *__restrict__ b_val, int cbits,
                                                   uint64_t *__restrict__ key, uint32_t *__restrict__ val)
{
    __shared__ uint32_t s_off[TB / 64][65];
    const int w = threadIdx.x >> 6;
    const uint64_t wave = (uint64_t)blockIdx.x * (TB / 64) + w;
    const uint32_t per = 256;
    const uint64_t e0 = wave * per;
    if (e0 >= total) return;
    const uint64_t e1 = e0 + per < total ? e0 + per : total;
    wave_visit(apos, nent, (uint32_t)e0, (uint32_t)e1, s_off[w], [&](bool valid, uint32_t q, uint32_t k, uint32_t e) {
        if (!valid) return;
        const EA r = ea[q];
        const uint64_t bp = (uint64_t)r.bst + k;
        key[e] = ((uint64_t)r.h << cbits) | (uint32_t)b_col[bp];
        val[e] = __float_as_uint(r.va * b_val[bp]);
    });
}

// Run heads of the sorted hub partials: left-to-right sums, emit flags.
__global__ void k_sg_hreduce(const uint64_t *__restrict__ key, const uint32_t *__restrict__ val, uint32_t n,
                             uint32_t *__restrict__ emit, float *__restrict__ sum)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > n) return;
    uint32_t e = 0;
    if (p < n) {
        const uint64_t k = key[p];
        if (p == 0 || key[p - 1] != k) {
            float s = 0.0f + __uint_as_float(val[p]);
            for (uint32_t q = p + 1; q < n && key[q] == k; ++q) s = s + __uint_as_float(val[q]);
            e = s != 0.0f ? 1u : 0u;
            sum[p] = s;
        }
    }
    emit[p] = e;
}

__global__ void k_sg_hwrite(const uint64_t *__restrict__ key, uint32_t n, int cbits, const uint32_t *__restrict__ emit,
                            const uint32_t *__restrict__ epos, const float *__restrict__ sum,
                            const uint32_t *__restrict__ hlist, const uint32_t *__restrict__ hpre,
                            const uint32_t *__restrict__ Poff, int32_t *__restrict__ tcol, float *__restrict__ tval)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !emit[p]) return;
    const uint64_t k = key[p];
    const uint32_t h = (uint32_t)(k >> cbits);
    const uint32_t first = epos[hpre[2 * h + 1]];  // emits before this hub row
    const uint64_t o = (uint64_t)Poff[hlist[h]] + (epos[p] - first);
    tcol[o] = (int32_t)(k & ((1ull << cbits) - 1ull));
    tval[o] = sum[p];
}

__global__ void k_sg_hcount(const uint32_t *__restrict__ hlist, uint32_t hn, const uint32_t *__restrict__ hpre,
                            const uint32_t *__restrict__ epos, uint32_t *__restrict__ cnt)
{
    const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= hn) return;
    cnt[hlist[h]] = epos[hpre[2 * (h + 1) + 1]] - epos[hpre[2 * h + 1]];
}

// C row table from the output counts.
// Row r of C: chunk [Poff[r], Poff[r] + P[r]) holding cnt[r] entries.
__global__ void k_sg_set_rows(uint32_t R, const uint32_t *__restrict__ Poff, const uint32_t *__restrict__ P,
                              const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ nnz, uint32_t ptot,
                              uint32_t *__restrict__ start, int32_t *__restrict__ len, int32_t *__restrict__ cap,
                              DevState *ds)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    // a row's chunk keeps its P_r slots at Poff[r]; the P_r - cnt_r it does
    // not fill are counted as garbage, so its capacity is its length (bump
    // = sum of capacities + garbage, as for every other producer)
    if (r < R) {
        start[r] = Poff[r];
        len[r] = (int32_t)cnt[r];
        cap[r] = (int32_t)cnt[r];
    }
    if (r == 0) {
        const uint32_t tot = nnz[0];
        ds->bump = ptot;
        ds->nnz = tot;
        ds->garbage = ptot - tot;
        ds->status &= ST_SCAN_FALLBACK;
    }
}

template <int KPL>
void launch_team(unsigned g, hipStream_t s, const uint32_t *list, const SgState *st, const uint32_t *P,
                 const uint32_t *Poff, dcsr_handle A, dcsr_handle B, int32_t *tcol, float *tval, uint32_t *cnt)
{
    k_sg_team<KPL><<<g, TB, 0, s>>>(list, st, P, Poff, A->row_start, A->row_len, A->pool_col, A->pool_val,
                                    B->row_start, B->row_len, B->pool_col, B->pool_val, tcol, tval, cnt);
}

}  // namespace

dcsr_status dcsr_spgemm(dcsr_handle A, dcsr_handle B, dcsr_handle C, dcsr_stream_t stream)
{
    if (!A || !B || !C || C == A || C == B) return DCSR_ERR_INVALID;
    if (A->cols != B->rows || C->rows != A->rows || C->cols != B->cols) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    A->last_stream = B->last_stream = C->last_stream = s;
    if (!A->state_known) DCSR_TRY(pull_state(A, s));
    if (!B->state_known) DCSR_TRY(pull_state(B, s));
    const uint32_t R = (uint32_t)A->rows;

    uint32_t *P, *Poff, *cnt, *coff, *lists, *part;
    SgState *st;
    unsigned long long *bsum;
    const unsigned nbR = grid_for(R, TB * SG_CR);  // classify workgroups
    DCSR_TRY(ws_get(C, WS_SG_P, (size_t)(R + 1) * 4, (void **)&P));
    DCSR_TRY(ws_get(C, WS_SG_POFF, (size_t)(R + 1) * 4, (void **)&Poff));
    DCSR_TRY(ws_get(C, WS_SG_CNT, (size_t)(R + 1) * 4, (void **)&cnt));
    DCSR_TRY(ws_get(C, WS_SG_COFF, (size_t)(R + 1) * 4, (void **)&coff));
    DCSR_TRY(ws_get(C, WS_SG_LIST, (size_t)(R + 1) * 4 * NCLS, (void **)&lists));
    DCSR_TRY(ws_get(C, WS_SG_STATE, sizeof(SgState), (void **)&st));
    DCSR_TRY(ws_get(C, WS_SG_BSUM, (size_t)nbR * 8, (void **)&bsum));
    DCSR_TRY(ws_get_zeroed(C, WS_PART, scan_part_words(R + 1, 2) * 4, (void **)&part, s));

    // 1. partial counts, classes, offsets
    k_sg_state_zero<<<1, 1, 0, s>>>(st);
    if (R > 0) {
        k_sg_count<<<grid_for(R, TB), TB, 0, s>>>(R, A->row_start, A->row_len, A->pool_col, B->row_len, P);
        k_sg_classify<<<nbR, TB, 0, s>>>(R, P, lists, R + 1, st, cnt, bsum, B->cols <= TEAM_MAX_COLS ? 1u : 0u);
        k_sg_total<<<1, 1024, 0, s>>>(st, bsum, (int)nbR);
    }
    DCSR_HIP(scan_exclusive_u32(P, Poff, R, 1, part, Poff + R, s, fault_word(C)));
    SgState hst;
    DCSR_HIP(hipMemcpyAsync(&hst, st, sizeof(SgState), hipMemcpyDeviceToHost, s));
    DCSR_HIP(hipStreamSynchronize(s));
    if (hst.total >= 0xfffffff0ull) return DCSR_ERR_UNSUPPORTED;
    const uint32_t Ptot = (uint32_t)hst.total;

    // C's rows are written in place in its pool at their partial offsets:
    // row r owns the P_r slots from Poff[r] (its chunk; the slots its merged
    // columns do not fill are garbage), so no copy assembles C afterwards
    if ((int64_t)Ptot > C->pool_cap) {
        const int64_t want = (int64_t)(Ptot * C->cfg.pool_factor);
        DCSR_TRY(alloc_pool(C, want > (int64_t)Ptot ? want : (int64_t)Ptot));
    }
    int32_t *tcol = C->pool_col;
    float *tval = C->pool_val;

    // 2. numeric phases
    for (int c = CL_TEAM; c < CL_TEAM + NTEAMCL; ++c) {
        if (!hst.count[c]) continue;
        const uint32_t per = TB / TEAM;
        const uint32_t want = (hst.count[c] + per - 1) / per;
        const unsigned g = want < 8192 ? want : 8192;
        const uint32_t *list = lists + c * (uint64_t)(R + 1);
        switch (c - CL_TEAM) {
        case 0: launch_team<2>(g, s, list, st, P, Poff, A, B, tcol, tval, cnt); break;
        case 1: launch_team<4>(g, s, list, st, P, Poff, A, B, tcol, tval, cnt); break;
        case 2: launch_team<6>(g, s, list, st, P, Poff, A, B, tcol, tval, cnt); break;
        default: launch_team<8>(g, s, list, st, P, Poff, A, B, tcol, tval, cnt); break;
        }
    }
    if (hst.count[CL_WAVE]) {
        const unsigned g = (unsigned)((hst.count[CL_WAVE] + 3) / 4 < 8192 ? (hst.count[CL_WAVE] + 3) / 4 : 8192);
        if (B->cols <= WAVE_MAX_COLS32)
            k_sg_wave<uint32_t><<<g, TB, 0, s>>>(lists + CL_WAVE * (uint64_t)(R + 1), st, P, Poff, A->row_start,
                                                 A->row_len, A->pool_col, A->pool_val, B->row_start, B->row_len,
                                                 B->pool_col, B->pool_val, tcol, tval, cnt);
        else
            k_sg_wave<uint64_t><<<g, TB, 0, s>>>(lists + CL_WAVE * (uint64_t)(R + 1), st, P, Poff, A->row_start,
                                                 A->row_len, A->pool_col, A->pool_val, B->row_start, B->row_len,
                                                 B->pool_col, B->pool_val, tcol, tval, cnt);
    }
    if (hst.count[CL_BLOCK]) {
        const unsigned g = hst.count[CL_BLOCK] < 4096 ? hst.count[CL_BLOCK] : 4096;
        k_sg_block<<<g, TB, 0, s>>>(lists + CL_BLOCK * (uint64_t)(R + 1), st, P, Poff, A->row_start, A->row_len,
                                    A->pool_col, A->pool_val, B->row_start, B->row_len, B->pool_col, B->pool_val, tcol,
                                    tval, cnt);
    }
    if (hst.count[CL_HUB]) {
        const uint32_t hn = hst.count[CL_HUB];
        const uint32_t *hlist = lists + CL_HUB * (uint64_t)(R + 1);
        uint32_t *hpre_in, *hpre, *blen, *apos, *emit, *epos, *v0, *v1;
        uint64_t *k0, *k1;
        EA *ea;
        DCSR_TRY(ws_get(C, WS_SG_HPRE, (size_t)(hn + 1) * 16, (void **)&hpre_in));
        hpre = hpre_in + 2 * (hn + 1);
        k_sg_hprep<<<grid_for(hn + 1, TB), TB, 0, s>>>(hlist, hn, A->row_len, P, hpre_in);
        uint32_t htot[2];
        uint32_t *htot_d;
        DCSR_TRY(ws_get(C, WS_TMP2, 64, (void **)&htot_d));
        DCSR_HIP(scan_exclusive_u32(hpre_in, hpre, hn + 1, 2, part, htot_d, s, fault_word(C)));
        DCSR_HIP(hipMemcpyAsync(htot, htot_d, 8, hipMemcpyDeviceToHost, s));
        DCSR_HIP(hipStreamSynchronize(s));
        const uint32_t nent = htot[0], np = htot[1];
        DCSR_TRY(ws_get(C, WS_SG_EA, (size_t)(nent + 1) * sizeof(EA), (void **)&ea));
        DCSR_TRY(ws_get(C, WS_SG_APOS, (size_t)(nent + 2) * 8, (void **)&blen));
        apos = blen + (nent + 1);
        // hpre (the exclusive prefix), not hpre_in (the lengths): the segment
        // search needs monotone offsets
        k_sg_hentries<<<grid_for(nent, 4 * 256) + 1, TB, 0, s>>>(hlist, hn, hpre, nent, A->row_start, A->pool_col,
                                                               A->pool_val, B->row_start, B->row_len, ea, blen);
        size_t pw = scan_part_words(nent + 1, 1);
        if (scan_part_words((int64_t)np + 1, 1) > pw) pw = scan_part_words((int64_t)np + 1, 1);
        if (radix_part_words(np) > pw) pw = radix_part_words(np);
        DCSR_TRY(ws_get_zeroed(C, WS_SCAN, pw * 4, (void **)&part, s));
        DCSR_HIP(scan_exclusive_u32(blen, apos, nent, 1, part, apos + nent, s, fault_word(C)));
        DCSR_TRY(ws_get(C, WS_SG_KEY0, (size_t)(np + 1) * 8, (void **)&k0));
        DCSR_TRY(ws_get(C, WS_SG_KEY1, (size_t)(np + 1) * 8, (void **)&k1));
        DCSR_TRY(ws_get(C, WS_SG_VAL0, (size_t)(np + 1) * 4, (void **)&v0));
        DCSR_TRY(ws_get(C, WS_SG_VAL1, (size_t)(np + 1) * 4, (void **)&v1));
        k_sg_hexpand<<<grid_for(np, 4 * 256) + 1, TB, 0, s>>>(apos, nent, np, ea, B->pool_col, B->pool_val, C->cbits,
                                                              k0, v0);
        uint32_t *hist;
        DCSR_TRY(ws_get(C, WS_HIST, radix_hist_words(np) * 4, (void **)&hist));
        int which = 0;
        DCSR_HIP(radix_sort_pairs(k0, v0, k1, v1, np, bits_for(hn) + C->cbits, hist, part, &which, s));
        uint64_t *key = which ? k1 : k0;
        uint32_t *val = which ? v1 : v0;
        float *sum;
        DCSR_TRY(ws_get(C, WS_SG_FLAG, (size_t)(np + 2) * 8, (void **)&emit));
        sum = (float *)(emit + (np + 2));
        DCSR_TRY(ws_get(C, WS_SG_EPOS, (size_t)(np + 2) * 4, (void **)&epos));
        k_sg_hreduce<<<grid_for(np + 1, TB), TB, 0, s>>>(key, val, np, emit, sum);
        DCSR_HIP(scan_exclusive_u32(emit, epos, np + 1, 1, part, nullptr, s, fault_word(C)));
        k_sg_hwrite<<<grid_for(np, TB), TB, 0, s>>>(key, np, C->cbits, emit, epos, sum, hlist, hpre, Poff, tcol, tval);
        k_sg_hcount<<<grid_for(hn, TB), TB, 0, s>>>(hlist, hn, hpre, epos, cnt);
    }
    DCSR_HIP(hipGetLastError());

    // 3. C's row table: chunks at the partial offsets; nnz from a scan of
    // the row counts
    DCSR_TRY(ws_get_zeroed(C, WS_PART, scan_part_words(R + 1, 1) * 4, (void **)&part, s));
    DCSR_HIP(scan_exclusive_u32(cnt, coff, R, 1, part, coff + R, s, fault_word(C)));
    k_sg_set_rows<<<grid_for(R + 1, TB), TB, 0, s>>>(R, Poff, P, cnt, coff + R, Ptot, C->row_start, C->row_len,
                                                     C->row_cap, C->ds);
    DCSR_HIP(hipGetLastError());
    DCSR_TRY(pull_state(C, s));
    C->canonical = true;
    C->flat_dirty = true;
    C->hubs_dirty = true;
    return DCSR_OK;
}

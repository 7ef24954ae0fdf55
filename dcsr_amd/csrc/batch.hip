// batch.hip -- many small DCSR matrices updated and multiplied in one
// persistent launch.
//
// No reference counterpart: the reference handles one matrix at a time with
// a process-global bin table (include/Matrix_Test.h:57, include/load.h:606).
// The matrices of a batch live block-diagonally in one DCSR handle (matrix m
// owns global rows [row_base[m], row_base[m+1]) and keeps local column
// indices), so they share one pool and one allocator.  One persistent kernel
// does a whole step: every workgroup takes matrix ids from a work counter and,
// for each, applies that matrix's insert-or-add batch entirely in LDS and
// per-workgroup scratch (stable LDS radix sort of (row, col) keys, left-to-
// right duplicate sums, a per-row merge against the stored row for the
// codes, prefix-sum chunk claims, a per-row merge into the new chunk,
// commit), then
// computes that matrix's SpMV with TEAM lanes per row and a whole wave for
// rows longer than a threshold.  Semantics per matrix are those of
// dcsr_update followed by dcsr_spmv.  The last workgroup to finish resets the
// counter, so the launch can be captured in a hipGraph and replayed.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>
#include <map>
#include <vector>
#include "dcsr_internal.h"
#include "prims.h"
#include "wave.h"

namespace {

constexpr int BT = 256;
constexpr uint32_t BMAX = 4096;  // triples per matrix per step
constexpr int KB = 12;           // bits of a local row / column / arrival index
constexpr uint32_t LMAX = 1u << KB;
constexpr int ITEMS = 17;        // scan items per thread (256 * 17 >= BMAX + 1)
enum { C_NOP = 0, C_ADD = 1, C_DEL = 2, C_INS = 3 };

// per-workgroup scratch (words)
enum {
    S_URC = 0, S_UVAL = S_URC + BMAX, S_UPOS = S_UVAL + BMAX, S_UCODE = S_UPOS + BMAX, S_UT = S_UCODE + BMAX,
    S_UIP = S_UT + BMAX, S_UDP = S_UIP + BMAX + 1, S_TROW = S_UDP + BMAX + 1, S_TUB = S_TROW + BMAX,
    S_TINS = S_TUB + BMAX + 1, S_TDEL = S_TINS + BMAX, S_TDST = S_TDEL + BMAX, S_TMP = S_TDST + BMAX + 1,
    S_TMODE = S_TMP + BMAX + 1, S_TNL = S_TMODE + BMAX, S_HF = S_TNL + BMAX, S_RF = S_HF + BMAX + 1,
    S_WORDS = S_RF + BMAX + 1
};

struct BatchState {
    unsigned int work;  // next matrix id
    unsigned int done;  // workgroups finished
};

#define BST_BAD_INDEX 0x1u
#define BST_TOO_BIG 0x2u
#define BST_EXHAUSTED 0x4u

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

// In-place exclusive scan of a[0..n) (n <= 256 * ITEMS) by one workgroup;
// returns the total.  Blocked: thread t owns [t*ITEMS, t*ITEMS+ITEMS).
__device__ uint32_t block_scan_inplace(uint32_t *a, uint32_t n, uint32_t *lw)
{
    const uint32_t t = threadIdx.x;
    uint32_t loc[ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t idx = t * ITEMS + i;
        const uint32_t v = idx < n ? a[idx] : 0u;
        loc[i] = s;
        s += v;
    }
    const uint32_t inc = wave_incl_scan_u32(s);
    if (lane_id() == 63) lw[t >> 6] = inc;
    __syncthreads();
    uint32_t carry = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < BT / 64; ++i) {
        if (i < (int)(t >> 6)) carry += lw[i];
        tot += lw[i];
    }
    const uint32_t base = carry + inc - s;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t idx = t * ITEMS + i;
        if (idx < n) a[idx] = base + loc[i];
    }
    __syncthreads();
    return tot;
}

// Largest j in [0, n) with a[j] <= e (a non-decreasing, a[0] <= e).
__device__ __forceinline__ uint32_t find_last_le(const uint32_t *a, uint32_t n, uint32_t e)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] <= e) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t *lw)
{
    const uint32_t w = wave_sum_u32(v);
    if (lane_id() == 0) lw[threadIdx.x >> 6] = w;
    __syncthreads();
    const uint32_t t = (lw[0] + lw[1]) + (lw[2] + lw[3]);
    __syncthreads();
    return t;
}

// Block-wide exclusive scan of one value per thread (256 threads); *total
// receives the sum.
__device__ __forceinline__ uint32_t block_excl(uint32_t v, uint32_t *lw, uint32_t *total)
{
    const uint32_t inc = wave_incl_scan_u32(v);
    if (lane_id() == 63) lw[threadIdx.x >> 6] = inc;
    __syncthreads();
    uint32_t carry = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < BT / 64; ++i) {
        if (i < (int)(threadIdx.x >> 6)) carry += lw[i];
        tot += lw[i];
    }
    __syncthreads();
    *total = tot;
    return carry + inc - v;
}

// One stable pass (8-bit digit at `shift`) of an LDS radix sort of n <= BMAX
// (key, value) pairs from (ks, vs) to (kd, vd).  Wave w ranks its contiguous
// quarter in order (ballot digit matching, running per-digit counts in
// wcnt[w]); digit bases come from a block scan of the digit totals.
__device__ void lds_radix_pass(const uint32_t *ks, const float *vs, uint32_t *kd, float *vd, uint32_t n, int shift,
                               uint32_t (*wcnt)[256], uint32_t *lw)
{
    constexpr int MAXJ = BMAX / (BT / 64) / 64;  // 64-item chunks per wave
    const int w = threadIdx.x >> 6;
    const int lane = lane_id();
    const uint32_t Q = (((n + 3) / 4) + 63) & ~63u;
    const uint32_t wb = (uint32_t)w * Q;
    const uint32_t we = wb + Q < n ? wb + Q : n;
    for (int i = threadIdx.x; i < (BT / 64) * 256; i += BT) (&wcnt[0][0])[i] = 0;
    __syncthreads();
    uint32_t key[MAXJ], rnk[MAXJ];
    float val[MAXJ];
    const uint64_t lt = lanemask_lt();
    // every item read before the ranking writes the counters (the buffers
    // arrive as generic pointers, so the compiler may not move a read past
    // a counter write itself); indices stay below 4 * Q <= BMAX
#pragma unroll
    for (int j = 0; j < MAXJ; ++j) {
        const uint32_t idx = wb + (uint32_t)j * 64 + lane;
        key[j] = ks[idx];
        val[j] = vs[idx];
    }
#pragma unroll
    for (int j = 0; j < MAXJ; ++j) {
        const uint32_t idx = wb + (uint32_t)j * 64 + lane;
        const bool ok = idx < we;
        if (!ok) {
            key[j] = 0u;
            val[j] = 0.0f;
        }
        rnk[j] = 0;
        if (wb + (uint32_t)j * 64 < we) {  // wave-uniform
            const uint32_t d = (key[j] >> shift) & 255u;
            const uint64_t act = __ballot(ok);
            const uint64_t peers = wave_match_digit8(d, act);
            const int leader = __ffsll((long long)peers) - 1;
            // every lane of the digit reads its count before the leader adds
            // to it (a wave's LDS accesses complete in order): no broadcast
            const uint32_t b = wcnt[w][d];
            if (ok && lane == leader) wcnt[w][d] = b + (uint32_t)__popcll(peers);
            rnk[j] = b + (uint32_t)__popcll(peers & lt);
        }
    }
    __syncthreads();
    {
        const int d = threadIdx.x;
        uint32_t c[BT / 64];
        uint32_t tot = 0;
#pragma unroll
        for (int i = 0; i < BT / 64; ++i) {
            c[i] = wcnt[i][d];
            tot += c[i];
        }
        uint32_t all;
        uint32_t b = block_excl(tot, lw, &all);
#pragma unroll
        for (int i = 0; i < BT / 64; ++i) {
            wcnt[i][d] = b;
            b += c[i];
        }
    }
    __syncthreads();
    // every digit base read before the first write (generic pointers again)
    uint32_t pos[MAXJ];
#pragma unroll
    for (int j = 0; j < MAXJ; ++j) pos[j] = wcnt[w][(key[j] >> shift) & 255u] + rnk[j];
#pragma unroll
    for (int j = 0; j < MAXJ; ++j) {
        const uint32_t idx = wb + (uint32_t)j * 64 + lane;
        if (idx < we) {
            kd[pos[j]] = key[j];
            vd[pos[j]] = val[j];
        }
    }
    __syncthreads();
}

// Counters of the matrices one workgroup updated, kept by its thread 0 and
// added to the device state once when the workgroup is done (one atomic per
// counter and workgroup instead of per matrix).
struct BatchAcc {
    long long dnnz, dgarbage;
    unsigned long long ins, add, del;
    uint32_t touched;
};

// Update of matrix m inside one workgroup.  Returns nothing; status bits go
// to mstat[m].  reserved: the host has made room in the pool for every
// matrix's largest claim, so claims take a plain atomic add; otherwise a
// compare-and-swap that advances the bump pointer only when the claim fits.
__device__ void batch_update(uint32_t m, const uint32_t *__restrict__ row_base, const uint32_t *__restrict__ col_base,
                             const int64_t *__restrict__ tri_off, int64_t ntri, const int32_t *__restrict__ trow,
                             const int32_t *__restrict__ tcol, const float *__restrict__ tval,
                             uint32_t *__restrict__ row_start, int32_t *__restrict__ row_len,
                             int32_t *__restrict__ row_cap, int32_t *__restrict__ pcol, float *__restrict__ pval,
                             unsigned long long pool_cap, DevState *ds, uint32_t *__restrict__ mstat, uint32_t *sc,
                             uint32_t (*skb)[BMAX], float (*svb)[BMAX], uint32_t (*wcnt)[256], uint32_t *lw,
                             uint32_t *sm, bool reserved, BatchAcc &acc)
{
    const uint32_t t = threadIdx.x;
    const uint32_t gbase = row_base[m];
    const uint32_t Rm = row_base[m + 1] - gbase;
    const uint32_t Cm = col_base[m + 1] - col_base[m];
    const int64_t t0 = tri_off[m];
    const int64_t t1 = tri_off[m + 1];
    if (t0 < 0 || t1 < t0 || t1 > ntri) {  // offsets outside the triple arrays
        if (t == 0) mstat[m] |= BST_BAD_INDEX;
        return;
    }
    const int64_t nb64 = t1 - t0;
    if (nb64 == 0) return;
    if (nb64 > (int64_t)BMAX) {
        if (t == 0) mstat[m] |= BST_TOO_BIG;
        return;
    }
    const uint32_t nb = (uint32_t)nb64;
    // U1: (row, col) keys and values into LDS, stable radix sort (3 passes
    // of 8 bits over the 24 key bits; stability keeps arrival order among
    // duplicates)
    if (t == 0) sm[0] = 0;
    __syncthreads();
    uint32_t bad = 0;
    {
        // a thread's triples (at most BMAX / BT) are all loaded before the
        // first LDS store; indices clamp to the matrix's last triple
        constexpr int TPT = BMAX / BT;
        int32_t rr[TPT], cc[TPT];
        float vv[TPT];
#pragma unroll
        for (int j = 0; j < TPT; ++j) {
            const uint32_t i = t + (uint32_t)j * BT;
            const int64_t gi = t0 + (i < nb ? i : nb - 1);
            rr[j] = trow[gi];
            cc[j] = tcol[gi];
            vv[j] = tval[gi];
        }
#pragma unroll
        for (int j = 0; j < TPT; ++j) {
            const uint32_t i = t + (uint32_t)j * BT;
            if (i < nb) {
                const int32_t r = rr[j], c = cc[j];
                uint32_t k = 0;
                if (r < 0 || (uint32_t)r >= Rm || c < 0 || (uint32_t)c >= Cm) bad = 1;
                else k = ((uint32_t)r << KB) | (uint32_t)c;
                skb[0][i] = k;
                svb[0][i] = vv[j];
            }
        }
    }
    if (block_sum(bad, lw)) {
        if (t == 0) mstat[m] |= BST_BAD_INDEX;
        return;
    }
    lds_radix_pass(skb[0], svb[0], skb[1], svb[1], nb, 0, wcnt, lw);
    lds_radix_pass(skb[1], svb[1], skb[0], svb[0], nb, 8, wcnt, lw);
    lds_radix_pass(skb[0], svb[0], skb[1], svb[1], nb, 16, wcnt, lw);
    const uint32_t *sk = skb[1];
    const float *sv = svb[1];
    // U2: distinct coordinates (heads) and touched rows (row heads)
    uint32_t *hf = sc + S_HF, *rf = sc + S_RF;
    for (uint32_t p = t; p <= nb; p += BT) {
        uint32_t h = 0, rh = 0;
        if (p < nb) {
            const uint32_t k = sk[p];
            h = (p == 0 || sk[p - 1] != k) ? 1u : 0u;
            rh = (p == 0 || (sk[p - 1] >> KB) != (k >> KB)) ? 1u : 0u;
        }
        hf[p] = h;
        rf[p] = rh;
    }
    __syncthreads();
    const uint32_t nu = block_scan_inplace(hf, nb + 1, lw);
    const uint32_t nt = block_scan_inplace(rf, nb + 1, lw);
    uint32_t *urc = sc + S_URC, *upos = sc + S_UPOS, *ucode = sc + S_UCODE, *ut = sc + S_UT;
    float *uval = (float *)(sc + S_UVAL);
    uint32_t *trw = sc + S_TROW, *tub = sc + S_TUB, *tins = sc + S_TINS, *tdel = sc + S_TDEL;
    for (uint32_t p = t; p < nb; p += BT) {
        const uint32_t k = sk[p];
        const bool h = p == 0 || sk[p - 1] != k;
        if (!h) continue;
        const bool rh = p == 0 || (sk[p - 1] >> KB) != (k >> KB);
        const uint32_t u = hf[p];
        const uint32_t tt = rf[p] + (rh ? 0u : 0xffffffffu);
        float s = 0.0f + sv[p];
        for (uint32_t q = p + 1; q < nb && sk[q] == k; ++q) s = s + sv[q];
        urc[u] = k;
        uval[u] = s;
        ut[u] = tt;
        if (rh) {
            trw[tt] = k >> KB;
            tub[tt] = u;
            tins[tt] = 0;
            tdel[tt] = 0;
        }
    }
    if (t == 0) tub[nt] = nu;
    __syncthreads();
    // U3: one lane per touched row merges the row's update run (sorted by
    // column) against its old entries (sorted by column): code and new value
    // per distinct coordinate, position of a hit, inserts / deletes per row.
    uint32_t my_adds = 0, my_ins = 0, my_del = 0;
    for (uint32_t tt = t; tt < nt; tt += BT) {
        const uint32_t g = gbase + trw[tt];
        const uint64_t st = row_start[g];
        const int32_t L = row_len[g];
        const uint32_t ub = tub[tt], ue = tub[tt + 1];
        int32_t i = 0;
        uint32_t ni = 0, nd = 0;
        for (uint32_t u = ub; u < ue; ++u) {
            const int32_t c = (int32_t)(urc[u] & (LMAX - 1));
            // advance past the old columns below c, eight per step: the
            // step's loads are issued together (indices clamp into the row)
            // and counted with unrolled compares; the first column not below
            // c is picked the same way
            int32_t at = -1;  // old column at position i (-1: i == L)
            while (i < L) {
                int32_t pc[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) pc[j] = pcol[st + (uint64_t)(i + j < L ? i + j : L - 1)];
                int32_t below = 0;
#pragma unroll
                for (int j = 0; j < 8; ++j) below += (i + j < L && pc[j] < c) ? 1 : 0;
                at = -1;
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (j == below && i + j < L) at = pc[j];
                i += below;
                if (below < 8) break;
                at = -1;
            }
            const bool hit = i < L && at == c;
            const float s = uval[u];
            uint32_t code;
            float nv;
            if (hit) {
                nv = pval[st + i] + s;
                code = nv != 0.0f ? C_ADD : C_DEL;
            } else {
                nv = s;
                code = s != 0.0f ? C_INS : C_NOP;
            }
            uval[u] = nv;
            upos[u] = (uint32_t)i;
            ucode[u] = code;
            ni += code == C_INS;
            nd += code == C_DEL;
            my_adds += code == C_ADD;
        }
        tins[tt] = ni;
        tdel[tt] = nd;
        my_ins += ni;
        my_del += nd;
    }
    __syncthreads();
    const uint32_t nadd = block_sum(my_adds, lw);
    const uint32_t nins = block_sum(my_ins, lw);
    const uint32_t ndel = block_sum(my_del, lw);
    // U4: plan, claim
    uint32_t *tdst = sc + S_TDST, *tmode = sc + S_TMODE, *tnl = sc + S_TNL;
    long long dnnz = 0, dgar = 0;
    for (uint32_t tt = t; tt <= nt; tt += BT) {
        uint32_t claim = 0;
        if (tt < nt) {
            const uint32_t g = gbase + trw[tt];
            const uint32_t L = (uint32_t)row_len[g];
            const uint32_t ni = tins[tt], nd = tdel[tt];
            const uint32_t nl = L + ni - nd;
            const uint32_t mode = (ni | nd) ? 1u : 0u;
            if (mode) {
                claim = chunk_round(nl);
                dnnz += (long long)nl - (long long)L;
                dgar += row_cap[g];
            }
            tmode[tt] = mode;
            tnl[tt] = nl;
        }
        tdst[tt] = claim;
    }
    __syncthreads();
    const uint32_t claim_tot = block_scan_inplace(tdst, nt + 1, lw);
    // 64-bit deltas: per-thread values are small; reduce through 32-bit halves
    const uint32_t dn_lo = block_sum((uint32_t)(dnnz & 0xffffffff), lw);
    const uint32_t dg_lo = block_sum((uint32_t)(dgar & 0xffffffff), lw);
    if (t == 0) {
        // claim by compare-and-swap: the bump pointer advances only when the
        // claim fits, so a matrix that does not fit leaves room for later
        // (smaller) matrices of the same step
        unsigned long long base = 0ull;
        bool ok = true;
        if (claim_tot && reserved) {
            base = atomicAdd(&ds->bump, (unsigned long long)claim_tot);
        } else if (claim_tot) {
            unsigned long long cur = __hip_atomic_load(&ds->bump, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            for (;;) {
                if (cur + claim_tot > pool_cap) {
                    ok = false;
                    break;
                }
                const unsigned long long seen = atomicCAS(&ds->bump, cur, cur + claim_tot);
                if (seen == cur) {
                    base = cur;
                    break;
                }
                cur = seen;  // another matrix claimed first: retry from its end
            }
        }
        sm[1] = ok ? 1u : 0u;
        sm[2] = (uint32_t)base;
        if (!ok) mstat[m] |= BST_EXHAUSTED;
        if (ok) {
            acc.dnnz += (long long)(int32_t)dn_lo;
            acc.dgarbage += (long long)(int32_t)dg_lo;
            acc.ins += nins;
            acc.add += nadd;
            acc.del += ndel;
            acc.touched += nt;
        }
    }
    __syncthreads();
    if (!sm[1]) return;
    const uint32_t base = sm[2];
    // U5: one lane per touched row: rewritten rows are merged into their new
    // chunk (old entries and the row's update run, both column-sorted);
    // unchanged-structure rows get their added values written in place
    for (uint32_t tt = t; tt < nt; tt += BT) {
        const uint32_t g = gbase + trw[tt];
        const uint64_t st = row_start[g];
        const uint32_t ub = tub[tt], ue = tub[tt + 1];
        if (!tmode[tt]) {
            for (uint32_t u = ub; u < ue; ++u)
                if (ucode[u] == C_ADD) pval[st + upos[u]] = uval[u];
            continue;
        }
        const int32_t L = row_len[g];
        const uint64_t d0 = (uint64_t)base + tdst[tt];
        uint64_t o = d0;
        int32_t i = 0;
        // old entries [i, to) to o.., eight at a time with all eight loads
        // issued before their stores (source and destination share the
        // pool arrays, so the compiler keeps each load after the stores
        // before it); to <= L, so a clamped index stays in the row
        auto copy_run = [&](int32_t to) {
            for (; i < to; i += 8, o += 8) {
                int32_t cc[8];
                float vv[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const uint64_t si = st + (uint64_t)(i + j < to ? i + j : to - 1);
                    cc[j] = pcol[si];
                    vv[j] = pval[si];
                }
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    if (i + j < to) {
                        pcol[o + j] = cc[j];
                        pval[o + j] = vv[j];
                    }
                }
            }
            o -= (uint64_t)(i - to);  // the last chunk stopped at to
            i = to;
        };
        for (uint32_t u = ub; u < ue; ++u) {
            const uint32_t k = upos[u];  // old entries before this coordinate
            if (i < (int32_t)k) copy_run((int32_t)k);
            const uint32_t code = ucode[u];
            if (code == C_INS) {
                pcol[o] = (int32_t)(urc[u] & (LMAX - 1));
                pval[o] = uval[u];
                ++o;
            } else if (code == C_ADD) {
                pcol[o] = pcol[st + i];
                pval[o] = uval[u];
                ++o;
                ++i;
            } else if (code == C_DEL) {
                ++i;
            }
        }
        if (i < L) copy_run(L);
    }
    __syncthreads();
    for (uint32_t tt = t; tt < nt; tt += BT) {
        if (!tmode[tt]) continue;
        const uint32_t g = gbase + trw[tt];
        row_start[g] = base + tdst[tt];
        row_len[g] = (int32_t)tnl[tt];
        row_cap[g] = (int32_t)chunk_round(tnl[tt]);
    }
    __syncthreads();
}

// SpMV of matrix m: TEAM lanes per row for rows up to `thresh` entries, one
// wave per longer row; fixed-order reductions.
template <int TEAM>
__device__ void batch_spmv(uint32_t m, const uint32_t *__restrict__ row_base, const uint32_t *__restrict__ col_base,
                           const uint32_t *__restrict__ row_start, const int32_t *__restrict__ row_len,
                           const int32_t *__restrict__ pcol, const float *__restrict__ pval,
                           const float *__restrict__ x, float *__restrict__ y, uint32_t thresh)
{
    const uint32_t gbase = row_base[m];
    const uint32_t Rm = row_base[m + 1] - gbase;
    const float *xm = x + col_base[m];
    const uint32_t t = threadIdx.x;
    const uint32_t tl = t % TEAM;
    for (uint32_t r0 = 0; r0 < Rm; r0 += BT / TEAM) {
        const uint32_t r = r0 + t / TEAM;
        float acc = 0.0f;
        bool mine = false;
        if (r < Rm) {
            const uint32_t L = (uint32_t)row_len[gbase + r];
            mine = L <= thresh;
            if (mine) {
                const uint64_t st = row_start[gbase + r];
                uint32_t k = tl;
                // four entries of the lane in flight; the lane still sums
                // its entries in ascending order
                for (; k + 3 * TEAM < L; k += 4 * TEAM) {
                    const int32_t c0 = pcol[st + k], c1 = pcol[st + k + TEAM], c2 = pcol[st + k + 2 * TEAM],
                                  c3 = pcol[st + k + 3 * TEAM];
                    const float v0 = pval[st + k], v1 = pval[st + k + TEAM], v2 = pval[st + k + 2 * TEAM],
                                v3 = pval[st + k + 3 * TEAM];
                    const float x0 = xm[c0], x1 = xm[c1], x2 = xm[c2], x3 = xm[c3];
                    acc = acc + v0 * x0;
                    acc = acc + v1 * x1;
                    acc = acc + v2 * x2;
                    acc = acc + v3 * x3;
                }
                for (; k < L; k += TEAM) acc = acc + pval[st + k] * xm[pcol[st + k]];
            }
        }
#pragma unroll
        for (int o = TEAM / 2; o >= 1; o >>= 1) acc += __shfl_xor(acc, o);
        if (mine && tl == 0) y[gbase + r] = acc;
    }
    const int w = t >> 6;
    for (uint32_t r = w; r < Rm; r += BT / 64) {
        const uint32_t L = (uint32_t)row_len[gbase + r];
        if (L <= thresh) continue;
        const uint64_t st = row_start[gbase + r];
        float acc = 0.0f;
        uint32_t k = lane_id();
        for (; k + 3 * 64 < L; k += 4 * 64) {
            const int32_t c0 = pcol[st + k], c1 = pcol[st + k + 64], c2 = pcol[st + k + 128], c3 = pcol[st + k + 192];
            const float v0 = pval[st + k], v1 = pval[st + k + 64], v2 = pval[st + k + 128], v3 = pval[st + k + 192];
            const float x0 = xm[c0], x1 = xm[c1], x2 = xm[c2], x3 = xm[c3];
            acc = acc + v0 * x0;
            acc = acc + v1 * x1;
            acc = acc + v2 * x2;
            acc = acc + v3 * x3;
        }
        for (; k < L; k += 64) acc = acc + pval[st + k] * xm[pcol[st + k]];
        acc = wave_sum_f32(acc);
        if (lane_id() == 0) y[gbase + r] = acc;
    }
}

template <int TEAM>
__global__ __launch_bounds__(BT) void k_batch_step(uint32_t M, const uint32_t *__restrict__ row_base,
                                                   const uint32_t *__restrict__ col_base,
                                                   const int64_t *__restrict__ tri_off, int64_t ntri,
                                                   const int32_t *__restrict__ trow, const int32_t *__restrict__ tcol,
                                                   const float *__restrict__ tval,
                                                   uint32_t *__restrict__ row_start, int32_t *__restrict__ row_len,
                                                   int32_t *__restrict__ row_cap, int32_t *__restrict__ pcol,
                                                   float *__restrict__ pval, unsigned long long pool_cap, DevState *ds,
                                                   uint32_t *__restrict__ mstat, uint32_t *__restrict__ scratch,
                                                   const float *__restrict__ x, float *__restrict__ y, uint32_t thresh,
                                                   BatchState *bs, int do_update, int do_spmv, int reserved)
{
    __shared__ uint32_t sk[2][BMAX];
    __shared__ float sv[2][BMAX];
    __shared__ uint32_t wcnt[BT / 64][256];
    __shared__ uint32_t lw[BT / 64];
    __shared__ uint32_t sm[4];
    uint32_t *sc = scratch + (uint64_t)blockIdx.x * S_WORDS;
    BatchAcc acc{0, 0, 0, 0, 0, 0};
    for (;;) {
        if (threadIdx.x == 0) sm[3] = atomicAdd(&bs->work, 1u);
        __syncthreads();
        const uint32_t m = sm[3];
        __syncthreads();
        if (m >= M) break;
        if (do_update)
            batch_update(m, row_base, col_base, tri_off, ntri, trow, tcol, tval, row_start, row_len, row_cap, pcol, pval,
                         pool_cap, ds, mstat, sc, sk, sv, wcnt, lw, sm, reserved != 0, acc);
        if (do_spmv) batch_spmv<TEAM>(m, row_base, col_base, row_start, row_len, pcol, pval, x, y, thresh);
        __syncthreads();
    }
    // the workgroup's counters, then: the last workgroup out resets the work
    // counters (no memset: replay-safe).  No fence: both counters are
    // atomics, every workgroup's last work claim returned before its finish
    // count, and the kernel boundary publishes the reset (a release fence
    // would write back the XCD's whole L2 here)
    if (threadIdx.x == 0) {
        if (acc.touched) {
            atomicAdd((unsigned long long *)&ds->nnz, (unsigned long long)acc.dnnz);
            atomicAdd((unsigned long long *)&ds->garbage, (unsigned long long)acc.dgarbage);
            atomicAdd((unsigned long long *)&ds->n_ins, acc.ins);
            atomicAdd((unsigned long long *)&ds->n_add, acc.add);
            atomicAdd((unsigned long long *)&ds->n_del, acc.del);
            atomicAdd(&ds->ntouched, acc.touched);
        }
        if (atomicAdd(&bs->done, 1u) == gridDim.x - 1) {
            bs->work = 0;
            bs->done = 0;
        }
    }
}

__global__ void k_batch_globalize(const int64_t *__restrict__ tri_off, uint32_t M, int64_t n,
                                  const uint32_t *__restrict__ row_base, const uint32_t *__restrict__ col_base,
                                  const int32_t *__restrict__ rows, const int32_t *__restrict__ cols,
                                  int32_t *__restrict__ grows, int32_t *__restrict__ gcols, DevState *ds)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t lo = 0, hi = M;  // matrix of entry i: largest m with tri_off[m] <= i
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tri_off[mid] <= i) lo = mid; else hi = mid;
    }
    const uint32_t Rm = row_base[lo + 1] - row_base[lo], Cm = col_base[lo + 1] - col_base[lo];
    const int32_t r = rows[i], c = cols[i];
    const bool bad = r < 0 || (uint32_t)r >= Rm || c < 0 || (uint32_t)c >= Cm;
    grows[i] = bad ? -1 : (int32_t)(row_base[lo] + r);
    gcols[i] = c;
    if (bad) atomicOr(&ds->status, ST_BAD_INDEX);
}

__global__ void k_batch_zero_stat(uint32_t *mstat, uint32_t M, DevState *ds)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) mstat[i] = 0;
    if (i == 0) {
        ds->n_ins = ds->n_add = ds->n_del = 0;
        ds->ntouched = 0;
    }
}

}  // namespace

dcsr_status load_coo_internal(dcsr_handle h, const int32_t *rows, const int32_t *cols, const float *vals, int64_t n,
                              hipStream_t s);

struct TuneKey {
    int mbucket, lbucket;
    bool operator<(const TuneKey &o) const { return mbucket != o.mbucket ? mbucket < o.mbucket : lbucket < o.lbucket; }
};
struct TuneVal {
    int team;
    int thresh;
};
static std::map<TuneKey, TuneVal> g_tune_cache;  // process-wide autotuner choices

struct dcsr_batch_s {
    uint32_t M;
    dcsr_handle h;  // block-diagonal storage
    uint32_t *row_base, *col_base;  // device, M+1
    std::vector<uint32_t> row_base_h, col_base_h;
    uint32_t *mstat;                // device, per-matrix status bits
    uint32_t *scratch;              // per-workgroup scratch
    BatchState *bs;
    int grid;
    int team, thresh;               // SpMV launch shape (autotuned)
    hipStream_t cs;                 // capture stream (graphs replay on the caller's stream)
    hipGraph_t graph;
    hipGraphExec_t gexec;
    const void *gkey[13];           // launch arguments the captured graph was built with
};

static int bucket_log2(double v)
{
    int b = 0;
    while (v >= 2.0 && b < 30) {
        v /= 2.0;
        ++b;
    }
    return b;
}

dcsr_status dcsr_batch_create(dcsr_batch *out, int32_t count, const int32_t *rows, const int32_t *cols,
                              const dcsr_config *cfg)
{
    if (!out || count <= 0 || !rows || !cols) return DCSR_ERR_INVALID;
    uint64_t R = 0, Cmax = 1;
    std::vector<uint32_t> rb(count + 1), cb(count + 1);
    uint64_t cacc = 0;
    for (int i = 0; i < count; ++i) {
        if (rows[i] < 0 || cols[i] < 0 || (uint32_t)rows[i] > LMAX || (uint32_t)cols[i] > LMAX)
            return DCSR_ERR_UNSUPPORTED;  // local indices are 12-bit
        rb[i] = (uint32_t)R;
        cb[i] = (uint32_t)cacc;
        R += (uint32_t)rows[i];
        cacc += (uint32_t)cols[i];
        if ((uint64_t)cols[i] > Cmax) Cmax = cols[i];
    }
    rb[count] = (uint32_t)R;
    cb[count] = (uint32_t)cacc;
    if (R >= (1ull << 31) || cacc >= (1ull << 31)) return DCSR_ERR_UNSUPPORTED;
    dcsr_batch b = new dcsr_batch_s();
    b->M = (uint32_t)count;
    b->row_base_h = rb;
    b->col_base_h = cb;
    b->team = 8;
    b->thresh = 128;
    dcsr_status st = dcsr_create(&b->h, (int64_t)R, (int64_t)Cmax, cfg);
    if (st != DCSR_OK) {
        delete b;
        return st;
    }
    int dev = 0, ncu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    b->grid = (ncu > 0 ? ncu : 256) * 2;
    if ((uint32_t)b->grid > b->M) b->grid = (int)b->M;
    if (hipMalloc(&b->row_base, (count + 1) * 4) != hipSuccess || hipMalloc(&b->col_base, (count + 1) * 4) != hipSuccess ||
        hipMalloc(&b->mstat, count * 4) != hipSuccess ||
        hipMalloc(&b->scratch, (size_t)b->grid * S_WORDS * 4) != hipSuccess ||
        hipMalloc(&b->bs, sizeof(BatchState)) != hipSuccess ||
        hipStreamCreateWithFlags(&b->cs, hipStreamNonBlocking) != hipSuccess) {
        dcsr_batch_destroy(b);
        return DCSR_ERR_OOM;
    }
    if (hipMemcpy(b->row_base, rb.data(), (count + 1) * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->col_base, cb.data(), (count + 1) * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(b->mstat, 0, count * 4) != hipSuccess || hipMemset(b->bs, 0, sizeof(BatchState)) != hipSuccess) {
        dcsr_batch_destroy(b);
        return DCSR_ERR_HIP;
    }
    *out = b;
    return DCSR_OK;
}

dcsr_status dcsr_batch_destroy(dcsr_batch b)
{
    if (!b) return DCSR_ERR_INVALID;
    (void)hipDeviceSynchronize();
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cs) (void)hipStreamDestroy(b->cs);
    (void)hipFree(b->row_base);
    (void)hipFree(b->col_base);
    (void)hipFree(b->mstat);
    (void)hipFree(b->scratch);
    (void)hipFree(b->bs);
    if (b->h) dcsr_destroy(b->h);
    delete b;
    return DCSR_OK;
}

dcsr_handle dcsr_batch_handle(dcsr_batch b) { return b ? b->h : nullptr; }

dcsr_status dcsr_batch_info(dcsr_batch b, int64_t *row_base, int64_t *col_base, int32_t *team, int32_t *thresh)
{
    if (!b) return DCSR_ERR_INVALID;
    for (uint32_t i = 0; i <= b->M; ++i) {
        if (row_base) row_base[i] = b->row_base_h[i];
        if (col_base) col_base[i] = b->col_base_h[i];
    }
    if (team) *team = b->team;
    if (thresh) *thresh = b->thresh;
    return DCSR_OK;
}

dcsr_status dcsr_batch_load_coo(dcsr_batch b, const int64_t *tri_off, const int32_t *rows, const int32_t *cols,
                                const float *vals, int64_t n, dcsr_stream_t stream)
{
    if (!b || !tri_off || n < 0 || (n > 0 && (!rows || !cols || !vals))) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    dcsr_handle h = b->h;
    int32_t *gr, *gc;
    DCSR_TRY(ws_get(h, WS_SG_KEY0, (size_t)(n + 1) * 4, (void **)&gr));
    DCSR_TRY(ws_get(h, WS_SG_KEY1, (size_t)(n + 1) * 4, (void **)&gc));
    if (n > 0)
        k_batch_globalize<<<grid_for(n, 256), 256, 0, s>>>(tri_off, b->M, n, b->row_base, b->col_base, rows, cols, gr,
                                                           gc, h->ds);
    DCSR_HIP(hipGetLastError());
    return load_coo_internal(h, gr, gc, vals, n, s);
}

template <int TEAM>
static void launch_step(dcsr_batch b, const int64_t *tri_off, int64_t n, const int32_t *rows, const int32_t *cols,
                        const float *vals, const float *x, float *y, int do_update, int do_spmv, int reserved,
                        hipStream_t s)
{
    dcsr_handle h = b->h;
    k_batch_step<TEAM><<<b->grid, BT, 0, s>>>(b->M, b->row_base, b->col_base, tri_off, n, rows, cols, vals, h->row_start,
                                              h->row_len, h->row_cap, h->pool_col, h->pool_val,
                                              (unsigned long long)h->pool_cap, h->ds, b->mstat, b->scratch, x, y,
                                              (uint32_t)b->thresh, b->bs, do_update, do_spmv, reserved);
}

static void launch_any(dcsr_batch b, const int64_t *tri_off, int64_t n, const int32_t *rows, const int32_t *cols,
                       const float *vals, const float *x, float *y, int do_update, int do_spmv, int reserved,
                       hipStream_t s)
{
    switch (b->team) {
    case 4: launch_step<4>(b, tri_off, n, rows, cols, vals, x, y, do_update, do_spmv, reserved, s); break;
    case 16: launch_step<16>(b, tri_off, n, rows, cols, vals, x, y, do_update, do_spmv, reserved, s); break;
    case 32: launch_step<32>(b, tri_off, n, rows, cols, vals, x, y, do_update, do_spmv, reserved, s); break;
    default: launch_step<8>(b, tri_off, n, rows, cols, vals, x, y, do_update, do_spmv, reserved, s); break;
    }
}

// Pool headroom for one step: every matrix may rewrite all its rows plus its
// batch.  Compacts/grows ahead of the launch so the kernel never runs out
// (*reserved = 1: the claims need no fit check).  With allow_grow = 0 the
// pool keeps its capacity (retired chunks are reclaimed first): if the room
// may still fall short (*reserved = 0), a matrix whose claim does not fit
// fails alone with its exhausted bit and leaves room for the others.
static dcsr_status batch_reserve(dcsr_batch b, int64_t n, hipStream_t s, int *reserved)
{
    dcsr_handle h = b->h;
    if (!h->state_known) DCSR_TRY(pull_state(h, s));
    const int64_t need = h->nnz_h + n + 4 * (int64_t)h->rows + 64;
    *reserved = 1;
    if (h->pool_cap - h->bump_h >= need) return DCSR_OK;
    if (!h->cfg.allow_grow) {
        if (h->bump_h > h->nnz_h) DCSR_TRY(compact_in_order(h, s));
        // after a compaction the bump pointer is the live count
        *reserved = h->pool_cap - h->nnz_h >= need ? 1 : 0;
        return DCSR_OK;
    }
    return pool_reserve(h, need, s, nullptr);
}

dcsr_status dcsr_batch_step(dcsr_batch b, const int64_t *tri_off, const int32_t *rows, const int32_t *cols,
                            const float *vals, int64_t n, const float *x, float *y, uint32_t flags,
                            dcsr_stream_t stream)
{
    if (!b) return DCSR_ERR_INVALID;
    const int do_update = tri_off != nullptr && n > 0;
    const int do_spmv = x != nullptr && y != nullptr;
    if (!do_update && !do_spmv) return DCSR_OK;
    if (do_update && (!rows || !cols || !vals)) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    dcsr_handle h = b->h;
    h->last_stream = s;
    if (!h->canonical) DCSR_TRY(canonicalize(h, s));
    int reserved = 0;
    if (do_update) DCSR_TRY(batch_reserve(b, n, s, &reserved));
    k_batch_zero_stat<<<grid_for(b->M, 256), 256, 0, s>>>(b->mstat, b->M, h->ds);
    if (flags & DCSR_BATCH_GRAPH) {
        const void *key[13] = {tri_off, rows, cols, vals, x, y, h->pool_col, (const void *)(intptr_t)b->team,
                               (const void *)(intptr_t)b->thresh,
                               (const void *)(intptr_t)(reserved * 4 + do_update * 2 + do_spmv),
                               (const void *)(intptr_t)h->pool_cap, (const void *)(intptr_t)n, h->row_start};
        bool same = b->gexec != nullptr;
        for (int i = 0; i < 13 && same; ++i) same = key[i] == b->gkey[i];
        if (!same) {
            if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
            if (b->graph) (void)hipGraphDestroy(b->graph);
            b->gexec = nullptr;
            b->graph = nullptr;
            // capture on the batch's own stream (the caller's may be the null
            // stream, which cannot be captured); nothing runs during capture
            DCSR_HIP(hipStreamBeginCapture(b->cs, hipStreamCaptureModeThreadLocal));
            launch_any(b, tri_off, n, rows, cols, vals, x, y, do_update, do_spmv, reserved, b->cs);
            DCSR_HIP(hipStreamEndCapture(b->cs, &b->graph));
            DCSR_HIP(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
            for (int i = 0; i < 13; ++i) b->gkey[i] = key[i];
        }
        DCSR_HIP(hipGraphLaunch(b->gexec, s));
    } else {
        launch_any(b, tri_off, n, rows, cols, vals, x, y, do_update, do_spmv, reserved, s);
    }
    DCSR_HIP(hipGetLastError());
    if (do_update) {
        state_unknown(h, n, h->pool_cap);
        h->flat_dirty = true;
    }
    return DCSR_OK;
}

// Status of the last step: OR of the per-matrix bits (syncs); `per_matrix`
// (host, M entries) receives them when not NULL.  Also refreshes the handle's
// counters (nnz, inserted / added / deleted entries and touched rows of the
// step, summed over the matrices).
dcsr_status dcsr_batch_status(dcsr_batch b, uint32_t *per_matrix)
{
    if (!b) return DCSR_ERR_INVALID;
    std::vector<uint32_t> st(b->M);
    DCSR_HIP(hipMemcpyAsync(st.data(), b->mstat, b->M * 4, hipMemcpyDeviceToHost, b->h->last_stream));
    DCSR_HIP(hipStreamSynchronize(b->h->last_stream));
    // the step's counters (summed over matrices) become the handle's
    // "last update" figures reported by dcsr_query
    DCSR_TRY(pull_state(b->h, b->h->last_stream));
    b->h->last_ins = b->h->hs->n_ins;
    b->h->last_add = b->h->hs->n_add;
    b->h->last_del = b->h->hs->n_del;
    b->h->last_touched = b->h->hs->ntouched;
    uint32_t any = 0;
    for (uint32_t i = 0; i < b->M; ++i) {
        any |= st[i];
        if (per_matrix) per_matrix[i] = st[i];
    }
    if (any & BST_BAD_INDEX) return DCSR_ERR_BAD_INDEX;
    if (any & BST_TOO_BIG) return DCSR_ERR_UNSUPPORTED;
    if (any & BST_EXHAUSTED) return DCSR_ERR_POOL_EXHAUSTED;
    return DCSR_OK;
}

// Autotuner: times the SpMV part of the step (read-only) for every
// (lanes per row, long-row threshold) pair on the batch as loaded, keeps the
// fastest, and caches the choice per shape class (matrix count and mean row
// length, log2 buckets) for the process.
dcsr_status dcsr_batch_autotune(dcsr_batch b, const float *x, float *y, int32_t *team_out, int32_t *thresh_out,
                                dcsr_stream_t stream)
{
    if (!b || !x || !y) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    dcsr_handle h = b->h;
    if (!h->state_known) DCSR_TRY(pull_state(h, s));
    const double mean_len = h->rows ? (double)h->nnz_h / (double)h->rows : 0.0;
    TuneKey key{bucket_log2((double)b->M), bucket_log2(mean_len)};
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
        b->team = it->second.team;
        b->thresh = it->second.thresh;
    } else {
        const int teams[4] = {4, 8, 16, 32};
        const int threshs[3] = {64, 256, 1024};
        hipEvent_t e0, e1;
        DCSR_HIP(hipEventCreate(&e0));
        DCSR_HIP(hipEventCreate(&e1));
        float best = 1e30f;
        int bt = 8, bth = 128;
        for (int ti = 0; ti < 4; ++ti)
            for (int hi = 0; hi < 3; ++hi) {
                b->team = teams[ti];
                b->thresh = threshs[hi];
                launch_any(b, nullptr, 0, nullptr, nullptr, nullptr, x, y, 0, 1, 0, s);  // warm
                DCSR_HIP(hipEventRecord(e0, s));
                for (int rep = 0; rep < 3; ++rep) launch_any(b, nullptr, 0, nullptr, nullptr, nullptr, x, y, 0, 1, 0, s);
                DCSR_HIP(hipEventRecord(e1, s));
                DCSR_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                DCSR_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) {
                    best = ms;
                    bt = teams[ti];
                    bth = threshs[hi];
                }
            }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        b->team = bt;
        b->thresh = bth;
        g_tune_cache[key] = TuneVal{bt, bth};
    }
    if (team_out) *team_out = b->team;
    if (thresh_out) *thresh_out = b->thresh;
    return DCSR_OK;
}

dcsr_status dcsr_batch_set_shape(dcsr_batch b, int32_t team, int32_t thresh)
{
    if (!b || !(team == 4 || team == 8 || team == 16 || team == 32) || thresh < 0) return DCSR_ERR_INVALID;
    b->team = team;
    b->thresh = thresh;
    return DCSR_OK;
}

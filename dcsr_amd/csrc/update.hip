// update.hip -- in-place insert-or-add of a batch of triples.
//
// Replaces device::UpdateMatrix -> UpdateMatrix_dcsr (include/load.h:258-325,
// include/sparse_update.inl:16-130).  The reference appends a batch row into
// a new segment claimed by an atomicAdd on the pool bump pointer, without
// merging duplicates, adding to existing coordinates or dropping zeros.  This
// engine produces exactly the host path's result (sort batch by row,
// COO->CSR, cusp::add; cusp/system/detail/sequential/elementwise.h:44-131):
//
//   1. pack (row, col) keys, stable LSD radix sort of (key, value) pairs;
//   2. run heads: one entry per distinct coordinate, summed left-to-right in
//      batch order starting from 0 (the host path's B_row[j] accumulation);
//   3. probe every distinct coordinate in its (column-sorted) row: existing
//      -> new = old + sum, removed when == 0; missing -> inserted when != 0;
//   4. plan per touched row: rows whose structure is unchanged get their
//      values written in place; the others claim a new chunk from the pool's
//      bump region (one wave-aggregated atomicAdd per 64 rows: DPP prefix sum
//      of the chunk sizes, lane 63 claims the total) and retire the old one;
//   5. a one-thread check rejects the whole batch (no-op) on a bad index or
//      pool exhaustion, so a batch is either fully applied or not at all;
//   6. move surviving old entries and write inserted ones at their merged
//      positions, then commit the row table.
#include "dcsr_internal.h"
#include "prims.h"
#include "segments.h"

namespace {

constexpr int TB = 256;
enum { C_NOP = 0, C_ADD = 1, C_DEL = 2, C_INS = 3 };
enum { M_VALUES = 0, M_REWRITE = 1 };

__global__ void k_upd_begin(DevState *ds)
{
    ds->d_nnz = 0;
    ds->d_garbage = 0;
    ds->n_ins = ds->n_add = ds->n_del = 0;
    ds->d_hubs = 0;
    ds->d_hub_chunks = 0;
    ds->ubad = 0;
    ds->go = 0;
    ds->need = 0;
    ds->hub_touch = 0;
    ds->bump_saved = ds->bump;
}

__global__ void k_upd_replan_reset(DevState *ds)
{
    ds->d_nnz = 0;
    ds->d_garbage = 0;
    ds->d_hubs = 0;
    ds->d_hub_chunks = 0;
    ds->go = 0;
    ds->need = 0;
    ds->hub_touch = 0;
    ds->bump_saved = ds->bump;
}

// Invalid indices become ~0 keys: with one key bit above row|col they sort
// strictly after every valid key.
__global__ void k_upd_pack(const int32_t *__restrict__ rows, const int32_t *__restrict__ cols,
                           const float *__restrict__ vals, int64_t n, int64_t R, int64_t C, int cbits,
                           uint64_t *__restrict__ key, uint32_t *__restrict__ val, DevState *ds)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (i < n) {
        int32_t r = rows[i], c = cols[i];
        bad = (r < 0) | (r >= R) | (c < 0) | (c >= C);
        key[i] = bad ? ~0ull : (((uint64_t)(uint32_t)r << cbits) | (uint32_t)c);
        val[i] = __float_as_uint(vals[i]);
    }
    if (__ballot(bad) && lane_id() == 0) atomicOr(&ds->ubad, 1u);
}

// flags[2i] = distinct-coordinate head, flags[2i+1] = row head.
__global__ void k_upd_flags(const uint64_t *__restrict__ key, int64_t n, int cbits, uint32_t *__restrict__ flags)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t k = key[i];
    bool valid = k != ~0ull;
    bool head = valid && (i == 0 || key[i - 1] != k);
    bool rhead = valid && (i == 0 || (key[i - 1] >> cbits) != (k >> cbits));
    flags[2 * i] = head ? 1u : 0u;
    flags[2 * i + 1] = rhead ? 1u : 0u;
}

// Distinct coordinates: ucol/usum/ut; touched rows: trow/tubeg.
__global__ void k_upd_compact(const uint64_t *__restrict__ key, const uint32_t *__restrict__ val, int64_t n,
                              int cbits, const uint32_t *__restrict__ flags, const uint32_t *__restrict__ pre,
                              int32_t *__restrict__ ucol, float *__restrict__ usum, uint32_t *__restrict__ ut,
                              uint32_t *__restrict__ trow, uint32_t *__restrict__ tubeg, uint32_t *__restrict__ tins,
                              uint32_t *__restrict__ tdel)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flags[2 * i]) return;
    const uint64_t k = key[i];
    const bool rhead = flags[2 * i + 1] != 0;
    const uint32_t u = pre[2 * i];
    const uint32_t t = pre[2 * i + 1] + (rhead ? 0u : 0xffffffffu);
    float s = 0.0f + __uint_as_float(val[i]);
    for (int64_t m = i + 1; m < n && key[m] == k; ++m) s = s + __uint_as_float(val[m]);
    ucol[u] = (int32_t)(k & ((1ull << cbits) - 1ull));
    usum[u] = s;
    ut[u] = t;
    if (rhead) {
        trow[t] = (uint32_t)(k >> cbits);
        tubeg[t] = u;
        tins[t] = 0;
        tdel[t] = 0;
    }
}

// Probe each distinct coordinate in its row.
__global__ __launch_bounds__(TB) void k_upd_probe(int64_t n, const DevState *__restrict__ ds_ro,
                                                  const uint32_t *__restrict__ row_start,
                                                  const int32_t *__restrict__ row_len,
                                                  const int32_t *__restrict__ pcol, const float *__restrict__ pval,
                                                  const int32_t *__restrict__ ucol, const float *__restrict__ usum,
                                                  const uint32_t *__restrict__ ut, const uint32_t *__restrict__ trow,
                                                  uint32_t *__restrict__ tubeg, float *__restrict__ uval,
                                                  uint32_t *__restrict__ upos, uint32_t *__restrict__ ucode,
                                                  uint32_t *__restrict__ uflag, uint32_t *__restrict__ tins,
                                                  uint32_t *__restrict__ tdel, DevState *ds)
{
    int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nu = ds_ro->nuniq;
    uint32_t code = C_NOP;
    if (u < nu) {
        uint32_t t = ut[u];
        uint32_t r = trow[t];
        uint32_t st = row_start[r];
        int32_t L = row_len[r];
        int32_t c = ucol[u];
        int32_t lo = 0, hi = L;  // lower bound of c in the row
        while (lo < hi) {
            int32_t mid = (lo + hi) >> 1;
            if (pcol[(uint64_t)st + mid] < c) lo = mid + 1; else hi = mid;
        }
        bool hit = lo < L && pcol[(uint64_t)st + lo] == c;
        float s = usum[u];
        float nv;
        if (hit) {
            nv = pval[(uint64_t)st + lo] + s;
            code = (nv != 0.0f) ? C_ADD : C_DEL;
        } else {
            nv = s;
            code = (s != 0.0f) ? C_INS : C_NOP;
        }
        uval[u] = nv;
        upos[u] = (uint32_t)lo;
        ucode[u] = code;
        if (code == C_INS) atomicAdd(&tins[t], 1u);
        if (code == C_DEL) atomicAdd(&tdel[t], 1u);
        if (u == 0) tubeg[ds_ro->ntouched] = nu;
    }
    if (u <= n) {  // index n is the scan's end sentinel
        uflag[2 * u] = code == C_INS ? 1u : 0u;
        uflag[2 * u + 1] = code == C_DEL ? 1u : 0u;
    }
    uint32_t ni = wave_sum_u32(code == C_INS), na = wave_sum_u32(code == C_ADD), nd = wave_sum_u32(code == C_DEL);
    if (lane_id() == 0) {
        if (ni) atomicAdd((unsigned long long *)&ds->n_ins, (unsigned long long)ni);
        if (na) atomicAdd((unsigned long long *)&ds->n_add, (unsigned long long)na);
        if (nd) atomicAdd((unsigned long long *)&ds->n_del, (unsigned long long)nd);
    }
}

// Per touched row: new length, mode, chunk claim.
__global__ __launch_bounds__(TB) void k_upd_plan(int64_t n, const uint32_t *__restrict__ trow,
                                                 const uint32_t *__restrict__ tins, const uint32_t *__restrict__ tdel,
                                                 const uint32_t *__restrict__ row_start,
                                                 const int32_t *__restrict__ row_len,
                                                 const int32_t *__restrict__ row_cap, uint32_t *__restrict__ tolds,
                                                 uint32_t *__restrict__ toldl, uint32_t *__restrict__ tdest,
                                                 uint32_t *__restrict__ tnewl, uint32_t *__restrict__ tmode,
                                                 uint32_t *__restrict__ tmoved, DevState *ds)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nt = ds->ntouched;
    uint32_t claim = 0, mode = M_VALUES, L = 0, nl = 0, cap = 0, r = 0;
    if (t < nt) {
        r = trow[t];
        L = (uint32_t)row_len[r];
        cap = (uint32_t)row_cap[r];
        uint32_t ni = tins[t], nd = tdel[t];
        nl = L + ni - nd;
        if (ni | nd) {
            mode = M_REWRITE;
            claim = chunk_round(nl);
        }
    }
    // wave-aggregated claim of the bump region
    uint32_t inc = wave_incl_scan_u32(claim);
    uint32_t wtot = wave_read_lane_u32(inc, 63);
    unsigned long long base = 0;
    if (lane_id() == 63 && wtot) base = atomicAdd(&ds->bump, (unsigned long long)wtot);
    base = __shfl(base, 63);
    if (t < nt) {
        tolds[t] = row_start[r];
        toldl[t] = L;
        tdest[t] = (uint32_t)(base + inc - claim);
        tnewl[t] = nl;
        tmode[t] = mode;
    }
    if (t <= n) tmoved[t] = (t < nt && mode == M_REWRITE) ? L : 0u;  // index n: end sentinel
    long long dn = (long long)nl - (long long)L;
    long long dg = mode == M_REWRITE ? (long long)cap : 0;
    int hub_now = nl > DCSR_HUB_LEN, hub_was = L > DCSR_HUB_LEN;
    int dh = mode == M_REWRITE ? hub_now - hub_was : 0;
    int touch = mode == M_REWRITE && (hub_now | hub_was);
    long long dc = 0;
    if (mode == M_REWRITE) {
        long long cn = hub_now ? (nl + DCSR_HUB_CHUNK - 1) / DCSR_HUB_CHUNK : 0;
        long long cw = hub_was ? (L + DCSR_HUB_CHUNK - 1) / DCSR_HUB_CHUNK : 0;
        dc = cn - cw;
    }
    // per-wave sums of the deltas (two's complement, |sum| < 2^31 per wave)
    int sdn = (int)wave_sum_u32((uint32_t)(int)dn);
    int sdg = (int)wave_sum_u32((uint32_t)dg);
    int sdh = (int)wave_sum_u32((uint32_t)dh);
    int sdc = (int)wave_sum_u32((uint32_t)dc);
    uint32_t stt = wave_sum_u32((uint32_t)touch);
    if (lane_id() == 0) {
        if (sdn) atomicAdd((unsigned long long *)&ds->d_nnz, (unsigned long long)(long long)sdn);
        if (sdg) atomicAdd((unsigned long long *)&ds->d_garbage, (unsigned long long)(long long)sdg);
        if (sdh) atomicAdd((unsigned long long *)&ds->d_hubs, (unsigned long long)(long long)sdh);
        if (sdc) atomicAdd((unsigned long long *)&ds->d_hub_chunks, (unsigned long long)(long long)sdc);
        if (stt) atomicAdd(&ds->hub_touch, stt);
    }
}

__global__ void k_upd_check(DevState *ds, unsigned long long pool_cap)
{
    unsigned long long claimed = ds->bump - ds->bump_saved;
    ds->need = claimed;
    if (ds->ubad) {
        ds->go = 0;
        ds->bump = ds->bump_saved;
        ds->status |= ST_BAD_INDEX;
    } else if (ds->bump > pool_cap) {
        ds->go = 0;
        ds->bump = ds->bump_saved;
        ds->status |= ST_POOL_EXHAUSTED;
    } else {
        ds->go = 1;
    }
}

// Move surviving old entries of rewritten rows to their merged positions.
__global__ __launch_bounds__(TB) void k_upd_move(const DevState *__restrict__ ds, const uint32_t *__restrict__ tmoff,
                                                 const uint32_t *__restrict__ tolds,
                                                 const uint32_t *__restrict__ tdest,
                                                 const uint32_t *__restrict__ tubeg, const int32_t *__restrict__ ucol,
                                                 const float *__restrict__ uval, const uint32_t *__restrict__ ucode,
                                                 const uint32_t *__restrict__ upre, int32_t *__restrict__ pcol,
                                                 float *__restrict__ pval, uint32_t per_wave)
{
    __shared__ uint32_t s_off[TB / 64][65];
    if (!ds->go) return;
    const uint32_t total = ds->nrewrite_old;
    const uint32_t nt = ds->ntouched;
    const int w = threadIdx.x >> 6;
    const uint64_t nwaves = (uint64_t)gridDim.x * (TB / 64);
    for (uint64_t wave = (uint64_t)blockIdx.x * (TB / 64) + w; wave * per_wave < total; wave += nwaves) {
        uint64_t e0 = wave * per_wave;
        uint64_t e1 = e0 + per_wave < total ? e0 + per_wave : total;
        wave_visit(tmoff, nt, (uint32_t)e0, (uint32_t)e1, s_off[w], [&](bool valid, uint32_t t, uint32_t k, uint32_t) {
            if (!valid) return;
            uint64_t sp = (uint64_t)tolds[t] + k;
            int32_t c = pcol[sp];
            float v = pval[sp];
            uint32_t ub = tubeg[t], ue = tubeg[t + 1];
            uint32_t lo = ub, hi = ue;
            while (lo < hi) {
                uint32_t mid = (lo + hi) >> 1;
                if (ucol[mid] < c) lo = mid + 1; else hi = mid;
            }
            if (lo < ue && ucol[lo] == c) {
                if (ucode[lo] == C_DEL) return;
                v = uval[lo];
            }
            int64_t shift = (int64_t)(upre[2 * lo] - upre[2 * ub]) - (int64_t)(upre[2 * lo + 1] - upre[2 * ub + 1]);
            uint64_t d = (uint64_t)((int64_t)tdest[t] + (int64_t)k + shift);
            pcol[d] = c;
            pval[d] = v;
        });
    }
}

// Inserted entries, in-place value updates, row-table commit.
__global__ __launch_bounds__(TB) void k_upd_write(DevState *ds, const uint32_t *__restrict__ ut,
                                                  const uint32_t *__restrict__ trow, const uint32_t *__restrict__ tubeg,
                                                  const uint32_t *__restrict__ tolds,
                                                  const uint32_t *__restrict__ tdest,
                                                  const uint32_t *__restrict__ tnewl,
                                                  const uint32_t *__restrict__ tmode, const int32_t *__restrict__ ucol,
                                                  const float *__restrict__ uval, const uint32_t *__restrict__ upos,
                                                  const uint32_t *__restrict__ ucode,
                                                  const uint32_t *__restrict__ upre, int32_t *__restrict__ pcol,
                                                  float *__restrict__ pval, uint32_t *__restrict__ row_start,
                                                  int32_t *__restrict__ row_len, int32_t *__restrict__ row_cap)
{
    if (!ds->go) return;
    int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nu = ds->nuniq;
    if (u == 0) {
        ds->nnz += ds->d_nnz;
        ds->garbage += ds->d_garbage;
    }
    if (u >= nu) return;
    uint32_t t = ut[u];
    uint32_t code = ucode[u];
    uint32_t mode = tmode[t];
    uint32_t ub = tubeg[t];
    if (mode == M_VALUES) {
        if (code == C_ADD) pval[(uint64_t)tolds[t] + upos[u]] = uval[u];
        return;
    }
    if (code == C_INS) {
        int64_t shift = (int64_t)(upre[2 * u] - upre[2 * ub]) - (int64_t)(upre[2 * u + 1] - upre[2 * ub + 1]);
        uint64_t d = (uint64_t)((int64_t)tdest[t] + (int64_t)upos[u] + shift);
        pcol[d] = ucol[u];
        pval[d] = uval[u];
    }
    if (u == ub) {
        uint32_t r = trow[t];
        uint32_t nl = tnewl[t];
        row_start[r] = tdest[t];
        row_len[r] = (int32_t)nl;
        row_cap[r] = (int32_t)chunk_round(nl);
    }
}

struct UpdBufs {
    uint64_t *k0, *k1;
    uint32_t *v0, *v1, *hist, *part, *flags, *pre;
    int32_t *ucol;
    float *usum, *uval;
    uint32_t *ut, *upos, *ucode, *uflag, *upre;
    uint32_t *trow, *tubeg, *tins, *tdel, *tolds, *toldl, *tdest, *tnewl, *tmode, *tmoved, *tmoff;
};

static dcsr_status upd_bufs(dcsr_handle h, int64_t n, UpdBufs &b)
{
    size_t n4 = (size_t)(n + 2) * 4, n8 = (size_t)(n + 2) * 8;
    DCSR_TRY(ws_get(h, WS_KEY_A, n8, (void **)&b.k0));
    DCSR_TRY(ws_get(h, WS_KEY_B, n8, (void **)&b.k1));
    DCSR_TRY(ws_get(h, WS_VAL_A, n4, (void **)&b.v0));
    DCSR_TRY(ws_get(h, WS_VAL_B, n4, (void **)&b.v1));
    size_t partw = scan_part_words(n + 2, 2);
    size_t rpw = radix_part_words(n);
    DCSR_TRY(ws_get(h, WS_HIST, radix_hist_words(n) * 4, (void **)&b.hist));
    DCSR_TRY(ws_get(h, WS_SCAN, (partw > rpw ? partw : rpw) * 4, (void **)&b.part));
    DCSR_TRY(ws_get(h, WS_FLAGS, n8, (void **)&b.flags));
    DCSR_TRY(ws_get(h, WS_TMP0, n8, (void **)&b.pre));
    DCSR_TRY(ws_get(h, WS_UCOL, n4, (void **)&b.ucol));
    DCSR_TRY(ws_get(h, WS_UPOS, n4, (void **)&b.upos));
    DCSR_TRY(ws_get(h, WS_UVAL, n4, (void **)&b.uval));
    DCSR_TRY(ws_get(h, WS_UT, n4, (void **)&b.ut));
    DCSR_TRY(ws_get(h, WS_UCODE, n4, (void **)&b.ucode));
    DCSR_TRY(ws_get(h, WS_UPRE, n8, (void **)&b.upre));
    DCSR_TRY(ws_get(h, WS_TMP1, n8, (void **)&b.uflag));
    DCSR_TRY(ws_get(h, WS_TROW, n4, (void **)&b.trow));
    DCSR_TRY(ws_get(h, WS_TUBEG, n4 + 4, (void **)&b.tubeg));
    DCSR_TRY(ws_get(h, WS_TINS, n4, (void **)&b.tins));
    DCSR_TRY(ws_get(h, WS_TDEL, n4, (void **)&b.tdel));
    DCSR_TRY(ws_get(h, WS_TOLDS, n4, (void **)&b.tolds));
    DCSR_TRY(ws_get(h, WS_TOLDL, n4, (void **)&b.toldl));
    DCSR_TRY(ws_get(h, WS_TDEST, n4, (void **)&b.tdest));
    DCSR_TRY(ws_get(h, WS_TNEWL, n4, (void **)&b.tnewl));
    DCSR_TRY(ws_get(h, WS_TMODE, n4, (void **)&b.tmode));
    DCSR_TRY(ws_get(h, WS_TOLDPRE, n4 + 4, (void **)&b.tmoff));
    DCSR_TRY(ws_get(h, WS_CANON, n4 + 4, (void **)&b.tmoved));
    DCSR_TRY(ws_get(h, WS_TMP2, n4, (void **)&b.usum));
    return DCSR_OK;
}

}  // namespace

static dcsr_status plan_and_check(dcsr_handle h, int64_t n, UpdBufs &b, hipStream_t s)
{
    k_upd_plan<<<grid_for(n + 1, TB), TB, 0, s>>>(n, b.trow, b.tins, b.tdel, h->row_start, h->row_len, h->row_cap, b.tolds,
                                               b.toldl, b.tdest, b.tnewl, b.tmode, b.tmoved, h->ds);
    DCSR_HIP(scan_exclusive_u32(b.tmoved, b.tmoff, n + 1, 1, b.part, &h->ds->nrewrite_old, s));
    k_upd_check<<<1, 1, 0, s>>>(h->ds, (unsigned long long)h->pool_cap);
    return hipGetLastError() == hipSuccess ? DCSR_OK : DCSR_ERR_HIP;
}

dcsr_status dcsr_update(dcsr_handle h, const int32_t *rows, const int32_t *cols, const float *vals, int64_t n,
                        uint32_t flags, dcsr_stream_t stream)
{
    if (!h || n < 0 || (n > 0 && (!rows || !cols || !vals))) return DCSR_ERR_INVALID;
    if (n >= (int64_t)0x7ffffff0ll) return DCSR_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    h->last_stream = s;
    if (!h->canonical) {
        // validate first: a rejected batch must not canonicalize the matrix
        if (n > 0) {
            uint64_t *k0;
            uint32_t *v0;
            DCSR_TRY(ws_get(h, WS_KEY_A, (size_t)n * 8, (void **)&k0));
            DCSR_TRY(ws_get(h, WS_VAL_A, (size_t)n * 4, (void **)&v0));
            k_upd_begin<<<1, 1, 0, s>>>(h->ds);
            k_upd_pack<<<grid_for(n, TB), TB, 0, s>>>(rows, cols, vals, n, h->rows, h->cols, h->cbits, k0, v0, h->ds);
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
    DCSR_TRY(upd_bufs(h, n, b));

    const int nbits = h->rbits + h->cbits + 1;
    k_upd_begin<<<1, 1, 0, s>>>(h->ds);
    k_upd_pack<<<grid_for(n, TB), TB, 0, s>>>(rows, cols, vals, n, h->rows, h->cols, h->cbits, b.k0, b.v0, h->ds);
    int which = 0;
    DCSR_HIP(radix_sort_pairs(b.k0, b.v0, b.k1, b.v1, n, nbits, b.hist, b.part, &which, s));
    uint64_t *key = which ? b.k1 : b.k0;
    uint32_t *val = which ? b.v1 : b.v0;
    k_upd_flags<<<grid_for(n, TB), TB, 0, s>>>(key, n, h->cbits, b.flags);
    DCSR_HIP(scan_exclusive_u32(b.flags, b.pre, n, 2, b.part, &h->ds->nuniq, s));
    k_upd_compact<<<grid_for(n, TB), TB, 0, s>>>(key, val, n, h->cbits, b.flags, b.pre, b.ucol, b.usum, b.ut, b.trow,
                                                  b.tubeg, b.tins, b.tdel);
    k_upd_probe<<<grid_for(n + 1, TB), TB, 0, s>>>(n, h->ds, h->row_start, h->row_len, h->pool_col, h->pool_val, b.ucol,
                                                b.usum, b.ut, b.trow, b.tubeg, b.uval, b.upos, b.ucode, b.uflag, b.tins,
                                                b.tdel, h->ds);
    DCSR_HIP(scan_exclusive_u32(b.uflag, b.upre, n + 1, 2, b.part, nullptr, s));
    DCSR_TRY(plan_and_check(h, n, b, s));

    const bool nosync = (flags & DCSR_UPDATE_NOSYNC) != 0;
    if (!nosync) {
        DCSR_TRY(pull_state(h, s));
        if (!h->hs->go) {
            unsigned bits = h->hs->status;
            if (h->hs->ubad) {
                return DCSR_ERR_BAD_INDEX;
            }
            if ((bits & ST_POOL_EXHAUSTED) && h->cfg.allow_grow && !(flags & DCSR_UPDATE_NOGROW)) {
                // make room, then plan again (row contents are unchanged, only
                // their locations move)
                int64_t need = (int64_t)h->hs->need;
                DCSR_TRY(pool_reserve(h, need + need / 8 + 64, s, nullptr));
                k_upd_replan_reset<<<1, 1, 0, s>>>(h->ds);
                DCSR_TRY(plan_and_check(h, n, b, s));
                DCSR_TRY(pull_state(h, s));
                if (!h->hs->go) return DCSR_ERR_POOL_EXHAUSTED;
            } else {
                return DCSR_ERR_POOL_EXHAUSTED;
            }
        }
    }
    // persistent move kernel: any grid size is correct
    {
        const uint32_t per_wave = 1024;
        int64_t bound = nosync ? h->pool_cap : (int64_t)h->hs->nrewrite_old;
        int64_t waves = (bound + per_wave - 1) / per_wave;
        int64_t blocks = (waves + 3) / 4;
        if (blocks > 4096) blocks = 4096;
        if (blocks < 1) blocks = 1;
        k_upd_move<<<(unsigned)blocks, TB, 0, s>>>(h->ds, b.tmoff, b.tolds, b.tdest, b.tubeg, b.ucol, b.uval, b.ucode,
                                                   b.upre, h->pool_col, h->pool_val, per_wave);
    }
    k_upd_write<<<grid_for(n, TB), TB, 0, s>>>(h->ds, b.ut, b.trow, b.tubeg, b.tolds, b.tdest, b.tnewl, b.tmode, b.ucol,
                                                b.uval, b.upos, b.ucode, b.upre, h->pool_col, h->pool_val, h->row_start,
                                                h->row_len, h->row_cap);
    DCSR_HIP(hipGetLastError());
    if (nosync) {
        h->state_known = false;
        h->hubs_dirty = true;
        return DCSR_OK;
    }
    // host mirrors from the check point
    const DevState &st = *h->hs;
    h->nnz_h = st.nnz + st.d_nnz;
    h->bump_h = (int64_t)st.bump;
    h->garbage_h = st.garbage + st.d_garbage;
    h->last_ins = st.n_ins;
    h->last_add = st.n_add;
    h->last_del = st.n_del;
    h->last_touched = st.ntouched;
    if (st.hub_touch) h->hubs_dirty = true;
    h->state_known = true;
    return DCSR_OK;
}

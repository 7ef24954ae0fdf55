// server.hip -- resident-matrix request loop with continuous batching.
//
// No reference counterpart (the reference drivers call the API directly,
// src/Fill_Test.inl).  A server owns one resident matrix and two streams:
//
//   * queries (SpMV, SpMM with k columns) that sit together in the queue
//     between two updates see the same matrix version and form one group,
//     packed into ONE multi-column SpMM (their x columns side by side in a
//     row-major buffer, at most max_cols columns): A is streamed once per
//     32 packed columns and each nonzero gathers one line of the packed X
//     row instead of one line per query; the result columns are scattered
//     back;
//   * an update is prepared (sort, merge plan, pool check -- reads the matrix
//     only) on the update stream while the previous query group still runs on
//     the query stream; its commit waits for that group (event), and the next
//     query group waits for the commit.  Every query therefore sees exactly
//     the updates submitted before it;
//   * admission by pool memory: when the prepared update needs more pool
//     than is free, the server waits for in-flight queries, compacts / grows
//     the pool, re-plans, then commits;
//   * requests that arrive while a group is on the GPU wait in the queue and
//     form the next group (continuous batching).  A completer thread watches
//     group events and timestamps every request (latency = completion time -
//     submit time on the host clock).
#include <hip/hip_runtime.h>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>
#include <string.h>
#include "dcsr_internal.h"

namespace {

struct PackDesc {
    const float *src;  // request's x (SpMV) or X (SpMM)
    float *dst;        // request's y / Y
    int64_t ld_src, ld_dst;
    int32_t k;         // columns
    int32_t col0;      // first column in the packed block
};

constexpr int MAXQ = 64;  // requests per group
// Query groups launched and not yet finished on the GPU: one running, one
// queued behind it.  The scheduler forms the next group only when fewer are
// in flight, so requests that arrive meanwhile join it (a group costs about
// one pass over A whatever its width up to 32 packed columns).
constexpr size_t MAX_INFLIGHT_GROUPS = 2;

// Packed X (C x Kp, row-major, leading dimension ld) from the group's
// inputs; columns [K, Kp) are zero padding.
__global__ void k_pack(const PackDesc *__restrict__ d, int nreq, int64_t C, int K, int Kp, int64_t ld,
                       float *__restrict__ X)
{
    __shared__ PackDesc sd[MAXQ];
    __shared__ int16_t owner[MAXQ * 32];
    for (int i = threadIdx.x; i < nreq; i += blockDim.x) sd[i] = d[i];
    __syncthreads();
    for (int c = threadIdx.x; c < K; c += blockDim.x) {
        int q = 0;
        while (q + 1 < nreq && sd[q + 1].col0 <= c) ++q;
        owner[c] = (int16_t)q;
    }
    __syncthreads();
    const int64_t n = C * Kp;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i / Kp;
        const int c = (int)(i - r * Kp);
        float v = 0.0f;
        if (c < K) {
            const PackDesc &p = sd[owner[c]];
            v = p.src[r * p.ld_src + (c - p.col0)];
        }
        X[r * ld + c] = v;
    }
}

// Scatter the packed Y (R x K of leading dimension ld) back to every
// request's output.
__global__ void k_unpack(const PackDesc *__restrict__ d, int nreq, int64_t R, int K, int64_t ld,
                         const float *__restrict__ Y)
{
    __shared__ PackDesc sd[MAXQ];
    __shared__ int16_t owner[MAXQ * 32];
    for (int i = threadIdx.x; i < nreq; i += blockDim.x) sd[i] = d[i];
    __syncthreads();
    for (int c = threadIdx.x; c < K; c += blockDim.x) {
        int q = 0;
        while (q + 1 < nreq && sd[q + 1].col0 <= c) ++q;
        owner[c] = (int16_t)q;
    }
    __syncthreads();
    const int64_t n = R * K;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i / K;
        const int c = (int)(i - r * K);
        const PackDesc &p = sd[owner[c]];
        p.dst[r * p.ld_dst + (c - p.col0)] = Y[r * ld + c];
    }
}

// Width the packed product runs at: a column-block width (1, 2, 4, 8, 16,
// 32) up to 32 columns, whole 32-column blocks beyond; rows of the packed
// buffers are padded to 32 floats (128 bytes, one cache line) so each
// nonzero's gather of a block never straddles two lines.
static int packed_width(int K)
{
    if (K > 32) return (K + 31) / 32 * 32;
    int w = 1;
    while (w < K) w <<= 1;
    return w;
}
static int64_t packed_ld(int Kp) { return (Kp + 31) / 32 * 32; }

using Clock = std::chrono::steady_clock;

struct Req {
    dcsr_request r;
    uint64_t ticket;
    Clock::time_point t_submit;
};

struct Done {
    Clock::time_point t_done;
    dcsr_status status;
    uint64_t group;
};

struct Inflight {
    hipEvent_t ev;
    std::vector<uint64_t> tickets;
    uint64_t group;
    dcsr_status status;
};

}  // namespace

struct dcsr_server_s {
    dcsr_handle h;
    int max_cols;
    hipStream_t sq, su;
    float *X, *Y;  // packing buffers
    PackDesc *dd;  // device descriptors (one slot per group in flight)
    PackDesc *hd;  // pinned host descriptors
    int nslots, slot;
    std::vector<hipEvent_t> slot_ev;  // last use of each descriptor slot
    std::vector<char> slot_used;
    std::mutex mu;
    std::condition_variable cv_sched, cv_done, cv_comp;
    std::deque<Req> queue;
    std::deque<Inflight> inflight;
    std::unordered_map<uint64_t, Done> done;
    std::unordered_map<uint64_t, Clock::time_point> submitted;
    uint64_t next_ticket = 1, groups = 0;
    bool stop = false;
    hipEvent_t last_query_ev;  // end of the last query group (current version)
    bool have_query_ev = false;
    std::thread sched, comp;
    dcsr_server_stats stats;
    int device = 0;
};

static void complete_group(dcsr_server_s *S, Inflight &f)
{
    (void)hipEventSynchronize(f.ev);
    const auto now = Clock::now();
    (void)hipEventDestroy(f.ev);
    std::lock_guard<std::mutex> lk(S->mu);
    for (uint64_t t : f.tickets) S->done[t] = Done{now, f.status, f.group};
    S->cv_done.notify_all();
}

static void completer_loop(dcsr_server_s *S)
{
    (void)hipSetDevice(S->device);
    for (;;) {
        Inflight f;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_comp.wait(lk, [&] { return S->stop || !S->inflight.empty(); });
            if (S->inflight.empty()) {
                if (S->stop) return;
                continue;
            }
            f = S->inflight.front();
            S->inflight.pop_front();
        }
        complete_group(S, f);
    }
}

static void push_inflight(dcsr_server_s *S, hipStream_t s, std::vector<uint64_t> &&tickets, dcsr_status st)
{
    Inflight f;
    (void)hipEventCreateWithFlags(&f.ev, hipEventDisableTiming);
    (void)hipEventRecord(f.ev, s);
    f.tickets = std::move(tickets);
    f.status = st;
    std::lock_guard<std::mutex> lk(S->mu);
    f.group = S->groups++;
    S->inflight.push_back(f);
    S->cv_comp.notify_one();
}

static dcsr_status run_update(dcsr_server_s *S, const Req &q)
{
    dcsr_handle h = S->h;
    const int64_t n = q.r.n;
    bool go = false;
    int64_t need = 0;
    // prepare overlaps with the query group still running on sq
    dcsr_status st = update_prepare(h, q.r.rows, q.r.cols, q.r.vals, n, S->su, &go, &need);
    if (st != DCSR_OK || n == 0) return st;
    if (!go) {
        // admission by pool memory: quiesce readers, make room, plan again
        (void)hipStreamSynchronize(S->sq);
        S->stats.compactions++;
        st = pool_reserve(h, need + need / 8 + 64, S->su, nullptr);
        if (st == DCSR_OK) st = update_replan(h, n, S->su, &go);
        if (st != DCSR_OK) return st;
        if (!go) return DCSR_ERR_POOL_EXHAUSTED;
    }
    if (S->have_query_ev) (void)hipStreamWaitEvent(S->su, S->last_query_ev, 0);
    st = update_commit(h, n, S->su, false);
    // the next query group sees this version
    hipEvent_t ev;
    (void)hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    (void)hipEventRecord(ev, S->su);
    (void)hipStreamWaitEvent(S->sq, ev, 0);
    (void)hipEventDestroy(ev);
    S->stats.updates++;
    S->stats.update_triples += n;
    return st;
}

static int req_cols(const dcsr_request &r) { return r.kind == DCSR_REQ_SPMV ? 1 : (int)r.k; }

static dcsr_status run_one(dcsr_server_s *S, const dcsr_request &r)
{
    if (r.kind == DCSR_REQ_SPMV) return dcsr_spmv(S->h, r.x, r.y, 1.0f, 0.0f, S->sq);
    return dcsr_spmm_dense(S->h, r.x, r.k, r.ldx, r.y, r.ldy, 1.0f, 0.0f, S->sq);
}

// One group (same matrix version).  Two or more queries are packed side by
// side into one multi-column product: A is streamed once per 32 packed
// columns instead of once per query, and each nonzero gathers one
// contiguous run of the packed X row (a 128-byte line holds 32 columns)
// instead of one line per query.
static dcsr_status run_queries(dcsr_server_s *S, const std::vector<Req> &g)
{
    dcsr_handle h = S->h;
    hipStream_t s = S->sq;
    dcsr_status st = DCSR_OK;
    std::vector<const dcsr_request *> pack;
    for (const Req &q : g) pack.push_back(&q.r);
    if (pack.size() == 1) {
        if (st == DCSR_OK) st = run_one(S, *pack[0]);
    } else if (pack.size() > 1 && st == DCSR_OK) {
        const int slot = S->slot;
        S->slot = (S->slot + 1) % S->nslots;
        // the host must not rewrite a slot whose previous copy is pending
        if (S->slot_used[slot]) (void)hipEventSynchronize(S->slot_ev[slot]);
        PackDesc *hd = S->hd + (size_t)slot * MAXQ;
        PackDesc *dd = S->dd + (size_t)slot * MAXQ;
        int K = 0;
        for (size_t i = 0; i < pack.size(); ++i) {
            const dcsr_request &r = *pack[i];
            const int k = req_cols(r);
            hd[i] = PackDesc{r.x, r.y, r.kind == DCSR_REQ_SPMV ? 1 : r.ldx, r.kind == DCSR_REQ_SPMV ? 1 : r.ldy, k, K};
            K += k;
        }
        (void)hipMemcpyAsync(dd, hd, pack.size() * sizeof(PackDesc), hipMemcpyHostToDevice, s);
        const int64_t C = h->cols, R = h->rows;
        const int Kp = packed_width(K);
        const int64_t ld = packed_ld(Kp);
        const unsigned gx = (unsigned)((C * Kp + 255) / 256 < 65536 ? (C * Kp + 255) / 256 : 65536);
        const unsigned gy = (unsigned)((R * K + 255) / 256 < 65536 ? (R * K + 255) / 256 : 65536);
        if (C > 0) k_pack<<<gx ? gx : 1, 256, 0, s>>>(dd, (int)pack.size(), C, K, Kp, ld, S->X);
        st = dcsr_spmm_dense(h, S->X, Kp, ld, S->Y, ld, 1.0f, 0.0f, s);
        if (R > 0) k_unpack<<<gy ? gy : 1, 256, 0, s>>>(dd, (int)pack.size(), R, K, ld, S->Y);
        if (hipGetLastError() != hipSuccess) st = DCSR_ERR_HIP;
        (void)hipEventRecord(S->slot_ev[slot], s);
        S->slot_used[slot] = 1;
    }
    S->stats.query_groups++;
    S->stats.queries += g.size();
    return st;
}

static void scheduler_loop(dcsr_server_s *S)
{
    (void)hipSetDevice(S->device);
    std::deque<hipEvent_t> groups_in_flight;  // oldest first
    for (;;) {
        std::vector<Req> group;
        Req upd;
        bool is_update = false;
        bool query_next;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_sched.wait(lk, [&] { return S->stop || !S->queue.empty(); });
            if (S->queue.empty() && S->stop) break;
            query_next = S->queue.front().r.kind != DCSR_REQ_UPDATE;
        }
        if (query_next) {
            // retire finished groups; with MAX_INFLIGHT_GROUPS still on the
            // GPU, wait for the oldest (outside the lock: arrivals keep
            // queueing and join the next group)
            while (!groups_in_flight.empty() && hipEventQuery(groups_in_flight.front()) == hipSuccess) {
                (void)hipEventDestroy(groups_in_flight.front());
                groups_in_flight.pop_front();
            }
            if (groups_in_flight.size() >= MAX_INFLIGHT_GROUPS) {
                (void)hipEventSynchronize(groups_in_flight.front());
                (void)hipEventDestroy(groups_in_flight.front());
                groups_in_flight.pop_front();
            }
        }
        {
            std::unique_lock<std::mutex> lk(S->mu);
            if (S->queue.empty()) continue;
            if (S->queue.front().r.kind == DCSR_REQ_UPDATE) {
                upd = S->queue.front();
                S->queue.pop_front();
                is_update = true;
            } else {
                int cols = 0;  // packed columns of the group
                while (!S->queue.empty() && S->queue.front().r.kind != DCSR_REQ_UPDATE && (int)group.size() < MAXQ) {
                    const int k = req_cols(S->queue.front().r);
                    if (!group.empty() && cols + k > S->max_cols) break;
                    cols += k;
                    group.push_back(S->queue.front());
                    S->queue.pop_front();
                }
            }
        }
        if (is_update) {
            dcsr_status st = run_update(S, upd);
            std::vector<uint64_t> t{upd.ticket};
            push_inflight(S, S->su, std::move(t), st);
        } else {
            dcsr_status st = run_queries(S, group);
            hipEvent_t ev, gev;
            (void)hipEventCreateWithFlags(&ev, hipEventDisableTiming);
            (void)hipEventRecord(ev, S->sq);
            if (S->have_query_ev) (void)hipEventDestroy(S->last_query_ev);
            S->last_query_ev = ev;
            S->have_query_ev = true;
            if (hipEventCreateWithFlags(&gev, hipEventDisableTiming) == hipSuccess) {
                (void)hipEventRecord(gev, S->sq);
                groups_in_flight.push_back(gev);
            }
            std::vector<uint64_t> t;
            for (auto &q : group) t.push_back(q.ticket);
            push_inflight(S, S->sq, std::move(t), st);
        }
    }
    for (hipEvent_t e : groups_in_flight) {
        (void)hipEventSynchronize(e);
        (void)hipEventDestroy(e);
    }
}

dcsr_status dcsr_server_create(dcsr_server *out, dcsr_handle h, int32_t max_cols)
{
    if (!out || !h || max_cols < 1 || max_cols > 32 * MAXQ) return DCSR_ERR_INVALID;
    dcsr_server_s *S = new dcsr_server_s();
    S->h = h;
    S->max_cols = max_cols;
    S->nslots = 64;
    (void)hipGetDevice(&S->device);
    dcsr_status st = DCSR_OK;
    if (hipStreamCreateWithFlags(&S->sq, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&S->su, hipStreamNonBlocking) != hipSuccess)
        st = DCSR_ERR_HIP;
    const size_t ldmax = (size_t)packed_ld(packed_width(max_cols));
    if (st == DCSR_OK && (hipMalloc(&S->X, (size_t)(h->cols > 0 ? h->cols : 1) * ldmax * 4) != hipSuccess ||
                          hipMalloc(&S->Y, (size_t)(h->rows > 0 ? h->rows : 1) * ldmax * 4) != hipSuccess ||
                          hipMalloc(&S->dd, (size_t)S->nslots * MAXQ * sizeof(PackDesc)) != hipSuccess ||
                          hipHostMalloc(&S->hd, (size_t)S->nslots * MAXQ * sizeof(PackDesc), hipHostMallocDefault) !=
                              hipSuccess))
        st = DCSR_ERR_OOM;
    S->slot_ev.assign(S->nslots, nullptr);
    S->slot_used.assign(S->nslots, 0);
    for (int i = 0; i < S->nslots && st == DCSR_OK; ++i)
        if (hipEventCreateWithFlags(&S->slot_ev[i], hipEventDisableTiming) != hipSuccess) st = DCSR_ERR_HIP;
    if (st != DCSR_OK) {
        dcsr_server_destroy(S);
        return st;
    }
    // pending work of the caller's streams on this matrix is finished first;
    // a matrix loaded with duplicates / explicit zeros is canonicalized now
    // (the first update would otherwise rewrite every row under running
    // queries)
    (void)hipDeviceSynchronize();
    if (!h->canonical) {
        if (!h->state_known) st = pull_state(h, S->su);
        if (st == DCSR_OK) st = canonicalize(h, S->su);
        (void)hipStreamSynchronize(S->su);
        if (st != DCSR_OK) {
            dcsr_server_destroy(S);
            return st;
        }
    }
    memset(&S->stats, 0, sizeof(S->stats));
    S->sched = std::thread(scheduler_loop, S);
    S->comp = std::thread(completer_loop, S);
    *out = S;
    return DCSR_OK;
}

dcsr_status dcsr_server_submit(dcsr_server S, const dcsr_request *r, uint64_t *ticket)
{
    if (!S || !r || !ticket) return DCSR_ERR_INVALID;
    if (r->kind == DCSR_REQ_UPDATE) {
        if (r->n < 0 || (r->n > 0 && (!r->rows || !r->cols || !r->vals))) return DCSR_ERR_INVALID;
    } else if (r->kind == DCSR_REQ_SPMV) {
        if (!r->x || !r->y) return DCSR_ERR_INVALID;
    } else if (r->kind == DCSR_REQ_SPMM) {
        if (!r->x || !r->y || r->k < 1 || r->ldx < r->k || r->ldy < r->k || r->k > S->max_cols)
            return DCSR_ERR_INVALID;
    } else {
        return DCSR_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) return DCSR_ERR_INVALID;
    Req q{*r, S->next_ticket++, Clock::now()};
    S->submitted[q.ticket] = q.t_submit;
    S->queue.push_back(q);
    *ticket = q.ticket;
    S->cv_sched.notify_one();
    return DCSR_OK;
}

dcsr_status dcsr_server_wait(dcsr_server S, uint64_t ticket, double *latency_ms)
{
    if (!S) return DCSR_ERR_INVALID;
    std::unique_lock<std::mutex> lk(S->mu);
    auto sub = S->submitted.find(ticket);
    if (sub == S->submitted.end()) return DCSR_ERR_INVALID;
    S->cv_done.wait(lk, [&] { return S->done.count(ticket) > 0; });
    const Done d = S->done[ticket];
    if (latency_ms) latency_ms[0] = std::chrono::duration<double, std::milli>(d.t_done - sub->second).count();
    S->done.erase(ticket);
    S->submitted.erase(sub);
    return d.status;
}

dcsr_status dcsr_server_drain(dcsr_server S)
{
    if (!S) return DCSR_ERR_INVALID;
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.wait(lk, [&] { return S->queue.empty() && S->inflight.empty() && S->done.size() == S->submitted.size(); });
    return DCSR_OK;
}

dcsr_status dcsr_server_get_stats(dcsr_server S, dcsr_server_stats *out)
{
    if (!S || !out) return DCSR_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->stats;
    return DCSR_OK;
}

dcsr_status dcsr_server_destroy(dcsr_server S)
{
    if (!S) return DCSR_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cv_sched.notify_all();
        S->cv_comp.notify_all();
    }
    if (S->sched.joinable()) S->sched.join();
    if (S->comp.joinable()) S->comp.join();
    if (S->sq) (void)hipStreamSynchronize(S->sq);
    if (S->su) (void)hipStreamSynchronize(S->su);
    if (S->have_query_ev) (void)hipEventDestroy(S->last_query_ev);
    for (hipEvent_t e : S->slot_ev)
        if (e) (void)hipEventDestroy(e);
    if (S->sq) (void)hipStreamDestroy(S->sq);
    if (S->su) (void)hipStreamDestroy(S->su);
    (void)hipFree(S->X);
    (void)hipFree(S->Y);
    (void)hipFree(S->dd);
    if (S->hd) (void)hipHostFree(S->hd);
    S->h->last_stream = nullptr;
    delete S;
    return DCSR_OK;
}

// lookback.h -- single-pass prefix machinery (decoupled look-back) for gfx950.
//
// A kernel that needs the exclusive prefix of per-tile aggregates takes its
// tile index from a ticket counter (so every lower tile has already started
// and will publish without waiting on anything), publishes its aggregate,
// looks back over its predecessors' published words until it meets an
// inclusive prefix, then publishes its own inclusive prefix.  A status word
// packs (epoch, flag) in the high half and the 32-bit value in the low half,
// so one 64-bit agent-scope atomic store publishes both and readers need no
// fence.  The epoch is a per-launch count kept on the device (ctl[3], bumped
// by the last workgroup of each launch, read by every workgroup at its
// start), so status arrays are never cleared, words from older launches read
// as "not ready", and replays of a captured graph get fresh epochs too.
//
// Every wait is bounded: after `limit` polls (LB_SPIN_LIMIT unless the host
// sets another) a tile gives up waiting and computes its exclusive prefix
// exactly from the raw inputs of every earlier tile instead (the caller's
// fallback functor: one strided pass of the wave over them, slow but never
// wrong and waiting for nothing), so no launch can hang and no consumer
// ever sees a short prefix.  The event is counted in ctl[2] of the launch
// (never reset) and raises LB_FALLBACK_BIT in the owner's sticky status word
// (`fault`, may be null), which the handle reports as a diagnostic.  A limit
// of 0 sends every tile past the first down the fallback without polling (a
// test of that path).
//
// Control words ctl[0..3] sit at the head of the launch's workspace:
// ctl[0] ticket, ctl[1] finished workgroups, ctl[2] fault count, ctl[3]
// epoch.  The ticket and finish counters are self-resetting (the last
// workgroup to finish a launch zeroes them); all start at zero because
// workspace buffers are zero-filled on allocation.  Launches that share a
// control block must be stream-ordered.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "wave.h"

#define LB_AGG 1u
#define LB_INC 2u
#define LB_SPIN_LIMIT (1 << 20)
// Sticky status bit raised by a tile that took the exact fallback (the
// handle's ST_SCAN_FALLBACK).
#define LB_FALLBACK_BIT 0x8u

__device__ __forceinline__ uint32_t lb_tag(uint32_t epoch) { return (epoch & 0x3fffffffu) << 2; }

__device__ __forceinline__ void lb_publish(uint64_t *p, uint32_t tag, uint32_t flag, uint32_t v)
{
    const uint64_t w = ((uint64_t)(tag | flag) << 32) | (uint64_t)v;
    __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t lb_load(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool lb_ready(uint64_t w, uint32_t tag)
{
    const uint32_t hi = (uint32_t)(w >> 32);
    return (hi & ~3u) == tag && (hi & 3u) != 0u;
}

__device__ __forceinline__ bool lb_inclusive(uint64_t w, uint32_t tag)
{
    const uint32_t hi = (uint32_t)(w >> 32);
    return (hi & ~3u) == tag && (hi & 3u) == LB_INC;
}

// Ticket of this workgroup and the launch's status tag (thread 0 draws and
// reads; broadcast through LDS, two words).  The epoch word was last written
// by the previous launch's last workgroup, which finished before this launch
// started, so every workgroup reads the same value.
__device__ __forceinline__ uint32_t lb_ticket(unsigned int *ctl, uint32_t *lds2, uint32_t *tag)
{
    if (threadIdx.x == 0) {
        lds2[0] = atomicAdd(&ctl[0], 1u);
        lds2[1] = __hip_atomic_load(&ctl[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    const uint32_t t = lds2[0];
    *tag = lb_tag(lds2[1]);
    __syncthreads();
    return t;
}

// Called by thread 0 of every workgroup once it has finished all its
// reads and writes of the launch's shared state: the last one out resets
// the counters and bumps the epoch for the next launch.  No fence: the
// ticket and finish counters are atomics, the status words are atomic
// stores read only within this launch (all before the last finish), and
// the kernel boundary publishes the reset to the next launch.  (A release
// fence here writes back the XCD's whole L2, dirty lines of earlier kernels
// included, in every workgroup.)  n: the workgroups of the launch that take
// part in the scan.
__device__ __forceinline__ void lb_finish_n(unsigned int *ctl, uint32_t n)
{
    if (atomicAdd(&ctl[1], 1u) == n - 1) {
        ctl[3] = ctl[3] + 1u;
        ctl[0] = 0u;
        ctl[1] = 0u;
    }
}

__device__ __forceinline__ void lb_finish(unsigned int *ctl) { lb_finish_n(ctl, gridDim.x); }

// Exclusive prefixes of tile `tile` (> 0) for NC components whose status
// words sit at st[j * NC + c] (j < tile), computed by one whole wave: 64
// predecessors are examined per round, nearest first, all components in the
// same round (each stops at its own nearest inclusive word).  pre[c] gets
// the same value in every lane.  After `limit` polls without progress the
// wave calls fb(c) instead, which must return (in every lane, called by all
// 64 lanes) component c's exact sum over every earlier tile's inputs.
template <int NC, typename Fallback>
__device__ __forceinline__ void lb_wave_lookback(const uint64_t *st, int64_t tile, uint32_t tag, unsigned int *ctl,
                                                 unsigned int *fault, int limit, Fallback &&fb, uint32_t *pre)
{
    const int lane = lane_id();
    bool done[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        pre[c] = 0;
        done[c] = false;
    }
    int64_t j = tile - 1;
    int spins = 0;
    bool gave_up = limit <= 0;  // 0: every tile takes the fallback (a test of it)
    while (j >= 0 && !gave_up) {
        const int64_t idx = j - lane;
        uint64_t w[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) w[c] = 0;
        for (;;) {
            bool ok = true;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                if (idx >= 0 && !done[c]) w[c] = lb_load(&st[idx * NC + c]);
                ok = ok && (idx < 0 || done[c] || lb_ready(w[c], tag));
            }
            if (__all(ok)) break;
            if (++spins > limit) {  // wave-uniform: every lane counts the same polls
                gave_up = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (gave_up) break;
        bool all_done = true;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            if (done[c]) continue;
            const uint64_t im = __ballot(idx >= 0 && lb_inclusive(w[c], tag));
            uint32_t v = idx >= 0 ? (uint32_t)w[c] : 0u;
            if (im) {
                const int L = __ffsll((long long)im) - 1;
                if (lane > L) v = 0u;
                done[c] = true;
            }
            pre[c] += wave_sum_u32(v);
            all_done = all_done && done[c];
        }
        if (all_done) break;
        j -= 64;
    }
    if (gave_up) {
        if (lane == 0) {
            atomicAdd(&ctl[2], 1u);
            if (fault) atomicOr(fault, LB_FALLBACK_BIT);
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) pre[c] = fb(c);
    }
}

// Tile prefix of NC components in one call: thread 0 publishes the tile's
// aggregates s_agg[0..NC) (LDS, written and synchronised by the caller),
// wave 0 resolves the exclusive prefixes into s_pre[0..NC) (LDS) and
// publishes the inclusive ones.  Ends with a __syncthreads.
template <int NC, typename Fallback>
__device__ __forceinline__ void lb_tile_prefix(const uint32_t *s_agg, uint32_t *s_pre, uint64_t *st, int64_t tile,
                                               uint32_t tag, unsigned int *ctl, unsigned int *fault, int limit,
                                               Fallback &&fb)
{
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NC; ++c) lb_publish(&st[tile * NC + c], tag, tile == 0 ? LB_INC : LB_AGG, s_agg[c]);
    }
    if (threadIdx.x < 64) {
        uint32_t p[NC];
        if (tile > 0) {
            lb_wave_lookback<NC>(st, tile, tag, ctl, fault, limit, fb, p);
        } else {
#pragma unroll
            for (int c = 0; c < NC; ++c) p[c] = 0u;
        }
        if (threadIdx.x == 0) {
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                s_pre[c] = p[c];
                if (tile > 0) lb_publish(&st[tile * NC + c], tag, LB_INC, p[c] + s_agg[c]);
            }
        }
    }
    __syncthreads();
}

// Exact sum, in every lane, of the uint32 words src[i * stride] for i < n,
// by one whole wave (the look-back fallback: wrapping arithmetic, as the
// scans themselves).
__device__ __forceinline__ uint32_t lb_wave_sum_prefix(const uint32_t *src, int64_t n, int stride)
{
    uint32_t s = 0;
    for (int64_t i = lane_id(); i < n; i += 64) s += src[i * stride];
    return wave_sum_u32(s);
}

// Poll limit of the look-back waits: LB_SPIN_LIMIT, or the value set by
// dcsr_set_scan_wait_limit (host side, process-wide).
int lb_spin_limit();

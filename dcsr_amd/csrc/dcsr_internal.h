// dcsr_internal.h -- handle layout, device state and shared host helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include "../../include/dcsr.h"

// Sticky device-side status bits.
#define ST_BAD_INDEX 0x1u
#define ST_POOL_EXHAUSTED 0x2u
#define ST_SEGMENT_OVERFLOW 0x4u
// a single-pass scan's bounded look-back wait gave up and the tile computed
// its prefix by the exact fallback (lookback.h LB_FALLBACK_BIT): results are
// correct, the bit is a diagnostic (dcsr_info.scan_fallback) and survives
// status reads
#define ST_SCAN_FALLBACK 0x8u

// Rows longer than this are counted as power-law hubs in dcsr_info.
#define DCSR_HUB_LEN 4096
// SpMV calls on one structure that pay for building its paneled copy
#define DCSR_PNL_REUSE 4
#define DCSR_HUB_CHUNK 8192
// Nonzeros owned by one wave in the flat SpMV / SpMM kernels.
#define DCSR_SPMV_PW 256u

// Pool chunk granularity: chunk starts stay 16-byte aligned.
#define DCSR_CHUNK_ALIGN 4u

// Device-resident counters of one handle.  Written by kernels, read back by
// the host at the documented sync points.
struct DevState {
    unsigned long long bump;        // next free pool slot
    unsigned long long bump_saved;  // bump before the current update's claims
    long long nnz;                  // live entries
    long long garbage;              // entries in retired chunks
    long long d_nnz;                // current update: inserted - deleted
    long long d_garbage;            // current update: retired capacity
    long long n_ins, n_add, n_del;  // current update counters
    long long d_hubs;               // current update: hub-row count delta
    long long d_hub_chunks;         // current update: hub chunk delta
    unsigned long long need;        // current update: claimed capacity
    unsigned int status;            // sticky ST_* bits
    unsigned int go;                // 1 if the current update may commit
    unsigned int nuniq;             // distinct coordinates in the batch
    unsigned int ntouched;          // rows touched
    unsigned int claim_total;       // pool slots claimed by rewritten rows
    unsigned int nrewrite_old;      // old entries moved
    unsigned int hub_count;         // entries in the hub list
    unsigned int hub_chunks;        // total hub chunks
    unsigned int ubad;              // current batch holds an out-of-range index
    unsigned int hub_touch;         // current update: rewritten rows that were/are hubs
    unsigned int ubad_last;         // current update's batch held an out-of-range index
    unsigned int ptotal;            // current update: probe lengths of the touched rows, summed
};

// Grow-only device scratch buffers owned by a handle.
enum WsSlot {
    WS_KEY_A, WS_KEY_B, WS_VAL_A, WS_VAL_B, WS_FLAGS, WS_SCAN, WS_PART, WS_HIST,
    WS_UCOL, WS_UVAL, WS_UT, WS_UPOS, WS_UCODE, WS_UPRE,
    WS_TROW, WS_TUBEG, WS_TINS, WS_TDEL, WS_TOLDS, WS_TMOVED, WS_TDEST, WS_TNEWL, WS_TMODE, WS_TOLDPRE,
    WS_HUB_ROWS, WS_HUB_PRE, WS_HUB_PART, WS_HUB_DONE,
    WS_PPROBE, WS_PPLAN, WS_UROW, WS_TPAIRPRE, WS_FOFF, WS_ROWMAP, WS_CARRY_ROW, WS_CARRY_VAL, WS_SPMM_CROW, WS_SPMM_CVAL,
    WS_SG_P, WS_SG_POFF, WS_SG_CNT, WS_SG_COFF, WS_SG_LIST, WS_SG_STATE, WS_SG_BSUM,
    WS_SG_HPRE, WS_SG_EA, WS_SG_APOS, WS_SG_KEY0, WS_SG_KEY1, WS_SG_VAL0, WS_SG_VAL1, WS_SG_FLAG, WS_SG_EPOS,
    WS_CANON, WS_TMP0, WS_TMP1, WS_TMP2,
    WS_PNL_OFF, WS_PNL_BND, WS_PNL_COL, WS_PNL_VAL, WS_PNL_MAP, WS_PNL_PART,
    WS_TRS, WS_TLEN, WS_PSPAN, WS_MSPAN, WS_TCAP, WS_ROWT,
    WS_COUNT
};

// A NOSYNC update's state snapshot, stored by its write kernel (seq last).
struct HostSnap {
    uint64_t nnz, bump, seq;
};
constexpr int SNAP_RING = 64;

struct Workspace {
    void *ptr[WS_COUNT];
    size_t bytes[WS_COUNT];
};

struct dcsr_matrix_s {
    int64_t rows, cols;
    int rbits, cbits;        // key = (row << cbits) | col
    // row table
    uint32_t *row_start;
    int32_t *row_len;
    int32_t *row_cap;
    // pool
    int32_t *pool_col;
    float *pool_val;
    int64_t pool_cap;
    // second pool of the same layout: compaction copies into it and swaps
    int32_t *spare_col;
    float *spare_val;
    int64_t spare_cap;
    // spare buffers replaced by larger ones while work may still be queued
    // on them: freed at the next point where the handle's stream is known
    // idle (a state read-back) or at destroy, never with a sync of their own
    void *retired[32];
    int nretired;
    // device/host state
    DevState *ds;            // device
    DevState *hs;            // pinned host mirror
    int64_t nnz_h, bump_h, garbage_h, hub_rows_h, hub_chunks_h;
    // upper bounds of the live entries and of the bump pointer, valid also
    // while state_known is false (a NOSYNC update of n triples adds at most n
    // entries and claims at most its claim bound): they size SpMV grids and
    // decide pool room without reading the device state back
    int64_t nnz_ub, bump_ub;
    // state snapshot of each NOSYNC update: its write kernel stores the live
    // count and bump pointer it leaves, then the update's sequence number,
    // into coherent host memory (no copy, no event).  The host keeps the
    // running sums of the bounds it added per update (cum_*, and their
    // values after update k in ring slot k % SNAP_RING), so a landed
    // snapshot k bounds the current state by its values plus what was added
    // after k.  Snapshots before snap_valid_from predate a change the sums
    // do not count (a read-back, load or compaction) and are ignored.
    HostSnap *hsnap;         // host view (pinned, coherent)
    HostSnap *dsnap;         // the same memory as the device addresses it
    uint64_t snap_seq;       // last sequence number handed to an update
    uint64_t snap_valid_from;
    uint64_t snap_used;      // last snapshot folded into the bounds
    int64_t cum_add, cum_claim;
    int64_t ring_add[SNAP_RING], ring_claim[SNAP_RING];
    bool canonical;
    bool hubs_dirty;
    int64_t ncompact;        // compactions and growths (dcsr_query reports it)
    uint32_t upd_tag;        // tag of the last update's row -> touched-row map
    bool flat_dirty;         // flat row offsets (SpMV binning) need a rebuild
    bool state_known;        // host mirrors valid (no NOSYNC update since last sync)
    // column-paneled copy for SpMV on matrices whose x exceeds the L2 (see
    // spmv.hip): valid until the next structural change
    bool pnl_valid;
    int pnl_calls;           // SpMV calls since the last structural change
    int pnl_prev_calls;      // the same for the last structure that had any
    int pnl_count;           // panels in the current copy
    int64_t pnl_cols;        // panel width (columns)
    int64_t pnl_min_nnz;     // smallest nnz worth a paneled copy
    dcsr_config cfg;
    Workspace ws;
    hipStream_t last_stream;
    int64_t last_ins, last_add, last_del, last_touched;
};

// ---------------------------------------------------------------------------
// host helpers (defined in matrix.hip)
// ---------------------------------------------------------------------------
#define DCSR_HIP(call)                                                     \
    do {                                                                   \
        hipError_t _e = (call);                                            \
        if (_e != hipSuccess) return dcsr_hip_fail(_e, #call, __FILE__, __LINE__); \
    } while (0)

#define DCSR_TRY(call)                    \
    do {                                  \
        dcsr_status _s = (call);          \
        if (_s != DCSR_OK) return _s;     \
    } while (0)

dcsr_status dcsr_hip_fail(hipError_t e, const char *what, const char *file, int line);
dcsr_status ws_get(dcsr_handle h, WsSlot slot, size_t bytes, void **out);
// For slots whose words must start at zero (look-back control and status
// words, finish counters): a grown buffer is zero-filled by a kernel on s.
dcsr_status ws_get_zeroed(dcsr_handle h, WsSlot slot, size_t bytes, void **out, hipStream_t s);
dcsr_status pull_state(dcsr_handle h, hipStream_t s);   // copy DevState to host + sync
// Grow the pool in stream order, without reading the device state back:
// the live rows are compacted into a new pool of at least new_cap entries.
dcsr_status grow_in_order(dcsr_handle h, int64_t new_cap, hipStream_t s);
// After launches whose device-side results the host does not read back: at
// most `added` entries were inserted and `claimed` pool slots claimed.
void state_unknown(dcsr_handle h, int64_t added, int64_t claimed);
// The host mirrors are exact (after a read-back, or after work whose result
// the host computed itself).
void state_exact(dcsr_handle h);
// Enqueue an asynchronous copy of the device state (NOSYNC paths), and fold
// a completed one into the bounds (never waits).
uint64_t snapshot_next(dcsr_handle h);
void snapshot_note(dcsr_handle h, uint64_t seq);
void refresh_bounds(dcsr_handle h);
// Upper bounds of the live entries and of the bump pointer (exact when the
// state is known).
static inline int64_t live_bound(dcsr_handle h) { return h->state_known ? h->nnz_h : h->nnz_ub; }
static inline int64_t bump_bound(dcsr_handle h) { return h->state_known ? h->bump_h : h->bump_ub; }
dcsr_status ensure_flat(dcsr_handle h, hipStream_t s);  // spmv.hip
// The SpMV's flat offsets, row map and row-scan look-back block (sized for
// the current pool), and the bookkeeping once a stream-ordered scan has
// rebuilt them (the update's move kernel can run that scan).
struct FlatBufs {
    uint32_t *foff, *rowmap, *part;
};
dcsr_status flat_buffers(dcsr_handle h, hipStream_t s, FlatBufs *fb);  // spmv.hip
void flat_rebuilt(dcsr_handle h);                                      // spmv.hip
dcsr_status pool_reserve(dcsr_handle h, int64_t min_free, hipStream_t s, const int32_t *reserve_len);
dcsr_status canonicalize(dcsr_handle h, hipStream_t s);
// Compaction at the current capacity, in stream order, without a read-back.
dcsr_status compact_in_order(dcsr_handle h, hipStream_t s);
dcsr_status alloc_pool(dcsr_handle h, int64_t cap);
// update.hip: the two phases of dcsr_update (prepare only reads the matrix)
dcsr_status update_prepare(dcsr_handle h, const int32_t *rows, const int32_t *cols, const float *vals, int64_t n,
                           hipStream_t s, bool *go, int64_t *need);
dcsr_status update_replan(dcsr_handle h, int64_t n, hipStream_t s, bool *go);
dcsr_status update_commit(dcsr_handle h, int64_t n, hipStream_t s, bool nosync);
// update.hip: an update with its scratch carved from a fixed device region
// (no allocation, no pool growth or compaction; canonical matrices only),
// and the bytes such an update of n triples needs.
dcsr_status update_fixed_scratch(dcsr_handle h, const int32_t *rows, const int32_t *cols, const float *vals,
                                 int64_t n, uint32_t flags, hipStream_t s, uint8_t *base, size_t cap);
size_t update_scratch_bytes(dcsr_handle h, int64_t n);
// Flat copy of R rows (src_start[r], length dst_off[r+1]-dst_off[r]) to dst_off.
void launch_copy_rows(const uint32_t *dst_off, int64_t R, int64_t total, const uint32_t *src_start,
                      const int32_t *scol, const float *sval, int32_t *dcol, float *dval, hipStream_t s);

// The handle's sticky status word (device): scans set ST_SCAN_FALLBACK in it.
static inline unsigned int *fault_word(dcsr_handle h) { return &h->ds->status; }

static inline int bits_for(int64_t n)
{
    int b = 1;
    while ((int64_t(1) << b) < n) ++b;
    return b;
}

static inline unsigned grid_for(int64_t n, int per_block)
{
    int64_t g = (n + per_block - 1) / per_block;
    return (unsigned)(g < 1 ? 1 : g);
}

__host__ __device__ static inline uint32_t chunk_round(uint32_t len)
{
    return (len + DCSR_CHUNK_ALIGN - 1u) & ~(DCSR_CHUNK_ALIGN - 1u);
}

/*
 * dcsr.h -- C ABI of the MI355X Dynamic-CSR engine (libdcsr.so, gfx950).
 *
 * Drop-in boundary for the reference's header-only C++ API in namespace
 * `device` (include/load.h, include/spmv.h, include/spmm.h of the reference).
 * Each entry point names the reference function it replaces.
 *
 * Conventions
 *   - Matrices are opaque handles; the library owns every device array of a
 *     handle (reference: dcsr_matrix owns row_sizes/row_offsets/pool,
 *     include/dcsr_matrix.h:8-39).
 *   - Index type int32, value type float32: the reference's instantiation
 *     (src/Fill_Test.inl:11-16 with PRECISION 32).
 *   - Every array argument is a DEVICE pointer unless stated otherwise.
 *   - `stream` is a hipStream_t (NULL = the null stream).  Calls are
 *     stream-ordered.  Entry points documented as "syncs" wait for the stream
 *     to read a device-side count back (pool accounting, output sizes).
 *   - Every entry point returns a dcsr_status; the reference returns void and
 *     prints errors (include/sparse.h:24-30).  Errors never leave a matrix
 *     half-updated: a rejected update is a no-op.
 *   - Threading: a handle may be used by one host thread at a time.
 *
 * Storage model: every row is one column-sorted segment
 * [start, start+len) inside a pooled (col, val) store with capacity cap >= len;
 * structural updates write the new row to a chunk claimed from the pool's
 * bump region and retire the old chunk; compaction rewrites all rows
 * contiguously in natural row order.
 */
#ifndef DCSR_H
#define DCSR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dcsr_matrix_s *dcsr_handle;
typedef void *dcsr_stream_t; /* hipStream_t */

typedef enum dcsr_status {
    DCSR_OK = 0,
    DCSR_ERR_INVALID = 1,        /* bad argument (null handle, negative size, size mismatch) */
    DCSR_ERR_BAD_INDEX = 2,      /* a row/column index outside the matrix */
    DCSR_ERR_POOL_EXHAUSTED = 3, /* the update needs more pool than reserved and growth is disabled */
    DCSR_ERR_SEGMENT_OVERFLOW = 4, /* a row would exceed the per-row limit (2^31-1) */
    DCSR_ERR_HIP = 5,            /* a HIP runtime call failed */
    DCSR_ERR_OOM = 6,            /* device allocation failed */
    DCSR_ERR_CAPACITY = 7,       /* caller-provided output buffer too small */
    DCSR_ERR_UNSUPPORTED = 8,
    DCSR_ERR_TOO_LARGE = 9       /* request exceeds the fixed scratch of dcsr_update_noalloc */
} dcsr_status;

/* Flags for dcsr_update. */
#define DCSR_UPDATE_NOSYNC   0x1u /* asynchronous: no read-back of the batch's
                                     outcome.  Pool room comes from host-side
                                     upper bounds (compaction in stream order,
                                     or one read-back and growth, when they
                                     cannot rule out exhaustion).  A batch
                                     with a bad index, or one that does not
                                     fit under NOGROW / allow_grow = 0,
                                     becomes a no-op and its error is
                                     reported by the next dcsr_get_status */
#define DCSR_UPDATE_NOGROW   0x2u /* never grow/compact automatically */
/* Flags for dcsr_export_csr. */
#define DCSR_EXPORT_CANONICAL 0x1u /* merge duplicate columns and drop exact
                                      zeros (only differs before the first
                                      update of a matrix loaded with duplicates
                                      or explicit zeros) */

typedef struct dcsr_config {
    double pool_factor;  /* pool capacity = pool_factor * nnz at load (default 2.0) */
    int64_t min_pool;    /* minimum pool entries (default 1024) */
    int32_t allow_grow;  /* grow the pool when an update needs more (default 1) */
    int32_t reserved;
} dcsr_config;

typedef struct dcsr_info {
    int64_t rows, cols;
    int64_t nnz;          /* live entries */
    int64_t pool_cap;     /* pool capacity (entries) */
    int64_t pool_used;    /* bump pointer: entries handed out since the last compaction */
    int64_t garbage;      /* entries in retired chunks (reclaimed by compaction) */
    int64_t hub_rows;     /* rows longer than the hub threshold */
    int64_t last_inserted; /* last update: new coordinates stored */
    int64_t last_added;    /* last update: existing coordinates updated in place */
    int64_t last_deleted;  /* last update: coordinates removed (sum == 0) */
    int64_t last_touched_rows;
    int32_t canonical;    /* 1 if rows hold no duplicate columns and no exact zeros */
    int32_t max_segments; /* segments per row: always 1 in this layout */
    int32_t scan_fallback; /* 1 if a single-pass scan on this handle ever gave up
                              its bounded look-back wait and took its exact
                              (slow) fallback: results stay correct; 0 in a
                              healthy run */
    int32_t spmv_paneled; /* 1 while SpMV runs over the column-paneled copy */
    int64_t compactions;  /* pool compactions and growths on this handle, the
                             explicit dcsr_compact included */
} dcsr_info;

const char *dcsr_status_string(dcsr_status s);
/* Sticky device-side status of the handle (syncs the handle's last stream),
 * then clears it. */
dcsr_status dcsr_get_status(dcsr_handle h);

/* Container setup.  Replaces dcsr_matrix::resize (include/dcsr_matrix.h:41-73,
 * 76-108) and the owning constructor/destructor (:33-39).  cfg may be NULL. */
dcsr_status dcsr_create(dcsr_handle *out, int64_t rows, int64_t cols, const dcsr_config *cfg);
dcsr_status dcsr_destroy(dcsr_handle h);

/* Uniform preallocation: every row empty with capacity bin_length.
 * Replaces device::Initialize_Matrix -> InitializeMatrix_dcsr
 * (include/load.h:8-22, include/sparse_update.inl:1155-1176). */
dcsr_status dcsr_initialize(dcsr_handle h, int32_t bin_length, dcsr_stream_t stream);

/* COO -> DCSR.  Triples may be unsorted; entries of a row keep their input
 * order among equal columns (duplicates and explicit zeros are kept, as the
 * host COO->CSR conversion keeps them).  Syncs.
 * Replaces device::LoadMatrix(dcsr&, rows, cols, vals, N) (include/load.h:27-59),
 * which additionally required row-sorted input. */
dcsr_status dcsr_load_coo(dcsr_handle h, const int32_t *rows, const int32_t *cols,
                          const float *vals, int64_t n, dcsr_stream_t stream);

/* CSR -> DCSR.  Syncs.  Replaces device::LoadMatrix(dcsr&, const csr_matrix&)
 * (include/load.h:207-227). */
dcsr_status dcsr_load_csr(dcsr_handle h, const int32_t *row_offsets, const int32_t *cols,
                          const float *vals, dcsr_stream_t stream);

/* DCSR -> CSR in natural row order, columns ascending inside every row.
 * Output arrays hold `cap` entries; *nnz_out receives the entry count (host
 * pointer).  Syncs.  Replaces device::LoadMatrix(csr_matrix&, const dcsr&)
 * (include/load.h:232-253), which was only correct for a freshly compacted
 * matrix. */
dcsr_status dcsr_export_csr(dcsr_handle h, int32_t *row_offsets, int32_t *cols, float *vals,
                            int64_t cap, int64_t *nnz_out, uint32_t flags, dcsr_stream_t stream);

/* In-place insert-or-add of a batch of n (row, col, val) triples in any order.
 * Result equals the host path "sort batch by row, COO->CSR, cusp::add(A, B)":
 * duplicate batch coordinates are summed in batch order, the sum is added to
 * the stored value, and an exact-zero result removes the coordinate.
 * Replaces device::UpdateMatrix(dcsr&, rows, cols, vals, permuted)
 * (include/load.h:258-325), which appended without merging.
 * Syncs once unless DCSR_UPDATE_NOSYNC. */
dcsr_status dcsr_update(dcsr_handle h, const int32_t *rows, const int32_t *cols,
                        const float *vals, int64_t n, uint32_t flags, dcsr_stream_t stream);

/* Defragmentation: rewrite every row contiguously in natural row order and
 * release retired chunks.  Replaces device::SortMatrixRow (include/load.h:404-581). */
dcsr_status dcsr_compact(dcsr_handle h, dcsr_stream_t stream);

/* Rebuild what SpMV/SpMM/SpGEMM derive from the row lengths: the flat
 * exclusive offsets of the live rows, the 64-nonzero granule -> row map and
 * the count of hub rows (longer than 4096).  Replaces device::BinRows ->
 * SetBins (include/load.h:583-610); the compute entry points call it
 * implicitly after any structural change, so calling it is optional. */
dcsr_status dcsr_rebin(dcsr_handle h, dcsr_stream_t stream);

/* y = alpha * A x + beta * y  (beta == 0: y is overwritten, never read).
 * Replaces device::spmv(const dcsr&, x, y, streams) (include/spmv.h:121-248). */
dcsr_status dcsr_spmv(dcsr_handle h, const float *x, float *y, float alpha, float beta,
                      dcsr_stream_t stream);

/* Y = alpha * A X + beta * Y with a dense multi-column right-hand side:
 * X is cols x k row-major with leading dimension ldx >= k, Y is rows x k
 * row-major with ldy >= k (beta == 0: Y is overwritten, never read).
 * No reference counterpart (the reference has SpMV only); each column equals
 * dcsr_spmv on that column up to FP32 summation order. */
dcsr_status dcsr_spmm_dense(dcsr_handle h, const float *X, int64_t k, int64_t ldx, float *Y, int64_t ldy,
                            float alpha, float beta, dcsr_stream_t stream);

/* C = A * B (SpGEMM).  C must be a distinct handle with C.rows == A.rows and
 * C.cols == B.cols (A.cols == B.rows); its previous contents are replaced.
 * Structure: every (i, k) with a partial product whose FP32 sum, accumulated
 * from 0 in A-row order then B-row order, is not exactly zero (host path:
 * cusp/system/detail/sequential/multiply/csr_spgemm.h:75-150); rows come out
 * column-sorted.  Rows with more than 4096 partial products, which the
 * reference skips, are included.  Syncs.
 * Replaces device::spmm(const dcsr& A, const dcsr& B, dcsr& C, streams)
 * (include/spmm.h:30-391). */
dcsr_status dcsr_spgemm(dcsr_handle A, dcsr_handle B, dcsr_handle C, dcsr_stream_t stream);

/* DCSR -> HYB (ELL + COO).  The ELL block holds `width` slots per row in
 * column-major order (slot k of row r at k * pitch + r, pitch >= rows; padding
 * column -1, value 0); entries past `width` go to the COO arrays (capacity
 * coo_cap) in row order; *coo_n (host) receives their count.  Rows are taken
 * with columns ascending, as the host CSR -> HYB conversion with the same
 * width (cusp/system/detail/generic/conversions/csr_to_other.h:226-303).
 * Syncs.  The reference converts only COO/CSR -> HYB
 * (include/load.h:61-106, 174-205); HYB <-> DCSR went through CSR. */
dcsr_status dcsr_export_hyb(dcsr_handle h, int32_t width, int64_t pitch, int32_t *ell_col, float *ell_val,
                            int32_t *coo_row, int32_t *coo_col, float *coo_val, int64_t coo_cap, int64_t *coo_n,
                            dcsr_stream_t stream);

/* HYB -> DCSR: every ELL slot with a column != -1 (slot order), then the COO
 * part, loaded as by dcsr_load_coo (duplicates kept).  Syncs. */
dcsr_status dcsr_load_hyb(dcsr_handle h, int32_t width, int64_t pitch, const int32_t *ell_col, const float *ell_val,
                          const int32_t *coo_row, const int32_t *coo_col, const float *coo_val, int64_t coo_n,
                          dcsr_stream_t stream);

/* ---------------------------------------------------------------------------
 * Batches of small matrices (no reference counterpart: the reference keeps
 * one process-global bin table, include/Matrix_Test.h:57, and so handles one
 * matrix at a time).  Matrix m has rows[m] x cols[m] (each <= 4096) and uses
 * local indices.  A step applies every matrix's own insert-or-add batch (at
 * most 4096 triples per matrix; semantics of dcsr_update) and then its SpMV,
 * all in ONE persistent launch; with DCSR_BATCH_GRAPH the launch is captured
 * once into a hipGraph and replayed on later calls with the same arguments.
 * Triples of all matrices are concatenated; tri_off (device, count+1) gives
 * each matrix's range.  x / y concatenate the matrices' vectors in order.
 * ------------------------------------------------------------------------- */
typedef struct dcsr_batch_s *dcsr_batch;
#define DCSR_BATCH_GRAPH 0x1u

dcsr_status dcsr_batch_create(dcsr_batch *out, int32_t count, const int32_t *rows, const int32_t *cols,
                              const dcsr_config *cfg); /* rows/cols: host arrays */
dcsr_status dcsr_batch_destroy(dcsr_batch b);
/* The block-diagonal handle holding the batch (matrix m = global rows
 * [row_base[m], row_base[m+1]) with local columns): export / query only. */
dcsr_handle dcsr_batch_handle(dcsr_batch b);
/* Host arrays of count+1 offsets; current SpMV shape (lanes per row, long-row
 * threshold).  Any pointer may be NULL. */
dcsr_status dcsr_batch_info(dcsr_batch b, int64_t *row_base, int64_t *col_base, int32_t *team, int32_t *thresh);
dcsr_status dcsr_batch_load_coo(dcsr_batch b, const int64_t *tri_off, const int32_t *rows, const int32_t *cols,
                                const float *vals, int64_t n, dcsr_stream_t stream);
/* One step: update (skipped when tri_off is NULL or n == 0) then SpMV
 * (skipped when x or y is NULL).  Syncs once to check pool headroom. */
dcsr_status dcsr_batch_step(dcsr_batch b, const int64_t *tri_off, const int32_t *rows, const int32_t *cols,
                            const float *vals, int64_t n, const float *x, float *y, uint32_t flags,
                            dcsr_stream_t stream);
/* OR of the per-matrix status of the last step (syncs); per_matrix (host,
 * count entries) may be NULL. */
dcsr_status dcsr_batch_status(dcsr_batch b, uint32_t *per_matrix);
/* Autotuner: times the step's SpMV for lanes-per-row in {4,8,16,32} and
 * long-row thresholds {64,256,1024} and keeps the fastest; choices are cached
 * per shape class (matrix count, mean row length) for the process. */
dcsr_status dcsr_batch_autotune(dcsr_batch b, const float *x, float *y, int32_t *team_out, int32_t *thresh_out,
                                dcsr_stream_t stream);
dcsr_status dcsr_batch_set_shape(dcsr_batch b, int32_t team, int32_t thresh);

/* ---------------------------------------------------------------------------
 * Serving loop (no reference counterpart).  A server owns a resident matrix
 * (use only the server API on that handle while the server lives) and runs a
 * scheduler thread with continuous batching: queries (SpMV / SpMM) queued
 * together between two updates are answered by one multi-column SpMM on the
 * same matrix version (at most max_cols columns per group); an update's
 * prepare phase (reads the matrix only) overlaps the previous query group and
 * its commit is ordered after it; every query sees exactly the updates
 * submitted before it.  Updates that need more pool than is free wait for
 * in-flight queries, then the pool is compacted / grown.  Request arrays are
 * device pointers whose contents must be complete when submitted (the server
 * reads them on its own streams) and that must stay valid until the request
 * completes.
 * ------------------------------------------------------------------------- */
typedef struct dcsr_server_s *dcsr_server;
enum { DCSR_REQ_UPDATE = 0, DCSR_REQ_SPMV = 1, DCSR_REQ_SPMM = 2 };
typedef struct dcsr_request {
    int32_t kind;          /* DCSR_REQ_* */
    int32_t k;             /* SpMM columns */
    int64_t n;             /* update: triples */
    const int32_t *rows;   /* update */
    const int32_t *cols;
    const float *vals;
    const float *x;        /* SpMV: x[cols]; SpMM: X (cols x k, row-major, ldx) */
    float *y;              /* SpMV: y[rows]; SpMM: Y (rows x k, row-major, ldy) */
    int64_t ldx, ldy;
} dcsr_request;
typedef struct dcsr_server_stats {
    int64_t updates, update_triples, queries, query_groups, compactions;
} dcsr_server_stats;

dcsr_status dcsr_server_create(dcsr_server *out, dcsr_handle h, int32_t max_cols);
/* Enqueue; *ticket identifies the request.  Never blocks on the GPU. */
dcsr_status dcsr_server_submit(dcsr_server s, const dcsr_request *r, uint64_t *ticket);
/* Block until the request finished; returns its status and (host pointer,
 * may be NULL) the submit-to-completion latency in milliseconds. */
dcsr_status dcsr_server_wait(dcsr_server s, uint64_t ticket, double *latency_ms);
dcsr_status dcsr_server_drain(dcsr_server s);
dcsr_status dcsr_server_get_stats(dcsr_server s, dcsr_server_stats *out);
dcsr_status dcsr_server_destroy(dcsr_server s);

/* ---------------------------------------------------------------------------
 * Dynamic HYB matrix (ELL + COO overflow), the reference's hyb_matrix
 * wrapper (include/dcsr_matrix.h:148-172) with its load, append-update and
 * SpMV.  Layout: `width` ELL slots per row, column-major with pitch = rows
 * rounded up to 32 (slot k of row r at k * pitch + r; padding column -1,
 * value 0); COO overflow sorted by row; row_sizes[r] = entries of row r
 * (ELL + COO), row_sizes[rows] = COO count.  The view's device pointers stay
 * valid until the next load / update of the store (an update may move the
 * COO part to its other buffer).
 * ------------------------------------------------------------------------- */
typedef struct dcsr_hyb_s *dcsr_hyb;
typedef struct dcsr_hyb_view {
    int64_t rows, cols, width, pitch;
    int32_t *ell_col;     /* width * pitch */
    float *ell_val;
    int32_t *row_sizes;   /* rows + 1 */
    int32_t *coo_row, *coo_col;
    float *coo_val;
    int64_t num_overflow; /* COO entries */
    int64_t coo_cap;
} dcsr_hyb_view;

/* An empty store; coo_cap is the initial COO capacity (grown on demand).
 * Replaces hyb_matrix::resize(rows, cols, num_coo_entries, num_cols_per_row)
 * (include/dcsr_matrix.h:160-171). */
dcsr_status dcsr_hyb_create(dcsr_hyb *out, int64_t rows, int64_t cols, int32_t width, int64_t coo_cap);
dcsr_status dcsr_hyb_destroy(dcsr_hyb h);
/* Device pointers and sizes of the store (no sync). */
dcsr_status dcsr_hyb_view_get(dcsr_hyb h, dcsr_hyb_view *v);
/* COO -> HYB: stable by row, the first `width` entries of each row (input
 * order) in its ELL slots, the rest in COO in row order; an out-of-range
 * index rejects the load (no-op, DCSR_ERR_BAD_INDEX).  Syncs.  Replaces
 * LoadMatrix(hyb&, rows, cols, vals, N) (include/load.h:61-106). */
dcsr_status dcsr_hyb_load_coo(dcsr_hyb h, const int32_t *rows, const int32_t *cols, const float *vals, int64_t n,
                              dcsr_stream_t stream);
/* CSR -> HYB, same placement.  Syncs.  Replaces LoadMatrix(hyb&, const csr&)
 * (include/load.h:174-205). */
dcsr_status dcsr_hyb_load_csr(dcsr_hyb h, const int32_t *row_offsets, const int32_t *cols, const float *vals,
                              dcsr_stream_t stream);
/* Append update: in stable row order each triple of row r fills ELL slot
 * row_sizes[r] while that is below the width, else it goes to r's COO
 * overflow after r's earlier entries; no merging (a coordinate may be stored
 * several times; the matrix is the sum of its entries).  Syncs once (COO
 * capacity).  Replaces UpdateMatrix(hyb&, rows, cols, vals, N)
 * (include/load.h:327-373 -> UpdateMatrix_hyb, include/sparse_update.inl:318-367). */
dcsr_status dcsr_hyb_update(dcsr_hyb h, const int32_t *rows, const int32_t *cols, const float *vals, int64_t n,
                            dcsr_stream_t stream);
/* y = alpha * A x + beta * y; per row the ELL slots in slot order then the
 * COO entries in order are added left to right from 0, the host hyb
 * multiply's order (cusp/system/detail/sequential/multiply/hyb_spmv.h:42-58).
 * Replaces cusp::multiply(hyb, x, y) as the reference's driver calls it
 * (src/Fill_Test.inl:167-170). */
dcsr_status dcsr_hyb_spmv(dcsr_hyb h, const float *x, float *y, float alpha, float beta, dcsr_stream_t stream);
/* Copy the store into caller buffers (ELL width * pitch, row_sizes rows + 1,
 * COO up to coo_cap; *ncoo receives the COO count).  Syncs. */
dcsr_status dcsr_hyb_export(dcsr_hyb h, int32_t *ell_col, float *ell_val, int32_t *row_sizes, int32_t *coo_row,
                            int32_t *coo_col, float *coo_val, int64_t coo_cap, int64_t *ncoo, dcsr_stream_t stream);
/* HYB -> DCSR (every stored entry; duplicates kept, as dcsr_load_hyb) and
 * DCSR -> HYB (the rows with columns ascending).  Sync. */
dcsr_status dcsr_hyb_to_dcsr(dcsr_hyb h, dcsr_handle dst, dcsr_stream_t stream);
dcsr_status dcsr_hyb_from_dcsr(dcsr_hyb h, dcsr_handle src, dcsr_stream_t stream);

/* Counters and sizes of the handle (syncs). */
dcsr_status dcsr_query(dcsr_handle h, dcsr_info *info);

/* The row table of the pool allocator: for each row its chunk start, live
 * length and chunk capacity (device arrays of `rows` entries).  With
 * dcsr_query's pool_used (bump pointer), garbage and pool_cap it exposes the
 * allocator's invariants: chunks [start, start + cap) are disjoint and lie
 * below the bump pointer, len <= cap, bump <= pool_cap, and bump equals the
 * sum of the capacities plus the garbage.  Syncs.  Diagnostic; the
 * reference's row_offsets / row_sizes arrays (include/dcsr_matrix.h:54-57)
 * were public members. */
dcsr_status dcsr_export_layout(dcsr_handle h, uint32_t *row_start, int32_t *row_len, int32_t *row_cap,
                               dcsr_stream_t stream);

/* Poll limit of the bounded look-back waits of the single-pass scans
 * (process-wide; default 2^20, negative restores it).  A tile that exceeds
 * it computes its prefix exactly from the earlier tiles' inputs instead, so
 * results never depend on it; 0 sends every tile down that fallback without
 * waiting (used to test it).  No reference counterpart. */
void dcsr_set_scan_wait_limit(int32_t polls);

/* dcsr_update without any allocation.  No reference counterpart: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, flags and
 * results as dcsr_update, with these limits: at most
 * dcsr_update_noalloc_max_triples() triples (65,536) per call, else
 * DCSR_ERR_TOO_LARGE and nothing happens; the matrix must be canonical
 * (updated, compacted or an SpGEMM output once), else DCSR_ERR_INVALID; the
 * pool is never grown or compacted, so a batch whose chunk claims do not fit
 * the free pool is rejected whole (DCSR_ERR_POOL_EXHAUSTED, or the sticky bit
 * with DCSR_UPDATE_NOSYNC).  The scratch is one fixed device array per GPU
 * (8 MiB, held until the process ends) shared by every call on that GPU:
 * calls that overlap on one GPU (two streams, two threads) race, so callers
 * serialize them. */
dcsr_status dcsr_update_noalloc(dcsr_handle h, const int32_t *rows, const int32_t *cols, const float *vals,
                                int64_t n, uint32_t flags, dcsr_stream_t stream);
int64_t dcsr_update_noalloc_max_triples(void);


/* Plain CSR SpMV comparator, y = A x.  Replaces device::spmv(const csr&, x, y)
 * -> spmv_csr_vector_kernel (include/spmv.h:310-332, include/spmv.inl:708-781). */
dcsr_status dcsr_csr_spmv(int64_t rows, const int32_t *row_offsets, const int32_t *cols,
                          const float *vals, const float *x, float *y, dcsr_stream_t stream);

/* CSR-rebuild comparators (the reference's CSR baselines of its DCSR-vs-CSR
 * experiment).  Device CSR inputs (any order within rows; duplicates are
 * summed in stored order); C comes out with columns ascending, unique;
 * C goes to caller buffers of c_cap entries (c_off: rows + 1); *c_nnz (host)
 * receives C's entry count, and when it exceeds c_cap nothing is written
 * and DCSR_ERR_CAPACITY is returned.  Each call allocates its temporaries
 * and syncs.
 *
 * C = A + B, B a batch of nb triples in any order (duplicates allowed), by
 * a whole rebuild: cusp's generic add (A and B concatenated as COO, sorted
 * by (row, column), reduced by key, zeros removed; cusp/system/detail/
 * generic/elementwise.inl:177-225), as src/Fill_Test.inl:240-258 times it.
 * Sums follow the host cusp::add order, so C equals dcsr_update's result
 * bit for bit.  An out-of-range triple: DCSR_ERR_BAD_INDEX, nothing written. */
dcsr_status dcsr_csr_add(int64_t rows, int64_t cols, const int32_t *a_off, const int32_t *a_col,
                         const float *a_val, const int32_t *b_row, const int32_t *b_col, const float *b_val,
                         int64_t nb, int32_t *c_off, int32_t *c_col, float *c_val, int64_t c_cap, int64_t *c_nnz,
                         dcsr_stream_t stream);
/* C = A * B (A: a_rows x a_cols, B: a_cols x b_cols) by global ESC: every
 * partial product expanded, sorted by (row, column), reduced by key
 * (cusp/system/detail/generic/multiply/spgemm.h:43-312, the comparator of
 * src/SPMM_Test.inl:142).  Partials are summed from 0 in A-row then B-row
 * order and exact zeros dropped (the host path's csr_spgemm order), so C
 * equals dcsr_spgemm's result on the same column-sorted rows bit for bit.
 * DCSR_ERR_UNSUPPORTED past 2^31 - 16 partial products. */
dcsr_status dcsr_csr_spgemm(int64_t a_rows, int64_t a_cols, int64_t b_cols, const int32_t *a_off,
                            const int32_t *a_col, const float *a_val, const int32_t *b_off, const int32_t *b_col,
                            const float *b_val, int32_t *c_off, int32_t *c_col, float *c_val, int64_t c_cap,
                            int64_t *c_nnz, dcsr_stream_t stream);

/* ---------------------------------------------------------------------------
 * MatrixMarket ingest (host; include/dcsr_mtx.hpp is the same reader for C++
 * callers).  Coordinate files of type real / integer / pattern (pattern
 * entries are 1), general / symmetric (off-diagonal entries mirrored) /
 * skew-symmetric (mirrored with the value negated) storage; the triples come
 * back 0-based, sorted by (row, column), with the file's values.  Replaces
 * cusp::io::read_matrix_market_file (cusp/io/detail/matrix_market.inl:
 * 117-269) and the drivers' ReadMatrixFile (include/load_matrix.h:54-116),
 * which forced every value to 1.0 and kept symmetric storage unexpanded.
 * ------------------------------------------------------------------------- */
typedef struct dcsr_mtx_s *dcsr_mtx;
/* Parse the file; *rows, *cols, *nnz (host, may be NULL) receive the shape
 * and the expanded entry count.  DCSR_ERR_INVALID on a format error (the
 * message: dcsr_mtx_error()). */
dcsr_status dcsr_mtx_read(const char *path, dcsr_mtx *out, int64_t *rows, int64_t *cols, int64_t *nnz);
/* Copy the triples into host arrays of nnz entries. */
dcsr_status dcsr_mtx_copy(dcsr_mtx m, int32_t *row, int32_t *col, float *val);
dcsr_status dcsr_mtx_free(dcsr_mtx m);
/* Message of this thread's last failed dcsr_mtx_read ("" after a success). */
const char *dcsr_mtx_error(void);

#ifdef __cplusplus
}
#endif
#endif /* DCSR_H */

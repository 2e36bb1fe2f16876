/*
 * qmgpu.h — C-ABI of the MI355X (gfx950) columnar operator library.
 *
 * This is the drop-in boundary for QueryMode's local operator hot path:
 * predicate filter -> hash group-by aggregate -> hash-join probe -> top-K.
 * Table columns are uploaded once and stay resident in HBM; a query plan is a
 * compact binary encoding of the QueryDescriptor subset the path needs; a
 * result is a device-resident columnar batch that is read out in the QMCB
 * wire layout.
 *
 * Conventions (the same ones the reference's native C-ABI uses,
 * wasm/src/nodejs.zig:101-742, and its WASM FFI, src/wasm-engine.ts:21-188):
 *   - opaque handles, explicit *_close / *_drop, a global qm_shutdown;
 *   - every call returns an int status (QM_OK == 0, negative on failure);
 *     no C++ exception crosses the ABI;
 *   - output buffers are caller-owned with an explicit capacity;
 *   - a failing call leaves a message readable through qm_last_error().
 * Handles may be used from several host threads. A context has four
 * execution lanes, each a HIP stream of its own: every call that runs device
 * work holds one lane for its duration, so up to four calls (from four host
 * threads) run on the device at once; a fifth waits for a free lane. Table
 * upload / drop and knob changes wait until no call runs.
 */
#ifndef QMGPU_H
#define QMGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: mirror QueryModeError codes (src/errors.ts:7-17) ---- */
#define QM_OK                    0
#define QM_ERR_TABLE_NOT_FOUND  -1  /* TABLE_NOT_FOUND  */
#define QM_ERR_COLUMN_NOT_FOUND -2  /* COLUMN_NOT_FOUND */
#define QM_ERR_INVALID_FORMAT   -3  /* INVALID_FORMAT (bad plan / QMCB bytes) */
#define QM_ERR_SCHEMA_MISMATCH  -4  /* SCHEMA_MISMATCH  */
#define QM_ERR_INVALID_FILTER   -5  /* INVALID_FILTER   */
#define QM_ERR_INVALID_AGGREGATE -6 /* INVALID_AGGREGATE */
#define QM_ERR_MEMORY_EXCEEDED  -7  /* MEMORY_EXCEEDED (HBM budget) */
#define QM_ERR_QUERY_FAILED     -8  /* QUERY_FAILED (device / runtime error) */
#define QM_ERR_UNSUPPORTED      -9  /* valid QueryMode query outside the GPU path */
#define QM_ERR_INVALID_ARG     -10  /* null handle, bad size, ... */
#define QM_ERR_OVERFLOW        -11  /* exact integer result does not fit its type */

/* ---- column dtypes: the QMCB tags (src/columnar.ts:22-32) plus the
 * narrow integer widths the QMCB tag set lacks (int8 columns are filtered by
 * the reference's JS path, src/operators.ts:531-543; they are widened to I32
 * when a result is written as QMCB). ---- */
#define QM_F64     0
#define QM_I64     1
#define QM_UTF8    2   /* not accepted on the device path */
#define QM_BOOL    3   /* stored one byte per row on the device */
#define QM_F32VEC  4   /* not accepted on the device path */
#define QM_NULLT   5
#define QM_I32     6
#define QM_F32     7
#define QM_I8     16
#define QM_I16    17

typedef struct qm_ctx    qm_ctx;
typedef struct qm_table  qm_table;
typedef struct qm_plan   qm_plan;
typedef struct qm_result qm_result;
typedef struct qm_batch  qm_batch;
typedef struct qm_prepared qm_prepared;

/* One host column handed to qm_table_create. `null_bitmap` follows QMCB:
 * ceil(nrows/8) bytes, bit (row & 7) of byte (row >> 3) set = row is null;
 * NULL means the column has no nulls. */
typedef struct qm_column_desc {
  const char    *name;
  int32_t        dtype;
  const void    *data;        /* nrows values, little-endian, host memory */
  const uint8_t *null_bitmap; /* optional */
} qm_column_desc;

/* Per-result execution statistics (the GPU analogue of QueryResult's
 * scan/pipeline timing fields, src/types.ts:262-295). */
typedef struct qm_stats {
  uint64_t rows_scanned;      /* logical table rows consumed (probe + build for joins) */
  uint64_t rows_selected;     /* rows that passed the filter */
  uint64_t groups;            /* output rows */
  uint64_t bytes_algorithmic; /* bytes of every referenced column, read once */
  double   device_ms;         /* device time of the query's kernels (HIP events) */
  double   scan_ms;           /* device time of the dominant scan kernel alone (batches: the longest launch reading the table) */
  double   wall_ms;           /* host wall time of qm_exec */
  uint32_t kernel_launches;
  uint32_t retries;           /* hash-table regrowths */
  uint32_t strategy;          /* group-by strategy picked by the tuner (QM_STRAT_*) */
  uint32_t pages_skipped;     /* zone-map pages (65,536 rows) not read (QueryResult.pagesSkipped); in a batch,
                                 the pages no fused query can match, which its shared-interval and
                                 partitioned-runs passes skip */
  double   pass_ms[4];        /* device time of the strategy's passes (QM_STRAT_PART_RUNS: row scatter,
                                 record scatter, group count, write), HIP events on the library stream */
  uint64_t pass_records;      /* QM_STRAT_PART_RUNS: rows the first pass kept as records (selected by any
                                 query of the launch) */
  uint32_t join_parts;        /* key-range passes of a join (1 = the whole build side in one index;
                                 more when the index would not fit the HBM budget, or forced by the
                                 join_parts knob) */
} qm_stats;

#define QM_STRAT_NONE        0
#define QM_STRAT_DIRECT_LDS  1  /* dense small key domain, per-workgroup LDS table */
#define QM_STRAT_HASH_LDS    2  /* global open addressing + per-workgroup LDS cache */
#define QM_STRAT_DIRECT_GLB  3  /* dense large key domain, direct global table */
#define QM_STRAT_HASH_GLB    4  /* global open addressing, no LDS cache */
#define QM_STRAT_PART_HASH   5  /* partitioned: scatter by key hash, LDS table per partition */
#define QM_STRAT_PART_DIRECT 6  /* partitioned: scatter by dense key slice, LDS slice per partition */
#define QM_STRAT_SHARED_INTERVAL 7 /* batched: one range filter + small key, aggregated once per range interval */
#define QM_STRAT_SORTED_RUNS 8  /* key column stored in order: segmented reduction over its runs */
#define QM_STRAT_PART_RUNS   9  /* batched: key in any row order, scattered into key partitions, runs sorted in LDS */

/* ---- lifecycle ----
 * Replaces the executor construction QueryMode.local() does
 * (src/local.ts:31-34, src/local-executor.ts:36-40 memoryBudgetBytes):
 * `hbm_budget_bytes` bounds resident tables + per-query device state. */
int qm_init(int device, uint64_t hbm_budget_bytes, qm_ctx **out);
int qm_shutdown(qm_ctx *ctx);
/* Copies the calling thread's last error message (NUL-terminated, truncated
 * to cap). Returns the full message length. ctx may be NULL. */
int qm_last_error(qm_ctx *ctx, char *buf, size_t cap);
/* Device name / CU count / budget bookkeeping. */
int qm_device_info(qm_ctx *ctx, char *name, size_t cap, uint32_t *num_cus,
                   uint64_t *budget, uint64_t *used);

/* ---- tables ----
 * Replaces FragmentSource + page decode (src/operators.ts:96-103,
 * 282-404): decoded columns are uploaded once into resident HBM instead of
 * being decoded into JS arrays page by page. */
int qm_table_create(qm_ctx *ctx, const char *name, uint64_t nrows,
                    uint32_t ncols, const qm_column_desc *cols, qm_table **out);
/* Fails with QM_ERR_INVALID_ARG while prepared batches over the table are open. */
int qm_table_drop(qm_table *t);
uint64_t qm_table_rows(const qm_table *t);

/* ---- plans ----
 * `plan` is the binary QueryDescriptor subset (src/client.ts:1254-1277):
 * filters / filterGroups (FilterOp, src/types.ts:130-134), groupBy,
 * aggregates (AggregateOp, src/types.ts:137-144), product computed columns,
 * join {leftKey,rightKey,type}, sortColumn/sortDirection, limit/offset,
 * projections. Its layout is given in the package's plan encoder. */
int qm_plan_compile(qm_ctx *ctx, const uint8_t *plan, size_t len, qm_plan **out);
int qm_plan_close(qm_plan *p);

/* ---- execution ----
 * Replaces LocalExecutor.execute / executeJoin (src/local-executor.ts:386-508,
 * 597-667) driving ScanOperator -> AggregateOperator / HashJoinOperator ->
 * TopKOperator / InMemorySortOperator -> ProjectOperator
 * (src/operators.ts:3073-3150). `right` is the join build side (the
 * reference builds the RIGHT side, src/operators.ts:2349-2436) or NULL. */
int qm_exec(qm_plan *p, qm_table *left, qm_table *right, qm_result **out);

/* ---- batched execution: many queries over one resident table in fused
 * persistent launches (no reference counterpart; the reference runs one
 * pipeline per query). Results come back in submission order. */
int qm_batch_exec(qm_ctx *ctx, qm_plan *const *plans, uint32_t nplans,
                  qm_table *table, qm_result **results_out);

/* Prepared batches (no reference counterpart; the nearest is the
 * reference's per-query assemblePipeline, src/operators.ts:3073-3150, rebuilt
 * for every execution): the qm_batch_exec of `plans` over `table` is run once
 * to size every buffer (the host plans with the result sizes it reads back),
 * then its whole launch sequence is captured into a HIP graph. Each
 * qm_prepared_run replays that graph with one host launch: every pass over
 * the table runs again, with no host planning, allocation or readback. Table
 * columns are immutable once uploaded, so a replay's results are the batch's
 * results. The nplans result handles qm_prepared_run returns are views of
 * buffers the prepared batch owns: valid until its next run or its close
 * (close them with qm_result_close as usual). Runs of one prepared batch are
 * serialised; the plans may be closed after qm_batch_prepare returns; the
 * table stays open until the prepared batch is closed. Joins are rejected
 * (QM_ERR_UNSUPPORTED) as in qm_batch_exec. */
int qm_batch_prepare(qm_ctx *ctx, qm_plan *const *plans, uint32_t nplans,
                     qm_table *table, qm_prepared **out);
int qm_prepared_run(qm_prepared *p, qm_result **results_out);
/* Plans, graph nodes and device bytes held by the prepared batch. */
int qm_prepared_info(qm_prepared *p, uint32_t *nplans, uint64_t *graph_nodes, uint64_t *device_bytes);
int qm_prepared_close(qm_prepared *p);

/* Asynchronous form of qm_batch_exec, for callers that keep working (or
 * queue the next batch) while one runs: the batch executes on a worker
 * thread of the library, on a lane of its own. The
 * plans and the table must stay open until the batch is done. Replaces the
 * reference's Promise-returning execute (src/client.ts:1280-1304) for a set
 * of queries. */
int qm_batch_submit(qm_ctx *ctx, qm_plan *const *plans, uint32_t nplans,
                    qm_table *table, qm_batch **out);
/* *done = 1 once the batch has finished (successfully or not), else 0. */
int qm_batch_poll(qm_batch *b, int *done);
/* Blocks until the batch has finished and returns its status (the message is
 * then qm_last_error's). On QM_OK the nplans result handles go to
 * results_out (the caller closes them); they can be taken once. */
int qm_batch_wait(qm_batch *b, qm_result **results_out);
/* Waits for the batch if it still runs; frees results not taken. */
int qm_batch_close(qm_batch *b);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): qm_exec of one filter -> GROUP BY -> aggregate plan (no
 * join, computed column, sort, limit / offset or IN list) that calls no
 * device allocator. The group table, descriptor, counters and compaction
 * temporaries live in a per-device static scratch array of 32 MiB; a plan
 * needing more fails with QM_ERR_MEMORY_EXCEEDED. The result columns go to
 * the caller's device buffers: out_data[k] holds `cap` values of output column
 * k (the dtype qm_exec's result gives that column), out_nulls[k] (optional,
 * 8-byte aligned, ceil(cap / 64) 64-bit words) receives its NULL bitmap
 * (QMCB layout, bit set = null); *rows_out = groups written (> cap:
 * QM_ERR_INVALID_ARG). Calls that overlap on one device race: they share the
 * scratch array. */
int qm_exec_noalloc(qm_plan *p, qm_table *t, void *const *out_data, uint32_t *const *out_nulls,
                    uint64_t cap, uint64_t *rows_out);

/* Device bytes one execution of `p` needs beyond the resident tables (no
 * reference counterpart: the reference checks its memoryBudget per operator,
 * src/operators.ts:2383-2462; the serving loop admits requests with this). */
int qm_plan_device_bytes(qm_plan *p, qm_table *left, qm_table *right, uint64_t *bytes);

/* Measurement probe (no reference counterpart): a plain streaming read of
 * `bytes` of device memory (16-byte loads, every byte once, `reps` times);
 * *gbps = bytes / (average launch time). The achievable read ceiling beside
 * the 8 TB/s nominal peak. */
int qm_measure_read_bw(qm_ctx *ctx, uint64_t bytes, uint32_t reps, double *gbps);

/* Tuning knobs of the device planner, set at run time by an empirical tuner
 * (querymode_amd/autotune.py) or read at qm_init from QM_TUNE_<NAME>. No
 * reference counterpart (the reference has no device planner). Knobs:
 * scatter_rpb, part_load, hot, qp_min, runs_min_q, runs_min_len,
 * runs_chunks_per_cu, runs_win, runs_qblock, runs_lds_pad, shared_lean, shared_waves, sorted1,
 * join_unique, join_parts, sort_parts, join_claim, join_payload, zone_maps, topk_bpc,
 * pr_max_s, pr_b2, pr_scatter_bpc;
 * values are clamped to each knob's range; an unknown name is QM_ERR_INVALID_ARG.
 * No knob changes a result: only how the planner lays the work out. */
int qm_tune_set(qm_ctx *ctx, const char *name, double value);
int qm_tune_get(qm_ctx *ctx, const char *name, double *value);

/* ---- results (QueryResult.columnarData, src/types.ts:262-295) ---- */
int qm_result_rows(qm_result *r, uint64_t *rows);
int qm_result_qmcb_size(qm_result *r, uint64_t *size);
/* Writes the QMCB encoding (src/columnar.ts:523-613) into out[0..cap). */
int qm_result_qmcb(qm_result *r, uint8_t *out, uint64_t cap, uint64_t *written);
/* QMCB of rows [row0, row0 + nrows) of a result (clamped to the result),
 * encoded on the device (NULL bitmaps shifted and BOOL columns packed by a
 * kernel, data copied by DMA) into pinned host memory the library owns:
 * *out / *size describe it until qm_qmcb_release(*out). The bytes are what
 * qm_result_qmcb writes for the same rows; a host runtime wraps them without
 * copying (Node: an external ArrayBuffer; Python: a memoryview), and a
 * cursor takes one range per batch. Replaces the host-side encode of
 * src/columnar.ts:523-613 and backs QueryExecutor.cursor (src/client.ts:1300). */
int qm_result_qmcb_pinned(qm_result *r, uint64_t row0, uint64_t nrows, const uint8_t **out, uint64_t *size);
/* Hands a buffer of qm_result_qmcb_pinned back (reused by later calls);
 * valid after the result and its context are closed. */
int qm_qmcb_release(const uint8_t *p);
int qm_result_stats(qm_result *r, qm_stats *out);
int qm_result_close(qm_result *r);

#ifdef __cplusplus
}
#endif
#endif /* QMGPU_H */

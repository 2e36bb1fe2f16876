// qm_device.h — the device-side query program: the form a bound query takes
// once the host planner has resolved column names, applied the reference's
// literal coercions and picked a group-by strategy. Passed to kernels by
// value (kernarg segment -> scalar loads), so every branch on it is uniform.
#pragma once
#include <stdint.h>
#include "qm_common.h"

namespace qm {

// device column types (QMCB tags + narrow ints; BOOL is one byte per row)
enum : int32_t {
  DT_F64 = 0, DT_I64 = 1, DT_BOOL = 3, DT_I32 = 6, DT_F32 = 7, DT_I8 = 16, DT_I16 = 17,
};

static constexpr int MAX_COLS = 12;
static constexpr int MAX_PREDS = 16;
static constexpr int MAX_STATES = 16;
static constexpr int MAX_KEYS = 2;
static constexpr int MAX_QUERIES = 16;     // per fused launch of the scan kernel
static constexpr int MAX_PSLOTS = 4;       // predicate columns held in registers

// Rows handled per thread per tile; tiles are blockDim.x * ROWS_PER_THREAD.
static constexpr int RPT = 8;
// Columns are allocated padded to this many rows so full-tile vector loads
// never leave the allocation.
static constexpr uint64_t ROW_PAD = 8192;
// rows per zone-map page (per-page min / max order codes kept at upload)
static constexpr uint64_t ZONE_ROWS = 65536;

struct DCol {
  const void *data;
  const uint8_t *nulls;  // QMCB bitmap: bit set = null; nullptr = no nulls
  int32_t dtype;
  int32_t side;          // 0 = scanned table row, 1 = join build row
};

// Predicate kinds after host coercion (src/decode.ts:485-578 and the WASM
// numeric filters, wasm/src/wasm/aggregates.zig:605-911).
enum : uint8_t {
  PK_INT_RANGE = 1,   // lo <= v <= hi on the integer value, ^ negate
  PK_F64_RANGE = 2,   // float interval with strictness flags, ^ negate; NaN -> nan_result
  PK_INT_IN = 3,      // v in sorted int64 list, ^ negate
  PK_F64_IN = 4,      // v in sorted f64 list (+NaN flag), ^ negate
  PK_IS_NULL = 5,
  PK_IS_NOT_NULL = 6,
  PK_FALSE = 7,       // folded constant
};

struct DPred {
  int8_t col;
  uint8_t kind;
  uint8_t group;       // 0 = top-level AND; g >= 1 = OR group g
  uint8_t negate;
  uint8_t nan_result;  // PK_F64_RANGE: result for NaN values (before negate is NOT applied)
  uint8_t lo_strict, hi_strict;
  uint8_t list_has_nan;
  int64_t ilo, ihi;
  double flo, fhi;
  const void *list;
  int32_t list_n;
  int8_t slot;         // register slot of the column in the AND-only fast path
  int8_t _pad[3];
};

// aggregate state kinds; each owns `words` u64 words of a table entry
enum : uint8_t {
  SK_COUNT = 1,       // count of non-null values (column) or rows (col < 0)
  SK_SUM_I64 = 2,     // exact int64 sum of an i8/i16/i32/bool column
  SK_SUM_128 = 3,     // 128-bit sum: int64 column, or fixed-point f64 (+flags word)
  SK_MAX_CODE = 4,    // unsigned max of an order code (MAX, or MIN via ~code)
  SK_SUM_WIDE = 5,    // f64 sum over a wide exponent range: flags word + nw signed 32-bit digits in 64-bit words
};

struct DState {
  uint8_t kind;
  uint8_t is_min;      // SK_MAX_CODE: store ~code (MIN)
  uint8_t fixed;       // SK_SUM_128: 1 = f64 fixed point (3 words: lo, hi, flags)
  uint8_t product;     // value = col_a * col_b in f64
  int8_t col_a, col_b;
  int16_t word;        // first word of this state in a table entry
  int32_t scale;       // fixed-point scale; SK_SUM_WIDE: exponent of digit 0's lowest bit
  int32_t nw;          // SK_SUM_WIDE: digit words
};

// Group-by addressing.
enum : uint8_t { GM_NONE = 0, GM_DIRECT = 1, GM_HASH = 2 };
enum : uint8_t { LM_NONE = 0, LM_DIRECT = 1, LM_HASH = 2 };

struct DKey {
  int8_t ncols;
  int8_t col[MAX_KEYS];
  uint8_t is_float[MAX_KEYS];
  uint8_t nullable[MAX_KEYS];
  // DIRECT: index = sum_i code_i * stride_i, code_i = v - min_i (or d_i - 1 for null)
  // HASH with 2 columns: packed = code_0 | code_1 << bits_0 (codes as above)
  int64_t kmin[MAX_KEYS];
  uint64_t dom[MAX_KEYS];     // domain size per key column incl. null code
  uint64_t stride[MAX_KEYS];
  int32_t bits[MAX_KEYS];
};

// One query inside a scan launch.
struct DQuery {
  int32_t npreds;
  int32_t ngroups;        // number of OR groups (0 = none)
  int32_t nstates;
  int32_t entry_words;    // u64 words per table entry (HASH: word 0 = key)
  int32_t count_word;     // word of the row-count state (always present)
  uint8_t gmode, lmode;
  uint8_t and_only;       // no OR groups and <= MAX_PSLOTS distinct predicate columns
  int8_t nslots;          // distinct predicate columns (fast path)
  int8_t slot_col[4];     // column of each slot
  DPred preds[MAX_PREDS];
  DState states[MAX_STATES];
  DKey key;
  uint64_t *table;        // global table: (gcap + 2) entries; slot gcap = unrepresentable key, gcap+1 = NULL key
  uint64_t gcap;          // DIRECT: domain size; HASH: power of two
  uint32_t *status;       // [0] overflow flag, [1] unused
  uint64_t *selected;     // selected-row counter
  int32_t lds_entries;    // LM_DIRECT: domain; LM_HASH: capacity (power of 2)
  int32_t lds_replicas;   // LM_DIRECT
  int32_t lds_stride;     // words between replicas (LM_DIRECT)
  int32_t lds_words;      // total LDS words used by this query
  int32_t lds_offset;     // word offset inside the block's LDS
  int32_t _p2;
};

// Partitioned aggregation (high-cardinality GROUP BY). Pass 1 compacts the
// selected rows into records {key code, value words..., [null mask]};
// pass 2 scatters them by key partition; pass 3 aggregates each partition
// in an LDS table and writes its groups into the global table.
static constexpr int MAX_SRC = 8;
struct DPart {
  int32_t nsrc;                 // distinct value sources of the states
  int8_t src_a[MAX_SRC], src_b[MAX_SRC];
  uint8_t src_float[MAX_SRC];   // value word holds f64 bits (else int64)
  int8_t state_src[MAX_STATES]; // per state: source index (-1 = COUNT(*))
  int32_t rec_words;            // 1 + nsrc (+1 mask word)
  int32_t has_mask;             // mask word: bit j = source j null; bit 63 side key, bit 62 NULL key
  int32_t pbits;                // HASH: partitions = 1 << pbits (top hash bits)
  int32_t sbits;                // DIRECT: partition = idx >> sbits, LDS slot = idx & (2^sbits - 1)
  int32_t nparts;
  int32_t lds_entries;          // pass-3 LDS table entries
  uint64_t *recs;               // pass-1 output: block b's records at recs + b * blk_cap * rec_words
  uint64_t blk_cap;             // records per block region (rows the block scans)
  uint64_t *blk_count;          // [grid] records written by each block
  uint64_t *hist;               // [nparts][grid] counts
  const uint64_t *part_off;     // exclusive scan of hist
  uint64_t *precs;              // pass-2 output (partition order)
  const uint64_t *seg;          // pass-3 segments: {part, start, count, exclusive} x nseg
  uint64_t *inserted;           // groups inserted into the global table
  // heavy hitters: keys frequent enough to aggregate in LDS during pass 2
  // instead of being scattered; their records count in partition `nparts`
  const uint64_t *hot_keys;     // HOT_SLOTS open-addressing slots (HASH: key code, DIRECT: code + 1; 0 empty)
  int32_t hot_n;                // 0 = no heavy hitters
  int32_t _ph;
};
static constexpr int HOT_SLOTS = 128;
static constexpr int HOT_PROBES = 8;

struct DScan {
  uint64_t nrows;
  int32_t ncols;
  int32_t nqueries;
  DCol cols[MAX_COLS];
  // join build side (fused probe): when join_table != nullptr every scanned
  // row is probed with key column `join_col`; build rows are located through
  // the join index and exposed as side-1 columns.
  const uint64_t *join_table;  // hash entries {stored key, start<<32|count} or DIRECT {start<<32|count}
  const uint32_t *join_rows;   // build row ids grouped by key
  uint64_t join_cap;           // HASH: power of 2; DIRECT: domain size
  int64_t join_min;            // DIRECT: key offset
  int32_t join_col;            // probe key column
  int32_t join_mode;           // 0 none, 1 DIRECT, 2 HASH
  int32_t join_key_float;      // probe key column is f64 (canonical bits)
  int32_t _p;
  // DIRECT with unique build keys (fused aggregates only): build row + 1 per
  // key slot (0 = no build row), one gather per probe row
  const uint32_t *join_unique; // nullptr: use join_table / join_rows
  // DIRECT with unique build keys and at most JOIN_PAY_BYTES of build-side
  // columns: per key slot 16 bytes, the columns' values at pay_off[c] and a
  // flags word at bytes 12..15 (bit 0: a build row exists, bit pay_null[c]:
  // side-1 column c is NULL; pay_null 0: never NULL). One 16-byte gather per
  // probe row then replaces the row-id gather and the per-column gathers.
  const uint32_t *join_pay;    // nullptr: no payload
  int8_t pay_off[MAX_COLS];
  int8_t pay_null[MAX_COLS];
  // zone maps: bit p set = the ZONE_ROWS-row page p (counted from row 0 of the
  // scan) holds no row the query can select; nullptr: every page is read
  const uint32_t *page_skip;
};
static constexpr int JOIN_PAY_BYTES = 12;

// Build-side columns packed into the join payload (host -> kernel).
struct DJoinPay {
  int32_t n;
  int32_t _p;
  DCol cols[MAX_COLS];
  int8_t off[MAX_COLS];
  int8_t nbit[MAX_COLS];
};

// ---- batched execution: many aggregate queries over one table in one pass.
// Every loaded column value becomes its order-preserving unsigned code
// (code_i64 / code_f64_bits), so each AND-ed range predicate is one unsigned
// compare `code - lo <= span`, for integer and float columns alike.
static constexpr int B_MAX_COLS = 4;     // distinct columns a batch launch loads
static constexpr int B_MAX_SLOTS = 4;    // AND-ed ranges per query
static constexpr int B_MAX_STATES = 4;   // states of a sorted-key query
static constexpr int B_MAX_SHARED_STATES = 8;
static constexpr int B_MAX_BOUNDS = 1024;
static constexpr int B_MAX_SHARED = 2;   // shared-interval groups per launch
static constexpr int B_BLOCK = 512;

// A query grouped by a sorted integer key: its own predicate, then a segmented
// reduction over the key's runs into a DIRECT table (entry layout of the
// single-query engine; slot = key - kmin).
struct BQuery {
  uint8_t nslots, key_col, nstates, _p0;
  uint8_t slot_col[B_MAX_SLOTS];
  uint8_t st_col[B_MAX_STATES];   // batch column of each state (0xFF: COUNT(*))
  uint32_t carry_off;             // word offset of this query's per-wave carry in LDS
  uint64_t lo[B_MAX_SLOTS], span[B_MAX_SLOTS];
  int64_t kmin;
  int32_t entry_words, _p1;
  DState st[B_MAX_STATES];
  uint64_t *table;                // zeroed, dom x entry_words
  uint64_t *selected;             // selected-row counter
};

// Queries whose filter is one range on column `col` and whose key is a small
// dense domain (or absent): one shared aggregation keyed by (elementary
// interval of `col`, key). Interval e of a code x = #{bounds <= x}; a query
// with code range [lo, hi] covers intervals [e(lo), e(hi + 1)).
struct BShared {
  uint8_t col, key_col, nstates, _p0;  // key_col 0xFF: no key (dom 1)
  int32_t nbounds;
  int64_t kmin;
  uint32_t dom;
  int32_t entry_words;                 // states only; state 0 is COUNT(*)
  uint8_t st_col[B_MAX_SHARED_STATES];
  DState st[B_MAX_SHARED_STATES];
  const uint64_t *bounds;              // nbounds ascending codes
  uint64_t *table;                     // (nbounds + 1) x dom x entry_words, zeroed
  uint32_t lds_off;                    // word offset of the workgroup table in LDS
  uint32_t bnd_off;                    // word offset of the bounds copy in LDS
};

static_assert(sizeof(BQuery) % 8 == 0, "BQuery is copied to LDS in 8-byte words");

struct BScan {
  uint64_t nrows;
  int32_t ncols, nq, nshared, carry_words;   // carry_words: per-wave LDS carry area
  const void *data[B_MAX_COLS];
  const uint8_t *nulls[B_MAX_COLS];
  int32_t dtype[B_MAX_COLS];
  uint64_t tiles_per_wave;             // contiguous 512-row tiles per wave range
  uint32_t carry_off;                  // word offset of the per-wave carries in LDS
  uint32_t qdesc_off;                  // word offset of the LDS copy of the query descriptors
  BShared sh[B_MAX_SHARED];
  // query-parallel sorted-key launches: the aggregate list every query of
  // the launch shares (states, their batch columns, entry layout)
  int32_t sig_nstates, sig_entry_words;
  DState sig_st[B_MAX_STATES];
  uint8_t sig_col[B_MAX_STATES];
  int32_t qp_pred_col;                 // uniform predicate column (single-range queries)
};
static constexpr int B_QP_SLOTS = 2;   // AND-ed ranges per query on the query-parallel path

// ---- run-sorted batch: queries grouped by a sorted integer key, each
// filtered by (at most) one range on a common column. A run is the rows of
// one key value. The sorted distinct query bounds cut the filter column into
// buckets (bucket of code x = #{bounds <= x}; bucket nbounds + 1 holds NULLs),
// so each query's range is a contiguous bucket range [e_lo, e_hi). A wave
// counting-sorts each run segment (<= 64 rows) by bucket; a query's rows in
// the segment are then one contiguous span of the sorted order, answered from
// prefix sums (additive states) and a sparse max table (order codes).
static constexpr int R_MAX_STATES = 4;    // state 0 (COUNT(*)) + 3 more
static constexpr int R_MAX_BUCKETS = 512; // nbounds + 2
static constexpr int R_MAX_QUERIES = 256;
static constexpr int R_MAX_OUTS = 8;
static constexpr int R_BLOCK = 256;       // 4 waves, one chunk each

// one output column of every query of the launch (same for all queries)
struct ROut {
  uint8_t kind;          // OK_KEY, OK_COUNT, OK_SUM_I64, OK_AVG_I64, OK_MIN, OK_MAX
  uint8_t src_is_float;
  int8_t state;          // state holding the value (OK_KEY: unused)
  int8_t cnt_state;      // state deciding null-ness / the divisor (-1: none)
  int32_t out_dtype;
};

struct RScan {
  uint64_t nrows;
  uint64_t chunk_rows;           // a wave owns the runs that start in its chunk
  uint32_t nchunks;
  int32_t nq;
  const void *key;
  int32_t key_dt;
  int32_t p_dt;                  // filter column (codes), its NULL bitmap
  const void *pcol;
  const uint8_t *pnull;
  int32_t nstates;               // 1 + nsum + nmax
  int32_t nsum;                  // states 1..nsum: additive; the rest: order codes
  DState st[R_MAX_STATES];       // st[0] = COUNT(*)
  const void *scol[R_MAX_STATES];
  const uint8_t *snull[R_MAX_STATES];
  int32_t s_dt[R_MAX_STATES];
  int32_t nbounds;               // sorted distinct bound codes
  int32_t bsteps;                // search steps: the bounds are padded to 2^bsteps - 1 entries
  const uint64_t *bounds;        // 2^bsteps - 1 entries (padding: all-ones)
  const uint32_t *qe;            // per query: e_lo | e_hi << 16
  uint32_t *counts;              // [nchunks][nq] groups per chunk (count pass)
  const uint64_t *offsets;       // [nchunks][nq] first output row (write pass)
  const uint64_t *totals;        // [nq] result rows (write pass bound)
  int32_t nouts, _p0;
  ROut outs[R_MAX_OUTS];
  void *out_base[R_MAX_OUTS];    // column o of every query: query q's row i at global row row_off[q] + i
  const uint64_t *row_off;       // [nq]
  uint32_t *out_nbase[R_MAX_OUTS]; // column o's NULL bitmap over the arena rows (nullptr: never null)
  uint64_t *selected;            // [nq] selected-row counters
  // staging: the write pass appends each finished run's groups (one record
  // per query that selected a row of it) to its chunk's record range; the
  // place pass moves them into the result columns, query by query
  // records are grouped by (chunk, query block of qbw queries; qbw 0: all
  // queries in one block), chunk-major
  int32_t nqb, qbw;              // query blocks, queries per block (0, 16, 32 or 64)
  uint64_t *chunk_groups;        // [nchunks][nqb] records per (chunk, query block) (count pass)
  const uint64_t *stg_base;      // [nchunks * nqb + 1] first record of each (chunk, query block)
  // key and state values are stored as (value - base) in key_w / val_w
  // bytes (4 where every value of the launch fits 32 bits after the base)
  uint16_t *stg_q;               // query
  void *stg_key;                 // run key - kbase
  uint32_t *stg_pos;             // row within the query's result
  uint64_t stg_total;            // records staged (the staging arrays' length)
  void *stg_val[R_MAX_STATES];   // state values - vbase (0: the row count)
  int64_t kbase;
  uint64_t vbase[R_MAX_STATES];
  int32_t key_w;                 // 4 or 8
  int32_t val_w[R_MAX_STATES];   // 4 or 8
};

// ---- lean shared-interval pass: one group of queries with a range on one
// common column and a small key domain; rows across lanes, raw column words
// loaded one row step ahead, (interval, key) table in LDS merged into
// `table` with atomics. Interval e of code x = #{bounds <= x}.
static constexpr int S_MAX_V = 4;     // distinct value columns (0 = the range column)
static constexpr int S_BLOCK = 256;
struct SScan {
  uint64_t nrows;
  uint64_t rows_per_block;
  const void *key;               // nullptr: no key (dom 1)
  int32_t key_dt;
  uint32_t dom;
  int64_t kmin;
  int32_t nv;
  int32_t nstates;
  const void *vcol[S_MAX_V];
  const uint8_t *vnull[S_MAX_V];
  int32_t v_dt[S_MAX_V];
  DState st[B_MAX_SHARED_STATES];
  int8_t st_v[B_MAX_SHARED_STATES];  // value column of each state (-1: COUNT(*))
  int32_t entry_words;
  int32_t nbounds, bsteps;           // bounds padded to 2^bsteps - 1 entries (all-ones)
  int32_t _p0;
  const uint64_t *bounds;
  uint64_t *table;                   // (nbounds + 1) x dom x entry_words, zeroed
  const uint32_t *page_skip;         // bit p: no query can match a row of ZONE_ROWS-row page p (nullptr: none)
};

// ---- one query grouped by a sorted key (lean): a wave owns the runs that
// start in its chunk; per 64-row window the rows of the current run that pass
// the AND-ed ranges are reduced across lanes (DPP), and a finished run with
// selected rows is written once into the query's DIRECT table (slot = key -
// kmin) with plain stores.
struct S1Scan {
  uint64_t nrows;
  uint64_t chunk_rows;
  uint32_t nchunks;
  int32_t nslots;                   // AND-ed ranges (code lo + span); -1: never true
  const void *key;
  int32_t key_dt;
  int32_t nstates;                  // states of the query's entry layout (state 0 COUNT(*))
  int64_t kmin;
  const void *pcol[B_MAX_SLOTS];
  const uint8_t *pnull[B_MAX_SLOTS];
  int32_t p_dt[B_MAX_SLOTS];
  uint64_t lo[B_MAX_SLOTS], span[B_MAX_SLOTS];
  DState st[R_MAX_STATES];
  const void *scol[R_MAX_STATES];
  const uint8_t *snull[R_MAX_STATES];
  int32_t s_dt[R_MAX_STATES];
  int32_t entry_words, _p0;
  uint64_t *table;                  // zeroed, dom x entry_words
  uint64_t *selected;
};

// ---- many queries grouped by one integer key stored in any row order
// (k_pruns.hip). Rows selected by at least one query become records
// (u32 words: word 0 = key bits below the level-1 digit | bucket << 16,
// then the raw words of the state columns, then a NULL-bit word when a state
// column has NULLs), scattered by the key's high bits into level-1
// partitions, then by the next bits into final partitions of 2^s keys; a
// workgroup then sorts one final partition by (key, bucket) in LDS and
// answers every query from prefix sums over that order (lanes = keys).
static constexpr int PR_MAX_V = 4;       // distinct state columns copied into records
static constexpr int PR_MAX_RW = 8;      // record words
static constexpr int PR_MAX_P = 1024;    // digits per scatter pass
static constexpr int PR_BLOCK = 512;     // scatter / partition workgroup
static constexpr int PR_STAGE_WORDS = 12288;  // LDS staging of one scatter tile (48 KB: two workgroups per CU)
static constexpr int PR_MAX_S = 6;       // key bits per final partition (one key per lane)
static constexpr int PR_MAX_STATES = 6;  // COUNT(*) + up to 5 more (COUNT(col) guards of nullable columns)
struct PRScan {
  uint64_t nrows;
  const void *key;
  int32_t key_dt;
  int32_t p_dt;                  // filter column (order codes) and its NULL bitmap
  const void *pcol;
  const uint8_t *pnull;
  int64_t kmin;
  int32_t nv;                    // state columns copied into records
  int32_t rw;                    // record words
  const void *vcol[PR_MAX_V];
  const uint8_t *vnull[PR_MAX_V];
  int32_t v_dt[PR_MAX_V];
  int32_t v_word[PR_MAX_V];      // first record word of column j (1 or 2 words)
  int32_t null_word;             // record word with the column NULL bits (-1: none)
  int32_t nbounds, bsteps;       // buckets: bucket(x) = #{bounds <= x}; NULL filter value: nbounds + 1
  const uint64_t *bounds;        // 2^bsteps - 1 entries (padding all-ones)
  uint32_t bucket_on[R_MAX_BUCKETS / 32];  // buckets some query covers (others are dropped)
  uint64_t *bucket_rows;         // [R_MAX_BUCKETS] rows per bucket (selected-row counts)
  int32_t s;                     // key bits per final partition
  int32_t sh1;                   // level-1 digit = key code >> sh1 (sh1 = s + level-2 bits)
  uint32_t np1, np;              // level-1 partitions, final partitions
  uint32_t tile;                 // records per scatter tile (PR_STAGE_WORDS / rw)
  uint32_t _p0;
  const uint64_t *start1;        // [np1] first record of each level-1 region (capacity from the key histogram)
  uint32_t *cur1;                // [np1] records written (zeroed before the batch)
  const uint64_t *start2;        // [np]
  uint32_t *cur2;                // [np]
  uint32_t *rec1, *rec2;         // record buffers
  const uint32_t *chunk0;        // [np1 + 1] first pass-2 workgroup of each level-1 partition
  uint32_t cap;                  // records a final partition may hold (LDS)
  int32_t nq;
  const uint32_t *qe;            // per query: e_lo | e_hi << 16 (bucket range [e_lo, e_hi))
  int32_t nstates, nsum;         // state 0 = COUNT(*); 1..nsum additive; the rest order codes
  DState st[PR_MAX_STATES];
  int8_t st_v[PR_MAX_STATES];    // record column of each state (-1: none)
  int32_t nouts;
  ROut outs[R_MAX_OUTS];
  uint64_t *gcount;              // [nq][np] groups per (query, partition); the write pass reads its
                                 // exclusive scan (in place): query q's rows of partition P start at
                                 // gcount[q][P] - gcount[q][0]
  void *out_base[R_MAX_OUTS];
  uint32_t *out_nbase[R_MAX_OUTS];
  const uint64_t *row_off;       // [nq] first arena row of each query
  uint32_t *status;              // [0] a partition exceeded cap
  const uint32_t *page_skip;     // bit p: no query can match a row of ZONE_ROWS-row page p (nullptr: none)
  // bucket lookup (bucket_slot): codes [bslot_lo + (k << bslot_shift), + (1 << bslot_shift)) hold
  // buckets bslot[k] .. bslot[k + 1] (bslot[k] = bounds below the slot's first code); nullptr: search
  const uint16_t *bslot;
  uint64_t bslot_lo;
  int32_t bslot_shift;
  uint32_t bslots;               // slots (<= PR_BSLOTS)
};
static constexpr uint32_t PR_BSLOTS = 1024;

}  // namespace qm

// qm_kernels.h — result descriptors and the host-side launcher entry points
// of the gfx950 kernels (all launches go on the caller's stream).
#pragma once
#include <hip/hip_runtime.h>
#include "qm_device.h"

namespace qm {

enum : uint8_t {
  OK_KEY = 1, OK_COUNT, OK_SUM_I64, OK_SUM_I128, OK_SUM_F64, OK_AVG_F64, OK_AVG_I64, OK_AVG_BIG,
  OK_MIN, OK_MAX, OK_SUM_WIDE, OK_AVG_WIDE,
};

struct DOut {
  uint8_t kind;
  uint8_t key_idx;
  uint8_t src_is_float;
  uint8_t nw;         // OK_SUM_WIDE / OK_AVG_WIDE: digit words (scale: exponent of digit 0)
  int32_t out_dtype;
  int16_t word;       // state word
  int16_t cnt_word;   // count word deciding null-ness / divisor (-1 = none)
  int32_t scale;      // fixed-point scale
  void *data;
  uint32_t *nulls;    // output null bitmap (QMCB layout, bit set = null), pre-zeroed, or nullptr
};

static constexpr int MAX_OUTS = 20;

struct DFinal {
  const uint64_t *table;
  uint64_t nslots;        // slots scanned (HASH: gcap + 2, DIRECT: domain)
  int32_t entry_words;
  int32_t count_word;     // row-count word (occupancy)
  uint8_t gmode;
  uint8_t always_one_row; // aggregate without GROUP BY: exactly one row
  uint8_t _p[2];
  int32_t nouts;
  DKey key;
  DOut outs[MAX_OUTS];
  uint32_t *status;       // [1] <- int128 SUM did not fit int64
};

// merge of a shared-interval table into one query's DIRECT table
struct BCombine {
  const uint64_t *src;     // (nbounds + 1) x dom x src_words
  uint64_t *dst;           // dom x dst_words (the query's entry layout)
  uint64_t *selected;      // rows selected by the query
  uint32_t dom, e_lo, e_hi;
  int32_t src_words, dst_words, nstates;
  int16_t src_word[B_MAX_SHARED_STATES], dst_word[B_MAX_SHARED_STATES];
  uint8_t kind[B_MAX_SHARED_STATES], fixed[B_MAX_SHARED_STATES];
};

// top-K candidate (composite order key, lexicographic on (a, b))
struct TopKey { uint64_t a, b; };

// HAVING program (k_post.hip): postfix ops over up to HP_MAX_COLS result
// columns; literals are JS numbers (HL_F64), bigints (HL_I64) or NULL.
enum : uint8_t { HL_NULL = 0, HL_F64 = 1, HL_I64 = 2 };
enum : uint8_t { HP_CMP = 1, HP_CMP_COL, HP_BETWEEN, HP_IN, HP_ISNULL, HP_AND, HP_OR, HP_NOT };
constexpr int HP_MAX_OPS = 32, HP_MAX_COLS = 8, HP_MAX_DEPTH = 16;
struct HPOp {
  uint8_t kind, cmp, col, col2, neg, lk, lk2, pad;
  uint32_t list_off, list_n;
  double f, f2;
  int64_t i, i2;
};
struct HProg {
  int32_t nops, ncols;
  DCol cols[HP_MAX_COLS];
  HPOp ops[HP_MAX_OPS];
  const uint8_t *list_kind;  // IN values of every HP_IN op
  const double *list_f;
  const int64_t *list_i;
};

}  // namespace qm

extern "C" {
hipError_t qmk_scan_agg(const qm::DScan *sc, const qm::DQuery *d_queries, int grid, int block,
                        size_t lds_bytes, hipStream_t st);
hipError_t qmk_fill_u64(uint64_t *p, uint64_t n, uint64_t v, hipStream_t st);
hipError_t qmk_zero_bytes(void *p, uint64_t n, hipStream_t st);

// SQL post-stages (k_post.hip)
hipError_t qmk_having(const qm::HProg *prog, uint64_t n, uint8_t *keep, hipStream_t st);
uint64_t qmk_mask_blocks(uint64_t n);
hipError_t qmk_mask_count(const uint8_t *keep, uint64_t n, uint64_t *counts, hipStream_t st);
hipError_t qmk_mask_write(const uint8_t *keep, uint64_t n, const uint64_t *offs, uint64_t *ids, hipStream_t st);
hipError_t qmk_iota(uint64_t *p, uint64_t n, hipStream_t st);
hipError_t qmk_order_codes(const qm::DCol *c, const uint64_t *perm, uint64_t n, int desc, int null_pass,
                           uint64_t *keys, hipStream_t st);
hipError_t qmk_distinct_marks(const qm::DCol *d_cols, int ncols, const uint64_t *perm, uint64_t n, uint8_t *keep,
                              hipStream_t st);
hipError_t qmk_gather_u64(const uint64_t *src, const uint64_t *idx, uint64_t n, uint64_t *dst, hipStream_t st);
hipError_t qmk_qmcb_bits(const uint8_t *src, int is_bool, uint64_t row0, uint64_t nrows, uint8_t *dst, uint32_t *any,
                         hipStream_t st);
size_t qmk_radix_pairs_tmp_bytes(uint64_t n);
// stable LSD radix sort of (key, 64-bit value) pairs on key bits [begin_bit, end_bit) (k_radix.hip);
// kin != kout, vin != vout; tmp of qmk_radix_tmp_bytes(n, key bytes)
size_t qmk_radix_tmp_bytes(uint64_t n, int key_bytes);
hipError_t qmk_radix_sort_u64(const uint64_t *kin, uint64_t *kout, const uint64_t *vin, uint64_t *vout, uint64_t n,
                              int begin_bit, int end_bit, void *tmp, size_t tmp_bytes, hipStream_t st);
hipError_t qmk_radix_sort_u8(const uint8_t *kin, uint8_t *kout, const uint64_t *vin, uint64_t *vout, uint64_t n,
                             int begin_bit, int end_bit, void *tmp, size_t tmp_bytes, hipStream_t st);
hipError_t qmk_radix_pairs(const uint64_t *kin, uint64_t *kout, const uint64_t *vin, uint64_t *vout, uint64_t n,
                           int end_bit, void *tmp, size_t tmp_bytes, hipStream_t st);
// Raises a kernel's dynamic-LDS limit to at least `bytes` (the attribute is
// set once per kernel and size: launches captured into a graph then make no
// attribute call).
hipError_t qmk_func_lds(const void *fn, size_t bytes);

hipError_t qmk_exscan_u64(const uint64_t *in, uint64_t n, uint64_t *out, uint64_t *total,
                          uint64_t *tmp_sums, hipStream_t st);
uint64_t qmk_exscan_tmp_words(uint64_t n);

uint64_t qmk_compact_blocks(uint64_t nslots);
hipError_t qmk_compact_count(const qm::DFinal *f, uint64_t *block_counts, hipStream_t st);
hipError_t qmk_compact_write(const qm::DFinal *f, const uint64_t *block_offsets, hipStream_t st);
// many tables of at most qmk_compact_small_slots() slots in one launch; the
// output columns hold nslots rows and their NULL bitmaps are cleared here
uint64_t qmk_compact_small_slots();
hipError_t qmk_compact_small(const qm::DFinal *d_f, uint32_t n, uint64_t *counts, hipStream_t st);

hipError_t qmk_select_count(const qm::DScan *sc, const qm::DQuery *d_q, uint64_t nblocks, uint64_t tpb,
                            uint64_t *block_counts, hipStream_t st);
hipError_t qmk_select_write(const qm::DScan *sc, const qm::DQuery *d_q, uint64_t nblocks, uint64_t tpb,
                            const uint64_t *block_offsets, uint64_t out_cap, uint64_t *out_rows,
                            hipStream_t st);
hipError_t qmk_gather(const qm::DCol *c, const uint64_t *rows, const uint32_t *idx, uint64_t n,
                      int32_t out_dt, void *out, uint32_t *nulls, hipStream_t st);

// top-K: order keys from one sort column (optionally filtered by a query)
// gthr: one word of device memory, set to ~0 by the caller before the launch
hipError_t qmk_topk_scan(const qm::DScan *sc, const qm::DQuery *d_q, int sort_col, int desc,
                         uint32_t k, int grid, qm::TopKey *cand, uint32_t *cand_counts, uint64_t *gthr,
                         hipStream_t st);
hipError_t qmk_topk_merge(const qm::TopKey *cand, const uint32_t *cand_counts, uint32_t nlists,
                          uint32_t lists_per_block, uint32_t k, qm::TopKey *out, uint32_t *out_count,
                          hipStream_t st);
hipError_t qmk_topk_rows(const qm::TopKey *keys, const uint32_t *count, uint32_t offset, uint64_t *rows,
                         uint32_t n, hipStream_t st);
uint32_t qmk_topk_cap(uint32_t k);

// full sort of n order keys (composite), returns permutation as row ids
hipError_t qmk_sort_keys(const qm::DCol *c, uint64_t n, int desc, uint64_t *rows_out, void *tmp,
                         size_t tmp_bytes, hipStream_t st);
size_t qmk_sort_tmp_bytes(uint64_t n);
hipError_t qmk_sort_hist(const qm::DCol *c, uint64_t n, int desc, uint32_t *hist, hipStream_t st);
hipError_t qmk_sort_range(const qm::DCol *c, uint64_t n, int desc, uint64_t lo, uint64_t hi, int nulls, uint8_t *keep,
                          hipStream_t st);
hipError_t qmk_sort_gather_keys(const qm::DCol *c, const uint64_t *ids, uint64_t m, int desc, uint64_t *keys,
                                hipStream_t st);

// join build
// sel: optional ascending build-row ids (rows passing the right side's filters);
// tmp: qmk_join_build_tmp_words(nent, nrows) u64 words; cursor: nent + 1 u32, zero
uint64_t qmk_join_build_tmp_words(uint64_t nent, uint64_t nrows);
hipError_t qmk_join_build(const qm::DCol *key, uint64_t nrows, const uint64_t *sel, int mode, int key_float,
                          int64_t kmin, uint64_t cap, uint64_t *table, uint32_t *rows, uint32_t *cursor,
                          uint64_t *tmp, uint32_t *status, hipStream_t st);
// unique-key DIRECT index straight from the build rows (uniq zeroed, cap + 16 u32);
// status[1] <- duplicate keys (the index is then incomplete)
hipError_t qmk_join_claim(const qm::DCol *key, uint64_t nrows, const uint64_t *sel, int64_t kmin, uint64_t cap,
                          uint32_t *uniq, uint32_t *status, hipStream_t st);
// unique-key DIRECT index (build row + 1 per slot); status[1] <- duplicate keys
hipError_t qmk_join_unique(const uint64_t *table, uint64_t nent, const uint32_t *rows, uint32_t *uniq,
                           uint32_t *status, hipStream_t st);
// join payload (DScan::join_pay) of `cap` DIRECT slots from the unique index: 4 u32 per slot
hipError_t qmk_join_payload(const uint32_t *uniq, uint64_t cap, const qm::DJoinPay *jp, uint32_t *out,
                            hipStream_t st);
// materialised inner-join pairs (probe row, build row) in probe order
hipError_t qmk_join_count(const qm::DScan *sc, const qm::DQuery *d_q, uint64_t nblocks, uint64_t tpb,
                          uint64_t *block_counts, hipStream_t st);
hipError_t qmk_join_write(const qm::DScan *sc, const qm::DQuery *d_q, uint64_t nblocks, uint64_t tpb,
                          const uint64_t *block_offsets, uint64_t out_cap, uint64_t *out_left,
                          uint64_t *out_right, hipStream_t st);

// partitioned aggregation
hipError_t qmk_part_scan(const qm::DScan *sc, const qm::DQuery *d_q, const qm::DPart *pp, uint32_t grid,
                         uint64_t tpb, int fast, int prefetch, hipStream_t st);
int qmk_part_scan_blocks_per_cu(int fast, size_t lds);
hipError_t qmk_part_scatter(const qm::DQuery *d_q, const qm::DPart *pp, uint32_t g1, uint32_t rpb, int entry_words,
                            hipStream_t st);
hipError_t qmk_part_starts(const uint64_t *part_off, uint32_t nparts, uint32_t grid, uint64_t total,
                           uint64_t *starts, hipStream_t st);
hipError_t qmk_part_agg(const qm::DQuery *d_q, const qm::DPart *pp, uint32_t nseg, uint32_t grid, int entry_words,
                        hipStream_t st);

// batched aggregate queries (one fused pass)
uint32_t qmk_batch_carry_words();
int qmk_batch_waves_per_cu(int ncols, int wide, int mode, size_t lds_bytes);
hipError_t qmk_batch_scan(const qm::BScan *bs, const qm::BQuery *d_q, uint32_t grid, size_t lds_bytes, int wide,
                          hipStream_t st);
hipError_t qmk_batch_combine(const qm::BCombine *d_cb, uint32_t n, uint32_t max_dom, hipStream_t st);
int qmk_batch_qpf_supported(int nadd, int nmax);
int qmk_batch_qpf_waves_per_cu(int nadd, int nmax, int qpl);
hipError_t qmk_batch_qpf_scan(const qm::BScan *bs, const qm::BQuery *d_q, uint32_t grid, int nadd, int nmax,
                              int qpl, hipStream_t st);
// run-sorted batch (sorted key, one range column): count pass, per-query
// offsets over chunks, write pass into the result columns
hipError_t qmk_runs_count(const qm::RScan *rs, hipStream_t st);
hipError_t qmk_runs_offsets(const uint32_t *counts, uint32_t nchunks, uint32_t nq, uint64_t *offsets,
                            uint64_t *totals, hipStream_t st);
int qmk_runs_supported(int nsum, int nmax);
hipError_t qmk_runs_write(const qm::RScan *rs, uint32_t lds_pad, hipStream_t st);
hipError_t qmk_runs_place(const qm::RScan *rs, int win, hipStream_t st);
// lean shared-interval pass (one group per launch)
size_t qmk_shared_lds_bytes(const qm::SScan *ss);
int qmk_shared_blocks_per_cu(const qm::SScan *ss);
hipError_t qmk_shared_scan(const qm::SScan *ss, uint32_t grid, hipStream_t st);
// one query grouped by a sorted key into its DIRECT table
hipError_t qmk_sorted1_scan(const qm::S1Scan *q, hipStream_t st);
// column statistics: min/max codes, finite max |x|, NaN count, HLL registers
hipError_t qmk_col_stats(const qm::DCol *c, uint64_t n, uint64_t *stats /*8 words*/, uint32_t *hll /*4096*/,
                         hipStream_t st);
// zone maps: per ZONE_ROWS-row page, min / max order code of the non-NULL values
hipError_t qmk_zone_maps(const qm::DCol *c, uint64_t n, uint64_t *zmin, uint64_t *zmax, hipStream_t st);
// key histogram of an integer column: hist[(v - kmin) >> g] += 1 (hist zeroed by the caller)
hipError_t qmk_key_hist(const qm::DCol *c, uint64_t n, int64_t kmin, int g, uint32_t nbins, uint32_t *hist,
                        hipStream_t st);
// batched group-by over a key in any row order (k_pruns.hip): scatter passes
// (src 0: table rows -> level-1 regions; src 1: level-1 -> final partitions),
// per-(query, partition) group counts, and the write pass
hipError_t qmk_stream_read(const void *p, uint64_t bytes, uint32_t grid, uint64_t *sink, hipStream_t st);
uint32_t qmk_pr_tile(int rw);
int qmk_pr_supported(int nsum, int nmax);
size_t qmk_pr_runs_lds(uint32_t nbin, int nsum, int nmax, uint32_t cap);
uint32_t qmk_pr_max_cap(uint32_t nbin, int nsum, int nmax);
hipError_t qmk_pr_scatter(const qm::PRScan *ps, int src, uint32_t grid, hipStream_t st);
hipError_t qmk_pr_count(const qm::PRScan *ps, hipStream_t st);
hipError_t qmk_pr_runs(const qm::PRScan *ps, hipStream_t st);
}

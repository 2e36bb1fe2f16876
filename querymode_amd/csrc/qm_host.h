// qm_host.h — host-side structures of the library: errors, the device
// memory pool, resident tables, parsed plans and results.
#pragma once
#include <hip/hip_runtime.h>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <condition_variable>
#include <shared_mutex>
#include <string>
#include <vector>
#include "../../include/qmgpu.h"
#include "qm_kernels.h"

namespace qm {

struct Error {
  int code;
  std::string msg;
};
[[noreturn]] void fail(int code, const char *fmt, ...);
void hip_check(hipError_t e, const char *what);
#define HIPX(expr) ::qm::hip_check((expr), #expr)

// ---- literal / plan model (the QueryDescriptor subset) ----
enum : int { LIT_NONE = 0, LIT_F64 = 1, LIT_I64 = 2, LIT_LIST = 3, LIT_PAIR = 4 };
enum : int {
  OP_EQ = 0, OP_NEQ, OP_LT, OP_LTE, OP_GT, OP_GTE, OP_BETWEEN, OP_NOT_BETWEEN, OP_IN, OP_NOT_IN,
  OP_IS_NULL, OP_IS_NOT_NULL, OP_LIKE, OP_NOT_LIKE,
};
enum : int { AGG_SUM = 0, AGG_AVG, AGG_MIN, AGG_MAX, AGG_COUNT, AGG_COUNT_DISTINCT, AGG_STDDEV, AGG_VARIANCE,
             AGG_MEDIAN, AGG_PERCENTILE };

struct Lit {
  int kind = LIT_NONE;  // LIT_F64 (JS number) or LIT_I64 (JS bigint)
  double f = 0;
  int64_t i = 0;
};

struct Filter {
  std::string col;
  int op = 0;
  int vkind = LIT_NONE;
  Lit v;                 // scalar
  std::vector<Lit> list; // in / not_in, between pair (2 entries)
};

struct Agg {
  int fn = 0;
  std::string col;
  std::string alias;
};

struct Computed {
  std::string alias;
  int kind = 1;  // 1 = multiply
  std::string a, b;
};

// HAVING: a postfix program over result columns, evaluated per row with the
// reference SQL evaluator's three-valued logic (src/sql/evaluator.ts:7-200).
enum : int { HK_CMP = 1, HK_CMP_COL, HK_BETWEEN, HK_IN, HK_ISNULL, HK_AND, HK_OR, HK_NOT };
struct HavingOp {
  int kind = 0;
  std::string col, col2;   // HK_CMP_COL compares col with col2
  int cmp = 0;             // 0 eq, 1 ne, 2 lt, 3 le, 4 gt, 5 ge
  bool negated = false;    // BETWEEN / IN / IS NULL
  Lit a, b;                // literal; BETWEEN low / high
  std::vector<Lit> list;   // IN values (kind LIT_NONE = NULL)
};

struct OrderKey {
  std::string col;
  bool desc = false;
};

struct Plan {
  std::vector<Filter> filters;
  std::vector<std::vector<Filter>> groups;
  std::vector<Computed> computed;
  std::vector<std::string> groupby;
  std::vector<Agg> aggs;
  bool has_join = false;
  std::string lkey, rkey;
  int jtype = 0;
  std::vector<Filter> rfilters;
  std::vector<std::vector<Filter>> rgroups;
  bool has_sort = false;
  std::string sort_col;
  bool desc = false;
  int64_t limit = -1, offset = 0;
  std::vector<std::string> proj;
  // SQL post-stages (src/sql/executor.ts:30-170, src/operators.ts:1177-1227)
  bool has_distinct = false;
  std::vector<std::string> distinct;   // empty with has_distinct: every column
  std::vector<OrderKey> order_by;      // multi-column ORDER BY (replaces sort_col)
  std::vector<HavingOp> having;
  bool post() const { return has_distinct || !order_by.empty() || !having.empty(); }
};

Plan parse_plan(const uint8_t *p, size_t n);

// predicate coercion: one reference FilterOp on a column of dtype `dt`
// -> one device predicate (kind PK_*), following the reference's WASM and
// JS filter paths. `list_storage` receives IN-list values to upload.
struct BoundPred {
  DPred p;
  std::vector<int64_t> ilist;
  std::vector<double> flist;
};
BoundPred coerce_filter(const Filter &f, int dt);

// ---- context / memory ----
struct qm_ctx_impl;
}  // namespace qm

// One execution lane of a context: a stream of its own and the events that
// time work on it. Each call that runs device work holds one lane for its
// whole duration, so up to QM_LANES calls overlap on the device. Buffers a
// call releases stay with its lane (reusable by the same call, whose stream
// orders the uses) until the lane's stream has drained; then they go back to
// the context pool that every lane draws from.
constexpr int QM_LANES = 4;

// How a lane runs the host side of a call. RECORD executes normally and keeps
// every device-to-host readback (the sizes and counts the host plans with) on
// a tape, plus how many buffers of each size were live at once; CAPTURE runs
// the same host logic while the lane's stream is being captured into a HIP
// graph: readbacks come from the tape instead of the device, waits and timing
// are skipped, zeroing is a fill kernel, host-to-device uploads are staged in
// pinned memory the graph keeps, and buffers come from a private pool
// reserved before the capture from the recorded sizes (no device allocator
// call while the stream captures); the graph owns that pool.
enum : int { LANE_NORMAL = 0, LANE_RECORD = 1, LANE_CAPTURE = 2 };
struct qm_capture {
  int mode = LANE_NORMAL;
  std::vector<std::vector<uint8_t>> tape;   // readbacks, in call order
  size_t pos = 0;                           // next tape entry (CAPTURE)
  uint64_t h2d_bytes = 0;                   // uploaded bytes (RECORD: sizes the staging area)
  uint8_t *staging = nullptr;               // pinned host memory (CAPTURE)
  size_t staging_cap = 0, staging_used = 0;
  std::map<void *, size_t> rec_live;        // RECORD: buffers the run allocated, still live
  std::map<size_t, int> live, peak;         // RECORD: live / most buffers of each size at once
  std::multimap<size_t, void *> free;       // CAPTURE: the private pool's free buffers
  uint64_t private_bytes = 0;               // the private pool's size
};

struct qm_lane {
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evk0 = nullptr, evk1 = nullptr, evk2 = nullptr;  // around the dominant scan kernel(s)
  bool busy = false;
  std::multimap<size_t, void *> held;  // released during the current call
  uint64_t held_bytes = 0;
  qm_capture *cap = nullptr;           // set while recording / capturing a prepared batch
};

struct qm_ctx {
  int device = 0;
  uint64_t budget = 0;
  std::atomic<uint64_t> used{0};
  qm_lane lanes[QM_LANES];
  std::mutex lanes_mu;
  std::condition_variable lanes_cv;
  // shared by executions, exclusive for table upload / drop and knob changes
  std::shared_mutex tables_mu;
  std::mutex cache_mu;     // lazily built per-column caches (partition geometries)
  std::mutex scratch_mu;   // the no-allocation scratch array
  std::mutex pool_mu;
  std::multimap<size_t, void *> pool;  // free device buffers by size
  uint64_t pooled = 0;     // bytes in `pool` and in the lanes' held lists
  int num_cus = 256;
  std::string name;
  void *scratch = nullptr;  // device address of this device's copy of the no-allocation scratch array
  // tuning parameters of the partitioned group-by (results do not depend on
  // them); defaults, overridable with QM_TUNE_* environment variables
  struct Tune {
    int scatter_rpb = 0;       // pass-1 regions per scatter workgroup (0 = auto)
    double part_load = 0.5;    // target fill of a pass-3 LDS table
    int hot = 1;               // aggregate heavy-hitter keys during the scatter
    int qp_min = 2;            // sorted-key launches with at least this many queries run query-parallel
    int runs_min_q = 2;        // ... and with at least this many (and runs this long) run-sorted
    double runs_min_len = 8.0; // mean rows per key run for the run-sorted kernel
    int runs_chunks_per_cu = 32;
    int runs_lds_pad = 0;      // extra LDS bytes per write-pass workgroup (caps waves in flight)
    int runs_win = 2048;       // place-pass window (records)
    int runs_qblock = 64;      // queries per staging block (0 = one block per chunk; 16, 32, 64)
    int shared_lean = 1;       // shared-interval groups (and single small-key queries) take the lean pass
    int shared_waves = 4;      // lean pass: workgroup rounds per resident slot
    int sorted1 = 1;           // a single sorted-key query takes the lean single-query pass
    int join_unique = 1;       // DIRECT join indexes of unique build keys hold the build row itself
    int join_parts = 0;        // key-range passes of a join (0 = as the HBM budget needs, else forced)
    int sort_parts = 0;        // passes of a full ORDER BY (0 = as the HBM budget needs, 1 = one, else ~this many)
    int join_claim = 1;        // ... built by one claiming pass over the build rows (fused aggregates)
    int join_payload = 1;      // ... and the fused probe gathers the build columns packed per key slot
    int zone_maps = 1;         // single aggregate queries skip the pages a range filter cannot match
    int topk_bpc = 2;          // top-K scan workgroups per CU (each keeps k candidates)
    int pr_max_s = 6;          // partitioned runs: at most 2^pr_max_s keys per final partition
    int pr_b2 = 0;             // ... key bits split by the second scatter pass (0 = from the partition count)
    int pr_scatter_bpc = 2;    // ... scatter workgroups per CU
  } tune;

  void *alloc(size_t bytes);
  void *alloc_shared(size_t bytes, qm_lane *L);  // the lane's released buffers, the pool, or the allocator
  void release(void *p, size_t bytes);
  void trim_pool();
  qm_lane &lane();        // the lane the calling thread holds (lane 0 outside any call)
  qm_lane *held_lane();   // ... or nullptr
};

namespace qm {

size_t pool_round(size_t bytes);

// RAII device buffer from the context pool
struct DevBuf {
  qm_ctx *ctx = nullptr;
  void *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(qm_ctx *c, size_t bytes) : ctx(c), n(pool_round(bytes)) { p = c->alloc(n); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept { *this = std::move(o); }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { reset(); ctx = o.ctx; p = o.p; n = o.n; o.p = nullptr; o.n = 0; o.ctx = nullptr; }
    return *this;
  }
  ~DevBuf() { reset(); }
  void reset() { if (p && ctx) ctx->release(p, n); p = nullptr; n = 0; }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

struct ColStats {
  bool has_int = false;      // imin/imax valid
  int64_t imin = 0, imax = 0;
  bool has_f = false;        // fmin/fmax valid (non-NaN values exist)
  double fmin = 0, fmax = 0;
  double max_abs = 0;        // largest finite |x|
  double min_abs = 0;        // smallest nonzero finite |x| (floats; 0 = none)
  uint64_t nan = 0, nulls = 0, nonnull = 0;
  double distinct = 0;       // HyperLogLog estimate
  double dup_frac = 0;       // share of a strided host sample whose value repeats in it
  std::vector<int64_t> heavy; // most frequent sample values (integer columns), >= 1/2048 of the sample
  bool sorted = false;        // integer column without NULLs stored in ascending order
};

struct TCol {
  std::string name;
  int dtype = 0;
  void *d = nullptr;
  size_t bytes = 0;
  uint8_t *dnull = nullptr;
  size_t null_bytes = 0;
  ColStats st;
  // zone maps: per ZONE_ROWS-row page, min / max order code of the non-NULL
  // values (empty page: min ~0, max 0)
  std::vector<uint64_t> zmin, zmax;
  // key histogram (integer columns without NULLs whose range spans < 2^32
  // values): rows per bin of 2^khist_g consecutive values from st.imin; sizes
  // the regions of the batched group-by's key partitions
  std::vector<uint32_t> khist;
  int khist_g = -1;
  // partition geometries of the batched group-by derived from khist, by
  // (buckets, states) of the launch (device copies of the region starts)
  struct PRGeo {
    uint32_t nbk = 0, nsn = 0, rw = 0;   // what the geometry was derived for
    int max_s = 0, want_b2 = 0;          // ... and under which knobs
    int s = -1, sh1 = 0;
    uint32_t np1 = 0, np = 0, cap = 0, tile = 0;
    uint64_t total = 0;                  // records the regions hold
    std::shared_ptr<DevBuf> start1, start2, chunk0;
    uint32_t nchunk2 = 0;                // pass-2 workgroup units
  };
  std::deque<PRGeo> prgeo;  // stable addresses: a running call keeps a pointer
};

}  // namespace qm

struct qm_table {
  qm_ctx *ctx = nullptr;
  std::string name;
  uint64_t rows = 0;
  std::atomic<int> prepared{0};  // prepared batches whose graphs read this table's columns
  // a zone-map view: pages of the view its query skips (device bitmap), or nullptr
  const uint32_t *page_skip = nullptr;
  std::shared_ptr<qm::DevBuf> page_skip_buf;
  std::vector<qm::TCol> cols;
  int find(const std::string &n) const {
    for (size_t i = 0; i < cols.size(); ++i) if (cols[i].name == n) return (int)i;
    return -1;
  }
};

struct qm_plan {
  qm_ctx *ctx = nullptr;
  qm::Plan plan;
};

namespace qm {
struct OutCol {
  std::string name;
  int dtype = 0;     // device dtype of `data`
  DevBuf data;
  DevBuf nulls;      // one byte per row, optional
};
}  // namespace qm

struct qm_result {
  qm_ctx *ctx = nullptr;
  uint64_t rows = 0;
  std::vector<qm::OutCol> cols;  // a column may be a view into one of `arenas`
  std::vector<std::shared_ptr<qm::DevBuf>> arenas;  // buffers shared by the results of one launch
  qm_stats stats{};
};

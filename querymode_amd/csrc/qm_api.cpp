// qm_api.cpp — the C-ABI: contexts, resident tables, plan binding, the
// executor that drives the gfx950 kernels, and QMCB result encoding.
//
// Execution mirrors the reference's pipeline assembly (src/operators.ts:
// 3073-3150, src/local-executor.ts:386-508 and 597-667):
//   filters -> [join probe] -> aggregate -> (TopK | sort | limit) -> project
// with every stage on the device; the host only binds names, picks the
// group-by strategy from column statistics and sizes the device tables.
#include <algorithm>
#include <cstring>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <thread>
#include <unordered_map>
#include "qm_host.h"

using namespace qm;

// ===========================================================================
// errors
// ===========================================================================
static thread_local std::string g_last_error;

static int set_error(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

#define API_BEGIN try {
#define API_END                                                         \
  }                                                                     \
  catch (const qm::Error &e) { return set_error(e.code, e.msg); }       \
  catch (const std::bad_alloc &) { return set_error(QM_ERR_MEMORY_EXCEEDED, "host out of memory"); } \
  catch (const std::exception &e) { return set_error(QM_ERR_QUERY_FAILED, e.what()); } \
  catch (...) { return set_error(QM_ERR_QUERY_FAILED, "unknown failure"); }  \
  return QM_OK;

// ===========================================================================
// device memory pool with the HBM budget
// ===========================================================================
namespace qm {
size_t pool_round(size_t bytes) {
  if (bytes < 256) bytes = 256;
  const size_t big = 64ull << 20;
  if (bytes <= big) {
    size_t r = 256;
    while (r < bytes) r <<= 1;
    return r;
  }
  return (bytes + big - 1) / big * big;
}
}  // namespace qm

namespace qm {
// the lane the calling thread holds, per context (set by LaneHold)
thread_local qm_ctx *tl_ctx = nullptr;
thread_local qm_lane *tl_lane = nullptr;
}  // namespace qm

qm_lane *qm_ctx::held_lane() { return qm::tl_ctx == this ? qm::tl_lane : nullptr; }
qm_lane &qm_ctx::lane() {
  qm_lane *l = held_lane();
  return l ? *l : lanes[0];
}

void *qm_ctx::alloc(size_t bytes) {
  qm_lane *L = held_lane();
  if (L && L->cap && L->cap->mode == LANE_CAPTURE) {  // no allocator call inside a stream capture
    auto it = L->cap->free.find(bytes);
    if (it == L->cap->free.end())
      fail(QM_ERR_QUERY_FAILED, "prepared batch: the captured run needs a %zu-byte buffer the recorded run did not",
           bytes);
    void *p = it->second;
    L->cap->free.erase(it);
    return p;
  }
  void *p = alloc_shared(bytes, L);
  if (L && L->cap && L->cap->mode == LANE_RECORD) {
    qm_capture &k = *L->cap;
    k.rec_live[p] = bytes;
    k.peak[bytes] = std::max(k.peak[bytes], ++k.live[bytes]);
  }
  return p;
}

void *qm_ctx::alloc_shared(size_t bytes, qm_lane *L) {
  {
    std::lock_guard<std::mutex> g(pool_mu);
    if (L) {  // a buffer this call released: its stream orders the reuse
      auto it = L->held.find(bytes);
      if (it != L->held.end()) {
        void *p = it->second;
        L->held.erase(it);
        L->held_bytes -= bytes;
        pooled -= bytes;
        return p;
      }
    }
    auto it = pool.find(bytes);
    if (it != pool.end()) {
      void *p = it->second;
      pool.erase(it);
      pooled -= bytes;
      return p;
    }
  }
  if (used.load() + bytes > budget) trim_pool();
  if (used.load() + bytes > budget)
    fail(QM_ERR_MEMORY_EXCEEDED, "HBM budget exceeded: %llu + %llu > %llu bytes", (unsigned long long)used.load(),
         (unsigned long long)bytes, (unsigned long long)budget);
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    trim_pool();
    e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      fail(QM_ERR_MEMORY_EXCEEDED, "device allocation of %llu bytes failed", (unsigned long long)bytes);
    }
  }
  used += bytes;
  return p;
}

void qm_ctx::release(void *p, size_t bytes) {
  qm_lane *L = held_lane();
  if (L && L->cap && L->cap->mode == LANE_CAPTURE) {  // back to the graph's private pool
    L->cap->free.emplace(bytes, p);
    return;
  }
  if (L && L->cap && L->cap->mode == LANE_RECORD) {
    auto it = L->cap->rec_live.find(p);
    if (it != L->cap->rec_live.end()) {
      --L->cap->live[it->second];
      L->cap->rec_live.erase(it);
    }
  }
  std::lock_guard<std::mutex> g(pool_mu);
  if (L) {  // kernels of this call may still use it: other lanes only get it once the stream drained
    L->held.emplace(bytes, p);
    L->held_bytes += bytes;
  } else {
    pool.emplace(bytes, p);
  }
  pooled += bytes;
}

// Frees the shared pool and, when the caller holds a lane, that lane's own
// released buffers once its stream has drained (hipFree of memory a pending
// kernel uses is never issued).
void qm_ctx::trim_pool() {
  qm_lane *L = held_lane();
  if (L && L->cap && L->cap->mode == LANE_CAPTURE) return;  // no frees (or waits) inside a stream capture
  if (L && !L->held.empty()) (void)hipStreamSynchronize(L->stream);
  std::lock_guard<std::mutex> g(pool_mu);
  for (auto &kv : pool) {
    (void)hipFree(kv.second);
    used -= kv.first;
    pooled -= kv.first;
  }
  pool.clear();
  if (L) {
    for (auto &kv : L->held) {
      (void)hipFree(kv.second);
      used -= kv.first;
    }
    pooled -= L->held_bytes;
    L->held.clear();
    L->held_bytes = 0;
  }
}

namespace qm {
// Holds one lane of `ctx` for the calling thread until destroyed; executions
// hold the table lock shared, table upload / drop and knob changes exclusive
// (then no other call runs). On release the lane's stream is drained and the
// buffers the call released go back to the context pool. A call made while
// the thread already holds a lane of the same context keeps that lane.
struct LaneHold {
  qm_ctx *ctx;
  qm_lane *lane = nullptr;
  qm_ctx *prev_ctx;
  qm_lane *prev_lane;
  std::shared_lock<std::shared_mutex> rd;
  std::unique_lock<std::shared_mutex> wr;
  LaneHold(qm_ctx *c, bool exclusive = false) : ctx(c), prev_ctx(tl_ctx), prev_lane(tl_lane) {
    if (tl_ctx == c) return;  // nested: the outer hold covers it
    if (exclusive) wr = std::unique_lock<std::shared_mutex>(c->tables_mu);
    else rd = std::shared_lock<std::shared_mutex>(c->tables_mu);
    {
      std::unique_lock<std::mutex> g(c->lanes_mu);
      for (;;) {
        for (qm_lane &l : c->lanes)
          if (!l.busy) { lane = &l; break; }
        if (lane) break;
        c->lanes_cv.wait(g);
      }
      lane->busy = true;
    }
    tl_ctx = c;
    tl_lane = lane;
    if (hipSetDevice(c->device) != hipSuccess) (void)hipGetLastError();
  }
  ~LaneHold() {
    if (!lane) return;
    (void)hipStreamSynchronize(lane->stream);
    {
      std::lock_guard<std::mutex> g(ctx->pool_mu);
      for (auto &kv : lane->held) ctx->pool.emplace(kv.first, kv.second);
      lane->held.clear();
      lane->held_bytes = 0;
    }
    tl_ctx = prev_ctx;
    tl_lane = prev_lane;
    {
      std::lock_guard<std::mutex> g(ctx->lanes_mu);
      lane->busy = false;
    }
    ctx->lanes_cv.notify_one();
  }
  LaneHold(const LaneHold &) = delete;
  LaneHold &operator=(const LaneHold &) = delete;
};
}  // namespace qm

// ===========================================================================
// helpers
// ===========================================================================
namespace {

int dt_size(int dt) {
  switch (dt) {
    case DT_F64: case DT_I64: return 8;
    case DT_I32: case DT_F32: return 4;
    case DT_I16: return 2;
    case DT_I8: case DT_BOOL: return 1;
  }
  return 0;
}
bool dt_int(int dt) { return dt == DT_I8 || dt == DT_I16 || dt == DT_I32 || dt == DT_I64 || dt == DT_BOOL; }
bool dt_float(int dt) { return dt == DT_F64 || dt == DT_F32; }

uint64_t pad_rows(uint64_t n) { return (n + ROW_PAD - 1) / ROW_PAD * ROW_PAD + ROW_PAD; }
uint64_t pow2ceil(uint64_t x) { uint64_t r = 1; while (r < x) r <<= 1; return r; }

double hll_estimate(const uint32_t *reg) {
  const int m = 4096;
  double sum = 0;
  int zeros = 0;
  for (int i = 0; i < m; ++i) { sum += std::ldexp(1.0, -(int)reg[i]); if (!reg[i]) ++zeros; }
  double alpha = 0.7213 / (1.0 + 1.079 / m);
  double e = alpha * m * (double)m / sum;
  if (e <= 2.5 * m && zeros > 0) e = m * std::log((double)m / zeros);
  return e;
}

qm_capture *lane_cap(qm_ctx *c) { return c->lane().cap; }
bool capturing(qm_ctx *c) {
  const qm_capture *k = lane_cap(c);
  return k && k->mode == LANE_CAPTURE;
}

// Stream waits and event timing; no-ops while the lane's stream is being
// captured (a graph has no host to wait for).
void stream_sync(qm_ctx *c) {
  if (!capturing(c)) HIPX(hipStreamSynchronize(c->lane().stream));
}
void ev_record(qm_ctx *c, hipEvent_t e) {
  if (!capturing(c)) HIPX(hipEventRecord(e, c->lane().stream));
}
void ev_sync(qm_ctx *c, hipEvent_t e) {
  if (!capturing(c)) HIPX(hipEventSynchronize(e));
}
void ev_ms(qm_ctx *c, float *ms, hipEvent_t a, hipEvent_t b) {
  *ms = 0;
  if (!capturing(c)) HIPX(hipEventElapsedTime(ms, a, b));
}

// Device-time bracket on the lane's stream.
struct Timer {
  qm_ctx *c;
  explicit Timer(qm_ctx *ctx) : c(ctx) { ev_record(c, c->lane().ev0); }
  double stop() {
    if (capturing(c)) return 0;
    HIPX(hipEventRecord(c->lane().ev1, c->lane().stream));
    HIPX(hipEventSynchronize(c->lane().ev1));
    float ms = 0;
    HIPX(hipEventElapsedTime(&ms, c->lane().ev0, c->lane().ev1));
    return ms;
  }
};

void h2d_bytes(qm_ctx *c, void *dst, const void *src, size_t bytes) {
  if (!bytes) return;
  qm_capture *k = lane_cap(c);
  if (k && k->mode == LANE_CAPTURE) {
    // the graph's copy node reads pinned memory that lives as long as the graph
    const size_t off = (k->staging_used + 255) & ~(size_t)255;
    if (off + bytes > k->staging_cap) fail(QM_ERR_QUERY_FAILED, "prepared batch: upload staging exhausted");
    memcpy(k->staging + off, src, bytes);
    k->staging_used = off + bytes;
    HIPX(hipMemcpyAsync(dst, k->staging + off, bytes, hipMemcpyHostToDevice, c->lane().stream));
    return;
  }
  if (k) k->h2d_bytes += ((bytes + 255) & ~(size_t)255) + 256;
  HIPX(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->lane().stream));
}
template <class T> void h2d(qm_ctx *c, void *dst, const T *src, size_t n) { h2d_bytes(c, dst, src, n * sizeof(T)); }

// Device-to-host readback of `rows` rows of `width` bytes, `pitch` apart on the
// device (pitch == width: one contiguous block). Recorded on the tape when
// recording, taken from it when capturing.
void d2h_rows(qm_ctx *c, void *dst, const void *src, size_t width, size_t pitch, size_t rows) {
  const size_t bytes = width * rows;
  if (!bytes) return;
  qm_capture *k = lane_cap(c);
  if (k && k->mode == LANE_CAPTURE) {
    if (k->pos >= k->tape.size() || k->tape[k->pos].size() != bytes)
      fail(QM_ERR_QUERY_FAILED, "prepared batch: the captured run diverged from the recorded one");
    memcpy(dst, k->tape[k->pos].data(), bytes);
    ++k->pos;
    return;
  }
  DevBuf rows_dev;
  if (pitch == width) {
    HIPX(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->lane().stream));
  } else {
    // the rows are gathered on the device first: a strided copy into pageable
    // host memory goes row by row (≈ 7 µs per row)
    rows_dev = DevBuf(c, bytes);
    HIPX(hipMemcpy2DAsync(rows_dev.p, width, src, pitch, width, rows, hipMemcpyDeviceToDevice, c->lane().stream));
    HIPX(hipMemcpyAsync(dst, rows_dev.p, bytes, hipMemcpyDeviceToHost, c->lane().stream));
  }
  HIPX(hipStreamSynchronize(c->lane().stream));
  if (k && k->mode == LANE_RECORD) {
    const uint8_t *b = static_cast<const uint8_t *>(dst);
    k->tape.emplace_back(b, b + bytes);
  }
}
template <class T> void d2h(qm_ctx *c, T *dst, const void *src, size_t n) {
  d2h_rows(c, dst, src, n * sizeof(T), n * sizeof(T), 1);
}
void zero(qm_ctx *c, void *p, size_t bytes) {
  if (!bytes) return;
  if (capturing(c)) HIPX(qmk_zero_bytes(p, bytes, c->lane().stream));  // no memset nodes in graphs
  else HIPX(hipMemsetAsync(p, 0, bytes, c->lane().stream));
}

int qmcb_tag(int dt) {
  switch (dt) {
    case DT_F64: return QM_F64;
    case DT_I64: return QM_I64;
    case DT_I32: return QM_I32;
    case DT_F32: return QM_F32;
    case DT_BOOL: return QM_BOOL;
  }
  return QM_I32;
}
// result dtype of a gathered column (QMCB has no I8/I16 tag)
int out_dt_of(int dt) { return (dt == DT_I8 || dt == DT_I16) ? DT_I32 : dt; }

// ===========================================================================
// column space: names -> DScan.cols slots (left = probe/scanned, right = build)
// ===========================================================================
struct Ref {
  int slot_a = -1, slot_b = -1;  // product when slot_b >= 0
  bool product() const { return slot_b >= 0; }
};

struct ColSpace {
  const qm_table *L = nullptr, *R = nullptr;
  std::string rkey;
  const Plan *plan = nullptr;
  std::vector<DCol> cols;
  std::vector<const TCol *> tcols;

  int slot_of(const TCol *t, int side) {
    for (size_t i = 0; i < tcols.size(); ++i)
      if (tcols[i] == t && cols[i].side == side) return (int)i;
    if ((int)cols.size() >= MAX_COLS) fail(QM_ERR_UNSUPPORTED, "query references more than %d columns", MAX_COLS);
    DCol c;
    c.data = t->d;
    c.nulls = t->dnull;
    c.dtype = t->dtype;
    c.side = side;
    cols.push_back(c);
    tcols.push_back(t);
    return (int)cols.size() - 1;
  }
  // plain column of the (joined) row, src/operators.ts:2278-2293 naming
  int plain(const std::string &name) {
    int i = L->find(name);
    if (i >= 0) return slot_of(&L->cols[i], 0);
    if (R) {
      if (name != rkey) {
        int j = R->find(name);
        if (j >= 0) return slot_of(&R->cols[j], 1);
      }
      if (name.rfind("right_", 0) == 0) {
        std::string base = name.substr(6);
        int j = R->find(base);
        if (j >= 0 && base != rkey && L->find(base) >= 0) return slot_of(&R->cols[j], 1);
      }
    }
    fail(QM_ERR_COLUMN_NOT_FOUND, "column not found: %s", name.c_str());
  }
  // aggregate input: plain column, computed product alias, or multiply(a,b)
  Ref value(const std::string &name) {
    Ref r;
    for (const Computed &c : plan->computed) {
      if (c.alias == name) { r.slot_a = plain(c.a); r.slot_b = plain(c.b); return r; }
    }
    if (name.rfind("multiply(", 0) == 0 && name.back() == ')') {
      std::string in = name.substr(9, name.size() - 10);
      size_t comma = in.find(',');
      if (comma != std::string::npos) {
        auto trim = [](std::string s) {
          while (!s.empty() && s.front() == ' ') s.erase(s.begin());
          while (!s.empty() && s.back() == ' ') s.pop_back();
          return s;
        };
        r.slot_a = plain(trim(in.substr(0, comma)));
        r.slot_b = plain(trim(in.substr(comma + 1)));
        return r;
      }
    }
    r.slot_a = plain(name);
    return r;
  }
  const ColStats &st(int slot) const { return tcols[slot]->st; }
  bool nullable(int slot) const { return tcols[slot]->st.nulls > 0; }
};

// ===========================================================================
// binding filters into a DQuery
// ===========================================================================
struct ListStore {
  std::vector<DevBuf> bufs;
};

void bind_filters(qm_ctx *ctx, ColSpace &cs, const std::vector<Filter> &ands,
                  const std::vector<std::vector<Filter>> &groups, DQuery &q, ListStore &ls, bool right_side) {
  q.npreds = 0;
  q.ngroups = 0;
  auto add = [&](const Filter &f, int group) {
    if (q.npreds >= MAX_PREDS) fail(QM_ERR_UNSUPPORTED, "more than %d filter predicates", MAX_PREDS);
    int slot;
    if (right_side) {
      int j = cs.R->find(f.col);
      if (j < 0) fail(QM_ERR_COLUMN_NOT_FOUND, "column not found: %s", f.col.c_str());
      slot = cs.slot_of(&cs.R->cols[j], 0);
    } else {
      int j = cs.L->find(f.col);
      if (j < 0) fail(QM_ERR_COLUMN_NOT_FOUND, "column not found: %s", f.col.c_str());
      slot = cs.slot_of(&cs.L->cols[j], 0);
    }
    BoundPred bp = coerce_filter(f, cs.cols[slot].dtype);
    bp.p.col = (int8_t)slot;
    bp.p.group = (uint8_t)group;
    if (!bp.ilist.empty()) {
      DevBuf b(ctx, bp.ilist.size() * 8);
      h2d(ctx, b.p, bp.ilist.data(), bp.ilist.size());
      bp.p.list = b.p;
      bp.p.list_n = (int)bp.ilist.size();
      ls.bufs.push_back(std::move(b));
    } else if (!bp.flist.empty() || bp.p.kind == PK_F64_IN) {
      size_t n = bp.flist.size();
      DevBuf b(ctx, (n ? n : 1) * 8);
      h2d(ctx, b.p, bp.flist.data(), n);
      bp.p.list = b.p;
      bp.p.list_n = (int)n;
      ls.bufs.push_back(std::move(b));
    }
    q.preds[q.npreds++] = bp.p;
  };
  for (const Filter &f : ands) add(f, 0);
  // An empty OR group is always true (Array.prototype.every on []).
  bool trivially_true = false;
  for (const auto &g : groups) if (g.empty()) trivially_true = true;
  if (!groups.empty() && !trivially_true) {
    if (groups.size() > 250) fail(QM_ERR_UNSUPPORTED, "too many OR groups");
    int gi = 1;
    for (const auto &g : groups) {
      for (const Filter &f : g) add(f, gi);
      ++gi;
    }
    q.ngroups = (int)groups.size();
  }
  // AND-only fast path: predicate columns held in registers
  q.and_only = 0;
  q.nslots = 0;
  if (q.ngroups == 0) {
    int ns = 0;
    bool ok = true;
    for (int i = 0; i < q.npreds && ok; ++i) {
      int s = -1;
      for (int k = 0; k < ns; ++k) if (q.slot_col[k] == q.preds[i].col) s = k;
      if (s < 0) {
        if (ns >= MAX_PSLOTS) { ok = false; break; }
        q.slot_col[ns] = q.preds[i].col;
        s = ns++;
      }
      q.preds[i].slot = (int8_t)s;
    }
    if (ok) {
      q.and_only = 1;
      q.nslots = (int8_t)ns;
      // merge the predicates of a column into one range where they are plain
      // (non-negated) ranges; IS NOT NULL is implied by any range
      DPred merged[MAX_PREDS];
      int nm = 0;
      bool rewrite = true;
      for (int s = 0; s < ns && rewrite; ++s) {
        int idx[MAX_PREDS];
        int nr = 0;
        bool is_false = false;
        for (int i = 0; i < q.npreds; ++i) {
          const DPred &p = q.preds[i];
          if (p.slot != s) continue;
          if (p.kind == PK_FALSE) is_false = true;
          else if (p.kind != PK_IS_NOT_NULL) idx[nr++] = i;
        }
        DPred out;
        memset(&out, 0, sizeof out);
        if (is_false) {
          out.kind = PK_FALSE;
        } else if (nr == 0) {
          out.kind = PK_IS_NOT_NULL;
        } else if (nr == 1) {
          out = q.preds[idx[0]];
        } else {
          DPred acc = q.preds[idx[0]];
          for (int k = 0; k < nr && rewrite; ++k) {
            const DPred &p = q.preds[idx[k]];
            const bool ir = p.kind == PK_INT_RANGE && !p.negate;
            const bool fr = p.kind == PK_F64_RANGE && !p.negate && !p.nan_result;
            if ((!ir && !fr) || p.kind != acc.kind) { rewrite = false; break; }
            if (k == 0) continue;
            if (ir) {
              if (p.ilo > acc.ilo) acc.ilo = p.ilo;
              if (p.ihi < acc.ihi) acc.ihi = p.ihi;
            } else {
              if (p.flo > acc.flo || (p.flo == acc.flo && p.lo_strict)) { acc.flo = p.flo; acc.lo_strict = p.lo_strict; }
              if (p.fhi < acc.fhi || (p.fhi == acc.fhi && p.hi_strict)) { acc.fhi = p.fhi; acc.hi_strict = p.hi_strict; }
            }
          }
          if (!rewrite) break;
          out = acc;
          if (out.kind == PK_INT_RANGE && out.ilo > out.ihi) { memset(&out, 0, sizeof out); out.kind = PK_FALSE; }
        }
        out.col = q.slot_col[s];
        out.slot = (int8_t)s;
        out.group = 0;
        merged[nm++] = out;
      }
      if (rewrite) {
        for (int i = 0; i < nm; ++i) q.preds[i] = merged[i];
        q.npreds = nm;
      }
    }
  }
}

// ===========================================================================
// aggregate binding and strategy
// ===========================================================================
struct AggOut {
  std::string name;
  DOut o;
};

struct AggBinding {
  DQuery q;
  std::vector<AggOut> outs;
  std::vector<int> key_slots;
  bool always_one_row = false;
  uint64_t nslots = 0;
  uint64_t table_words = 0;
  // partitioned aggregation (high-cardinality keys)
  bool part = false;
  int pbits = 0, sbits = 0, nparts = 0, part_lc = 0;
  bool has_mask = false;
  std::vector<std::pair<int, int>> srcs;  // (col_a, col_b)
  int8_t state_src[MAX_STATES] = {};
  double est_groups = 0;
  std::vector<uint64_t> hot_table;  // heavy-hitter slots (HOT_SLOTS), empty = none
};

const int PART_LDS_WORDS = 128 * 1024 / 8;  // pass-3 LDS table budget (one 1024-thread workgroup per CU)

// Decide the partitioned strategy and its record layout.
void plan_partitioned(const qm_ctx *ctx, ColSpace &cs, uint64_t nrows, double est_groups, bool fused_join,
                      AggBinding &ab) {
  DQuery &q = ab.q;
  ab.part = false;
  if (fused_join || nrows < (1u << 20) || q.key.ncols == 0) return;
  const int W = q.entry_words;
  int lc = 1;
  while (lc * 2 * W <= PART_LDS_WORDS) lc *= 2;
  if (q.gmode == GM_HASH) {
    if (est_groups <= 32768.0) return;
    // pass-3 tables use the whole LDS budget at the target fill; the count
    // is a multiple of the CU count so the persistent pass 3 ends in step
    uint64_t parts = (uint64_t)std::ceil(est_groups / (lc * ctx->tune.part_load));
    const uint64_t cus = (uint64_t)std::max(1, ctx->num_cus);
    if (parts > cus) parts = (parts + cus - 1) / cus * cus;
    if (parts > 8192) return;
    ab.nparts = (int)std::max<uint64_t>(parts, 2);
    ab.pbits = 1;  // hash-partitioned
  } else {
    if (q.lmode != LM_NONE || q.gcap < 32768) return;
    uint64_t parts = (q.gcap + lc - 1) / lc;
    if (parts > 8192) return;
    int sb = 0;
    while ((1 << sb) < lc) ++sb;
    ab.sbits = sb;
    ab.nparts = (int)parts;
  }
  ab.part_lc = lc;
  // value sources of the states
  ab.srcs.clear();
  bool nullable_src = false;
  for (int si = 0; si < q.nstates; ++si) {
    const DState &s = q.states[si];
    if (s.col_a < 0) { ab.state_src[si] = -1; continue; }
    std::pair<int, int> key{s.col_a, s.product ? s.col_b : -1};
    int idx = -1;
    for (size_t j = 0; j < ab.srcs.size(); ++j) if (ab.srcs[j] == key) idx = (int)j;
    if (idx < 0) { ab.srcs.push_back(key); idx = (int)ab.srcs.size() - 1; }
    ab.state_src[si] = (int8_t)idx;
    if (cs.nullable(s.col_a) || (s.product && cs.nullable(s.col_b))) nullable_src = true;
  }
  if ((int)ab.srcs.size() > 6) return;  // record = key + <= 6 sources + mask in 8 words
  bool side_keys = false;
  if (q.gmode == GM_HASH) {
    for (int i = 0; i < q.key.ncols; ++i) if (q.key.nullable[i]) side_keys = true;
    if (q.key.ncols == 1) {
      const ColStats &st = cs.st(q.key.col[0]);
      if (q.key.is_float[0]) side_keys = true;
      else if (st.has_int && (int64_t)KEY_UNREPRESENTABLE >= st.imin && (int64_t)KEY_UNREPRESENTABLE <= st.imax)
        side_keys = true;
    }
  }
  ab.has_mask = nullable_src || side_keys;
  // heavy hitters of a single integer key (from the upload sample)
  ab.hot_table.clear();
  if (ctx->tune.hot && q.key.ncols == 1 && !q.key.is_float[0]) {
    const ColStats &st = cs.st(q.key.col[0]);
    if (st.heavy.size() >= 2) {
      std::vector<uint64_t> tab(HOT_SLOTS, 0);
      int placed = 0;
      for (int64_t v : st.heavy) {
        uint64_t hk;
        if (q.gmode == GM_DIRECT) {
          if (v < q.key.kmin[0] || (uint64_t)(v - q.key.kmin[0]) >= q.key.dom[0]) continue;
          hk = (uint64_t)(v - q.key.kmin[0]) + 1;
        } else {
          hk = (uint64_t)v + KEY_SALT;
          if (hk == 0) continue;
        }
        uint32_t h = (uint32_t)mix64(hk ^ 0xA5A5A5A5A5A5A5A5ull) & (HOT_SLOTS - 1);
        for (int pr = 0; pr < HOT_PROBES; ++pr) {
          uint32_t sl = (h + pr) & (HOT_SLOTS - 1);
          if (tab[sl] == 0) { tab[sl] = hk; ++placed; break; }
        }
      }
      if (placed >= 2) ab.hot_table = tab;
    }
  }
  q.lmode = LM_NONE;
  q.lds_words = 0;
  ab.part = true;
}

const int LDS_BLOCK_BUDGET_WORDS = 48 * 1024 / 8;  // 48 KiB of LDS per workgroup for tables

struct StateReq {
  uint8_t kind, is_min, fixed, product;
  int8_t a, b;
  int scale;
  int nw = 0;  // SK_SUM_WIDE digit words
};

void bind_aggregate(qm_ctx *ctx, ColSpace &cs, const Plan &pl, uint64_t nrows, bool fused_join, AggBinding &ab,
                    uint32_t regrow) {
  DQuery &q = ab.q;
  std::vector<StateReq> reqs;
  auto want = [&](StateReq r) -> int {
    for (size_t i = 0; i < reqs.size(); ++i) {
      const StateReq &x = reqs[i];
      if (x.kind == r.kind && x.is_min == r.is_min && x.a == r.a && x.b == r.b && x.product == r.product)
        return (int)i;
    }
    reqs.push_back(r);
    return (int)reqs.size() - 1;
  };
  // state 0: row count (COUNT(*), occupancy)
  want(StateReq{SK_COUNT, 0, 0, 0, -1, -1, 0});

  // ---- keys ----
  const int nk = (int)pl.groupby.size();
  if (nk > MAX_KEYS) fail(QM_ERR_UNSUPPORTED, "GROUP BY over more than %d columns", MAX_KEYS);
  DKey &key = q.key;
  memset(&key, 0, sizeof key);
  key.ncols = (int8_t)nk;
  bool all_int = true;
  long double dom_prod = 1;
  double distinct_prod = 1;
  for (int i = 0; i < nk; ++i) {
    int s = cs.plain(pl.groupby[i]);
    ab.key_slots.push_back(s);
    key.col[i] = (int8_t)s;
    const TCol *t = cs.tcols[s];
    key.is_float[i] = dt_float(t->dtype) ? 1 : 0;
    key.nullable[i] = t->st.nulls > 0 ? 1 : 0;
    if (dt_float(t->dtype)) {
      all_int = false;
    } else {
      uint64_t d = t->st.has_int ? (uint64_t)(t->st.imax - t->st.imin) + 1 : 0;
      if (t->st.has_int && d == 0) d = ~0ull;  // full 64-bit range
      key.kmin[i] = t->st.has_int ? t->st.imin : 0;
      key.dom[i] = d + (key.nullable[i] ? 1 : 0);
      if (key.dom[i] == 0) key.dom[i] = 1;
      dom_prod *= (long double)key.dom[i];
    }
    distinct_prod *= std::max(1.0, t->st.distinct * 1.05 + (key.nullable[i] ? 1 : 0));
  }

  // ---- states and outputs ----
  struct Pending { int req; int cnt_req; uint8_t kind; int out_dt; bool src_float; int scale; std::string name; };
  std::vector<Pending> pend;
  auto nonnull_count = [&](const Ref &r) -> int {
    bool nullable = cs.nullable(r.slot_a) || (r.product() && cs.nullable(r.slot_b));
    if (!nullable) return 0;
    return want(StateReq{SK_COUNT, 0, 0, (uint8_t)(r.product() ? 1 : 0), (int8_t)r.slot_a, (int8_t)r.slot_b, 0});
  };
  for (const Agg &a : pl.aggs) {
    if (a.fn >= AGG_COUNT_DISTINCT)
      fail(QM_ERR_UNSUPPORTED, "aggregate %s: only sum/avg/min/max/count run on the device", a.alias.c_str());
    if (a.col == "*") {
      if (a.fn != AGG_COUNT) fail(QM_ERR_INVALID_AGGREGATE, "aggregate %s over '*'", a.alias.c_str());
      pend.push_back(Pending{-1, 0, OK_COUNT, DT_I64, false, 0, a.alias});
      continue;
    }
    Ref r = cs.value(a.col);
    const int dta = cs.cols[r.slot_a].dtype;
    const bool isf = r.product() || dt_float(dta);
    if (dta == DT_BOOL && a.fn != AGG_COUNT)
      fail(QM_ERR_UNSUPPORTED, "aggregate %s over a boolean column", a.alias.c_str());
    int cnt = nonnull_count(r);
    if (a.fn == AGG_COUNT) { pend.push_back(Pending{-1, cnt, OK_COUNT, DT_I64, false, 0, a.alias}); continue; }
    if (a.fn == AGG_SUM || a.fn == AGG_AVG) {
      int scale = 0;
      StateReq sr{0, 0, 0, (uint8_t)(r.product() ? 1 : 0), (int8_t)r.slot_a, (int8_t)r.slot_b, 0};
      uint8_t ok;
      int odt;
      if (isf) {
        double bound = cs.st(r.slot_a).max_abs;
        if (r.product()) {
          const int dtb = cs.cols[r.slot_b].dtype;
          double ba = dt_float(dta) ? cs.st(r.slot_a).max_abs
                                    : std::max(std::fabs((double)cs.st(r.slot_a).imin), std::fabs((double)cs.st(r.slot_a).imax));
          double bb = dt_float(dtb) ? cs.st(r.slot_b).max_abs
                                    : std::max(std::fabs((double)cs.st(r.slot_b).imin), std::fabs((double)cs.st(r.slot_b).imax));
          bound = ba * bb * 1.0000001;
          if (!std::isfinite(bound)) bound = 1.7976931348623157e308;
        }
        scale = fixed_scale_for(bound);
        // the smallest nonzero |x| (a product's: of the operands' smallest,
        // halved for its rounding; integers: 1)
        auto low_of = [&](int slot) {
          const ColStats &st = cs.st(slot);
          return dt_float(cs.cols[slot].dtype) ? st.min_abs : 1.0;
        };
        double low = low_of(r.slot_a);
        if (r.product()) low = low * low_of(r.slot_b) * 0.5;
        int wbase = 0, wnw = 0;
        if (low > 0 && bound > 0) wide_layout(low, bound, &wbase, &wnw);
        if (low > 0 && bound > 0 && wbase < -scale) {
          // some value has bits below the 128-bit fixed point's resolution:
          // digit words over the column's whole exponent range instead
          sr.kind = SK_SUM_WIDE; sr.scale = wbase; sr.nw = wnw;
          scale = wbase;
          ok = (a.fn == AGG_SUM) ? OK_SUM_WIDE : OK_AVG_WIDE;
        } else {
          sr.kind = SK_SUM_128; sr.fixed = 1; sr.scale = scale;
          ok = (a.fn == AGG_SUM) ? OK_SUM_F64 : OK_AVG_F64;
        }
        odt = DT_F64;
      } else if (dta == DT_I64) {
        sr.kind = SK_SUM_128;
        ok = (a.fn == AGG_SUM) ? OK_SUM_I128 : OK_AVG_BIG;
        odt = (a.fn == AGG_SUM) ? DT_I64 : DT_F64;
      } else {
        sr.kind = SK_SUM_I64;
        ok = (a.fn == AGG_SUM) ? OK_SUM_I64 : OK_AVG_I64;
        odt = (a.fn == AGG_SUM) ? DT_I64 : DT_F64;
      }
      int ri = want(sr);
      reqs[ri].scale = scale;
      reqs[ri].nw = sr.nw;
      pend.push_back(Pending{ri, cnt, ok, odt, isf, scale, a.alias});
      continue;
    }
    // MIN / MAX
    StateReq sr{SK_MAX_CODE, (uint8_t)(a.fn == AGG_MIN ? 1 : 0), 0, (uint8_t)(r.product() ? 1 : 0),
                (int8_t)r.slot_a, (int8_t)r.slot_b, 0};
    int ri = want(sr);
    int odt = r.product() ? DT_F64 : out_dt_of(dta);
    pend.push_back(Pending{ri, cnt, (uint8_t)(a.fn == AGG_MIN ? OK_MIN : OK_MAX), odt, isf, 0, a.alias});
  }
  if (reqs.size() > (size_t)MAX_STATES) fail(QM_ERR_UNSUPPORTED, "too many aggregate states");

  // ---- strategy ----
  const bool direct_ok = nk == 0 || (all_int && dom_prod <= (long double)std::max<uint64_t>(1ull << 16, 2 * nrows) &&
                                     dom_prod <= (long double)(1ull << 30));
  q.gmode = direct_ok ? GM_DIRECT : GM_HASH;
  const int key_words = (q.gmode == GM_HASH) ? 1 : 0;
  int w = key_words;
  q.nstates = (int)reqs.size();
  for (size_t i = 0; i < reqs.size(); ++i) {
    DState &s = q.states[i];
    memset(&s, 0, sizeof s);
    s.kind = reqs[i].kind;
    s.is_min = reqs[i].is_min;
    s.fixed = reqs[i].fixed;
    s.product = reqs[i].product;
    s.col_a = reqs[i].a;
    s.col_b = reqs[i].b;
    s.scale = reqs[i].scale;
    s.nw = reqs[i].nw;
    s.word = (int16_t)w;
    w += (s.kind == SK_SUM_128) ? (s.fixed ? 3 : 2) : (s.kind == SK_SUM_WIDE ? 1 + s.nw : 1);
  }
  bool wide = false;
  for (int i = 0; i < q.nstates; ++i) wide = wide || q.states[i].kind == SK_SUM_WIDE;
  if (w > (wide ? 160 : 24)) fail(QM_ERR_UNSUPPORTED, "aggregate state too wide (%d words)", w);
  q.entry_words = w;
  q.count_word = q.states[0].word;

  if (q.gmode == GM_DIRECT) {
    uint64_t dom = 1;
    for (int i = nk - 1; i >= 0; --i) { key.stride[i] = dom; dom *= key.dom[i]; }
    q.gcap = dom;
    ab.nslots = dom;
    ab.always_one_row = (nk == 0);
    // LDS: replicated dense table when the domain fits
    uint64_t ent_words = dom * (uint64_t)w;
    if (ent_words <= (uint64_t)LDS_BLOCK_BUDGET_WORDS) {
      uint64_t stride = ent_words;
      if (stride % 32 != 1) stride += (33 - stride % 32) % 32;  // bank spread between replicas
      int reps = 1;
      while (reps < 64 && (uint64_t)(reps * 2) * stride <= (uint64_t)LDS_BLOCK_BUDGET_WORDS) reps *= 2;
      q.lmode = LM_DIRECT;
      q.lds_entries = (int)dom;
      q.lds_replicas = reps;
      q.lds_stride = (int)stride;
      q.lds_words = (int)(stride * reps);
    } else {
      q.lmode = LM_NONE;
    }
    ab.table_words = dom * w;
  } else {
    // HASH: size from the HLL distinct estimate (src/partial-agg.ts keeps one
    // Map entry per distinct key; here the table is sized once up front).
    if (!all_int && nk > 1) fail(QM_ERR_UNSUPPORTED, "multi-column GROUP BY with a float key");
    if (nk == 2) {
      for (int i = 0; i < 2; ++i) {
        int bits = 1;
        while (bits < 63 && (bits >= 64 ? false : ((1ull << bits) < key.dom[i]))) ++bits;
        key.bits[i] = bits;
      }
      if (key.bits[0] + key.bits[1] > 63) fail(QM_ERR_UNSUPPORTED, "GROUP BY key domain exceeds 63 bits");
    }
    double est = std::min((double)nrows, distinct_prod);
    ab.est_groups = est;
    uint64_t cap = pow2ceil((uint64_t)std::max(2048.0, est * 2.0));
    cap <<= (2 * regrow);
    q.gcap = cap;
    ab.nslots = cap + 2;
    ab.table_words = (cap + 2) * (uint64_t)w;
    // LDS cache for skewed keys
    int lw = LDS_BLOCK_BUDGET_WORDS / w;
    int le = 1;
    while (le * 2 <= lw) le *= 2;
    bool skewed = false;
    for (int s : ab.key_slots) if (cs.st(s).dup_frac >= 0.05) skewed = true;
    if (est <= (double)le * 8) skewed = true;
    if (skewed && !fused_join && le >= 64) {
      q.lmode = LM_HASH;
      q.lds_entries = le;
      q.lds_words = le * w;
    } else {
      q.lmode = LM_NONE;
    }
  }
  if (fused_join && q.lmode == LM_HASH) q.lmode = LM_NONE;
  if (wide && w > 24) {  // entries wider than the LDS merge buffer update the global table directly
    q.lmode = LM_NONE;
    q.lds_words = 0;
  }

  if (!wide) plan_partitioned(ctx, cs, nrows, ab.est_groups, fused_join, ab);

  // ---- outputs ----
  for (int i = 0; i < nk; ++i) {
    AggOut ao;
    ao.name = pl.groupby[i];
    memset(&ao.o, 0, sizeof ao.o);
    ao.o.kind = OK_KEY;
    ao.o.key_idx = (uint8_t)i;
    const int dt = cs.cols[ab.key_slots[i]].dtype;
    ao.o.out_dtype = dt_float(dt) ? DT_F64 : out_dt_of(dt);
    ao.o.cnt_word = -1;
    ab.outs.push_back(ao);
  }
  for (const Pending &p : pend) {
    AggOut ao;
    ao.name = p.name;
    memset(&ao.o, 0, sizeof ao.o);
    ao.o.kind = p.kind;
    ao.o.out_dtype = p.out_dt;
    ao.o.src_is_float = p.src_float ? 1 : 0;
    ao.o.scale = p.scale;
    ao.o.nw = (p.req >= 0 && q.states[p.req].kind == SK_SUM_WIDE) ? (uint8_t)q.states[p.req].nw : 0;
    ao.o.word = (p.req >= 0) ? q.states[p.req].word : 0;
    ao.o.cnt_word = q.states[p.cnt_req].word;
    ab.outs.push_back(ao);
  }
  if ((int)ab.outs.size() > MAX_OUTS) fail(QM_ERR_UNSUPPORTED, "too many output columns");
  (void)ctx;
}

// ===========================================================================
// result assembly helpers
// ===========================================================================
OutCol make_out(qm_ctx *ctx, const std::string &name, int dt, uint64_t rows, bool with_nulls, bool zero_nulls = true) {
  OutCol oc;
  oc.name = name;
  oc.dtype = dt;
  uint64_t pr = pad_rows(rows);
  oc.data = DevBuf(ctx, pr * dt_size(dt));
  if (with_nulls) {
    oc.nulls = DevBuf(ctx, pr / 8 + 64);
    if (zero_nulls) zero(ctx, oc.nulls.p, pr / 8 + 64);
  }
  return oc;
}

DCol dcol_of(const OutCol &o) {
  DCol c;
  c.data = o.data.p;
  c.nulls = o.nulls.as<uint8_t>();
  c.dtype = o.dtype;
  c.side = 0;
  return c;
}

// Gather every column of `res` through `rows` (n entries) into a new result.
void permute_result(qm_result *res, const uint64_t *d_rows, uint64_t n) {
  qm_ctx *ctx = res->ctx;
  std::vector<OutCol> nc;
  for (OutCol &o : res->cols) {
    OutCol x = make_out(ctx, o.name, o.dtype, n, o.nulls.p != nullptr);
    DCol c = dcol_of(o);
    HIPX(qmk_gather(&c, d_rows, nullptr, n, o.dtype, x.data.p, x.nulls.as<uint32_t>(), ctx->lane().stream));
    nc.push_back(std::move(x));
  }
  res->cols = std::move(nc);
  res->rows = n;
}

// rows of a sorted/limited order over one column `c` of length n:
// returns device row ids (ascending order positions) for [offset, offset+limit)
DevBuf select_kept(qm_ctx *ctx, const uint8_t *keep, uint64_t n, uint64_t *n_out);

// Rows per pass of a full ORDER BY whose one-pass temporaries would not fit
// half the HBM budget left (0: one pass), or n / sort_parts when that knob is
// above 1 (1: always one pass). A pass holds its rows' ids and keys twice plus
// the radix sort's scratch.
uint64_t sort_pass_rows(qm_ctx *ctx, uint64_t n, size_t one_pass_bytes) {
  if (n < 2 || ctx->tune.sort_parts == 1) return 0;
  if (ctx->tune.sort_parts > 1) return (n + ctx->tune.sort_parts - 1) / ctx->tune.sort_parts;
  uint64_t used;
  {
    std::lock_guard<std::mutex> g(ctx->pool_mu);
    used = ctx->used.load() - ctx->pooled;
  }
  const uint64_t room = (ctx->budget > used ? ctx->budget - used : 0) / 2;
  if (one_pass_bytes + n * 8 <= room) return 0;
  const uint64_t fixed = 2 * n + (1u << 20);  // the pass mask (pool-rounded), the key histogram
  // ids, keys twice, the sorted ids and the radix scratch: 48 bytes a row,
  // up to 64 once the pool rounds each buffer up to a power of two
  return room > fixed ? std::max<uint64_t>(1, (room - fixed) / 64) : 1;
}

// The stable sort of `order_rows` in passes over ranges of the sort key (the
// reference sorts runs that fit its budget and merges them from spill files,
// src/operators.ts:2811-3010; here the column stays resident and each pass
// takes the rows of one key range, so no merge is needed): a histogram of the
// top 16 key bits cuts the key space into ranges of at most `per_pass` rows
// (a single hotter bin is one range); each pass selects its rows in row order,
// sorts them stably by key, and its place in the order is the rows before it;
// NULL and NaN rows come last in row order. Only the passes overlapping
// [start, start + take) are sorted and copied into `rows`.
void sort_in_passes(qm_ctx *ctx, const DCol &c, uint64_t n, bool desc, uint64_t start, uint64_t take,
                    uint64_t per_pass, DevBuf &rows, uint32_t *launches) {
  const int d = desc ? 1 : 0;
  DevBuf hist(ctx, 65536 * 4);
  zero(ctx, hist.p, 65536 * 4);
  HIPX(qmk_sort_hist(&c, n, d, hist.as<uint32_t>(), ctx->lane().stream));
  std::vector<uint32_t> h(65536);
  d2h(ctx, h.data(), hist.p, 65536);
  std::vector<std::pair<uint64_t, uint64_t>> ranges;  // inclusive key bounds
  uint64_t acc = 0;
  uint32_t b0 = 0;
  for (uint32_t b = 0; b < 65536; ++b) {
    if (acc && acc + h[b] > per_pass) {
      ranges.push_back({(uint64_t)b0 << 48, ((uint64_t)b << 48) - 1});
      b0 = b;
      acc = 0;
    }
    acc += h[b];
  }
  ranges.push_back({(uint64_t)b0 << 48, ~0ull});
  DevBuf keep(ctx, n);
  uint64_t at = 0;
  const uint64_t end = start + take;
  for (size_t pi = 0; pi <= ranges.size() && at < end; ++pi) {
    const bool nulls = pi == ranges.size();
    HIPX(qmk_sort_range(&c, n, d, nulls ? 0 : ranges[pi].first, nulls ? 0 : ranges[pi].second, nulls ? 1 : 0,
                        keep.as<uint8_t>(), ctx->lane().stream));
    uint64_t m = 0;
    DevBuf ids = select_kept(ctx, keep.as<uint8_t>(), n, &m);
    *launches += 4;
    if (at + m <= start) {  // wholly before the slice
      at += m;
      continue;
    }
    const uint64_t *src = ids.as<uint64_t>();
    DevBuf k0, k1, sorted, tmp;
    if (!nulls && m > 1) {
      k0 = DevBuf(ctx, m * 8);
      k1 = DevBuf(ctx, m * 8);
      sorted = DevBuf(ctx, m * 8);
      tmp = DevBuf(ctx, std::max<size_t>(qmk_radix_pairs_tmp_bytes(m), 256));
      HIPX(qmk_sort_gather_keys(&c, ids.as<uint64_t>(), m, d, k0.as<uint64_t>(), ctx->lane().stream));
      HIPX(qmk_radix_pairs(k0.as<uint64_t>(), k1.as<uint64_t>(), ids.as<uint64_t>(), sorted.as<uint64_t>(), m, 64,
                           tmp.p, tmp.n, ctx->lane().stream));
      src = sorted.as<uint64_t>();
      *launches += 3;
    }
    const uint64_t lo = std::max(start, at), hi = std::min(end, at + m);
    if (hi > lo)
      HIPX(hipMemcpyAsync(rows.as<uint64_t>() + (lo - start), src + (lo - at), (hi - lo) * 8, hipMemcpyDeviceToDevice,
                          ctx->lane().stream));
    at += m;
  }
}

DevBuf order_rows(qm_ctx *ctx, const DCol &c, uint64_t n, bool desc, int64_t limit, int64_t offset, uint64_t *out_n,
                  const DQuery *d_noq, uint32_t *launches) {
  uint64_t start = (uint64_t)offset;
  uint64_t want = (limit >= 0) ? (uint64_t)limit : (n > start ? n - start : 0);
  uint64_t avail = n > start ? n - start : 0;
  uint64_t take = std::min(want, avail);
  *out_n = take;
  DevBuf rows(ctx, std::max<uint64_t>(take, 1) * 8);
  if (take == 0) return rows;
  uint64_t k = start + take;
  if (limit >= 0 && qmk_topk_cap((uint32_t)std::min<uint64_t>(k, 1u << 30)) && k < (1u << 30)) {
    DScan sc;
    memset(&sc, 0, sizeof sc);
    sc.nrows = n;
    sc.ncols = 1;
    sc.nqueries = 1;
    sc.cols[0] = c;
    const uint64_t tiles = (n + 2047) / 2048;
    int grid = (int)std::min<uint64_t>(tiles, (uint64_t)ctx->num_cus * (uint64_t)ctx->tune.topk_bpc);
    if (grid < 1) grid = 1;
    DevBuf cand(ctx, (size_t)grid * k * sizeof(TopKey));
    DevBuf counts(ctx, (size_t)grid * 4 + 64), gthr(ctx, 64);
    HIPX(qmk_fill_u64(gthr.as<uint64_t>(), 1, ~0ull, ctx->lane().stream));
    HIPX(qmk_topk_scan(&sc, d_noq, 0, desc ? 1 : 0, (uint32_t)k, grid, cand.as<TopKey>(), counts.as<uint32_t>(),
                       gthr.as<uint64_t>(), ctx->lane().stream));
    ++*launches;
    const TopKey *lists = cand.as<TopKey>();
    const uint32_t *lcounts = counts.as<uint32_t>();
    uint32_t nlists = (uint32_t)grid;
    DevBuf m1, m1c;
    if (nlists > 16) {
      uint32_t g2 = (nlists + 15) / 16;
      m1 = DevBuf(ctx, (size_t)g2 * k * sizeof(TopKey));
      m1c = DevBuf(ctx, (size_t)g2 * 4 + 64);
      HIPX(qmk_topk_merge(lists, lcounts, nlists, 16, (uint32_t)k, m1.as<TopKey>(), m1c.as<uint32_t>(), ctx->lane().stream));
      ++*launches;
      lists = m1.as<TopKey>();
      lcounts = m1c.as<uint32_t>();
      nlists = g2;
    }
    DevBuf fin(ctx, (size_t)k * sizeof(TopKey));
    DevBuf finc(ctx, 64);
    HIPX(qmk_topk_merge(lists, lcounts, nlists, nlists, (uint32_t)k, fin.as<TopKey>(), finc.as<uint32_t>(),
                        ctx->lane().stream));
    HIPX(qmk_topk_rows(fin.as<TopKey>(), finc.as<uint32_t>(), (uint32_t)start, rows.as<uint64_t>(), (uint32_t)take,
                       ctx->lane().stream));
    *launches += 2;
    return rows;
  }
  // full stable sort, then slice
  size_t tb = qmk_sort_tmp_bytes(n);
  const uint64_t per_pass = sort_pass_rows(ctx, n, tb);
  if (per_pass) {
    sort_in_passes(ctx, c, n, desc, start, take, per_pass, rows, launches);
    return rows;
  }
  DevBuf tmp(ctx, tb);
  DevBuf all(ctx, n * 8);
  HIPX(qmk_sort_keys(&c, n, desc ? 1 : 0, all.as<uint64_t>(), tmp.p, tmp.n, ctx->lane().stream));
  *launches += 4;
  HIPX(hipMemcpyAsync(rows.p, all.as<uint64_t>() + start, take * 8, hipMemcpyDeviceToDevice, ctx->lane().stream));
  return rows;
}

int find_out(const qm_result *r, const std::string &name) {
  for (size_t i = 0; i < r->cols.size(); ++i) if (r->cols[i].name == name) return (int)i;
  return -1;
}

// ---- SQL post-stages over a materialised result (k_post.hip) ----

// Ascending ids of the rows whose keep byte is set.
DevBuf select_kept(qm_ctx *ctx, const uint8_t *keep, uint64_t n, uint64_t *n_out) {
  const uint64_t nb = std::max<uint64_t>(1, qmk_mask_blocks(n));
  DevBuf counts(ctx, nb * 8), offs(ctx, nb * 8), stmp(ctx, qmk_exscan_tmp_words(nb) * 8 + 64);
  uint64_t total = 0;
  if (n) {
    HIPX(qmk_mask_count(keep, n, counts.as<uint64_t>(), ctx->lane().stream));
    uint64_t *total_d = stmp.as<uint64_t>() + qmk_exscan_tmp_words(nb);
    HIPX(qmk_exscan_u64(counts.as<uint64_t>(), nb, offs.as<uint64_t>(), total_d, stmp.as<uint64_t>(),
                        ctx->lane().stream));
    d2h(ctx, &total, total_d, 1);
  }
  DevBuf ids(ctx, std::max<uint64_t>(total, 1) * 8);
  if (total) HIPX(qmk_mask_write(keep, n, offs.as<uint64_t>(), ids.as<uint64_t>(), ctx->lane().stream));
  *n_out = total;
  return ids;
}

// HAVING (src/sql/executor.ts:108-112): the rows whose program value is true,
// in their order.
void having_stage(qm_result *res, const Plan &pl) {
  qm_ctx *ctx = res->ctx;
  HProg prog;
  memset(&prog, 0, sizeof prog);
  if (pl.having.size() > (size_t)HP_MAX_OPS) fail(QM_ERR_UNSUPPORTED, "HAVING has more than %d terms", HP_MAX_OPS);
  std::vector<int> slot_of(res->cols.size(), -1);
  auto col = [&](const std::string &name) -> uint8_t {
    const int ci = find_out(res, name);
    if (ci < 0) fail(QM_ERR_COLUMN_NOT_FOUND, "HAVING column not found: %s", name.c_str());
    if (slot_of[ci] < 0) {
      if (prog.ncols >= HP_MAX_COLS) fail(QM_ERR_UNSUPPORTED, "HAVING references more than %d columns", HP_MAX_COLS);
      slot_of[ci] = prog.ncols;
      prog.cols[prog.ncols++] = dcol_of(res->cols[ci]);
    }
    return (uint8_t)slot_of[ci];
  };
  auto lit = [](const Lit &l, uint8_t &kind, double &f, int64_t &i) {
    kind = l.kind == LIT_I64 ? HL_I64 : (l.kind == LIT_F64 ? HL_F64 : HL_NULL);
    f = l.f;
    i = l.i;
  };
  std::vector<uint8_t> lk;
  std::vector<double> lf;
  std::vector<int64_t> li;
  int depth = 0, max_depth = 0;
  for (const HavingOp &h : pl.having) {
    HPOp &o = prog.ops[prog.nops++];
    o.kind = (uint8_t)h.kind;  // HK_* and HP_* share their values
    o.cmp = (uint8_t)h.cmp;
    o.neg = h.negated ? 1 : 0;
    switch (h.kind) {
      case HK_CMP: o.col = col(h.col); lit(h.a, o.lk, o.f, o.i); ++depth; break;
      case HK_CMP_COL: o.col = col(h.col); o.col2 = col(h.col2); ++depth; break;
      case HK_BETWEEN: o.col = col(h.col); lit(h.a, o.lk, o.f, o.i); lit(h.b, o.lk2, o.f2, o.i2); ++depth; break;
      case HK_IN:
        o.col = col(h.col);
        o.list_off = (uint32_t)lk.size();
        o.list_n = (uint32_t)h.list.size();
        for (const Lit &l : h.list) {
          uint8_t k;
          double f;
          int64_t i;
          lit(l, k, f, i);
          lk.push_back(k);
          lf.push_back(f);
          li.push_back(i);
        }
        ++depth;
        break;
      case HK_ISNULL: o.col = col(h.col); ++depth; break;
      case HK_AND: case HK_OR: --depth; break;
      default: break;
    }
    max_depth = std::max(max_depth, depth);
  }
  if (max_depth > HP_MAX_DEPTH) fail(QM_ERR_UNSUPPORTED, "HAVING nests deeper than %d", HP_MAX_DEPTH);
  const size_t nl = std::max<size_t>(lk.size(), 1);
  DevBuf dk(ctx, nl), df(ctx, nl * 8), di(ctx, nl * 8);
  if (!lk.empty()) {
    h2d(ctx, dk.p, lk.data(), lk.size());
    h2d(ctx, df.p, lf.data(), lf.size());
    h2d(ctx, di.p, li.data(), li.size());
  }
  prog.list_kind = dk.as<uint8_t>();
  prog.list_f = df.as<double>();
  prog.list_i = di.as<int64_t>();
  const uint64_t n = res->rows;
  DevBuf keep(ctx, std::max<uint64_t>(n, 1));
  if (n) HIPX(qmk_having(&prog, n, keep.as<uint8_t>(), ctx->lane().stream));
  uint64_t kept = 0;
  DevBuf ids = select_kept(ctx, keep.as<uint8_t>(), n, &kept);
  permute_result(res, ids.as<uint64_t>(), kept);
}

// Stable LSD passes that order `perm` (the identity on entry) by the keys:
// per key, from the last to the first, its value code and then its NULL flag.
void stable_order(qm_ctx *ctx, const std::vector<std::pair<DCol, bool>> &keys, uint64_t n, DevBuf &perm) {
  if (n < 2 || keys.empty()) return;
  DevBuf ka(ctx, n * 8), kb(ctx, n * 8), pb(ctx, n * 8);
  const size_t tb = std::max<size_t>(qmk_radix_pairs_tmp_bytes(n), 256);
  DevBuf tmp(ctx, tb);
  for (size_t j = keys.size(); j-- > 0;) {
    const DCol &c = keys[j].first;
    for (int pass = 0; pass < (c.nulls ? 2 : 1); ++pass) {
      HIPX(qmk_order_codes(&c, perm.as<uint64_t>(), n, keys[j].second ? 1 : 0, pass, ka.as<uint64_t>(),
                           ctx->lane().stream));
      HIPX(qmk_radix_pairs(ka.as<uint64_t>(), kb.as<uint64_t>(), perm.as<uint64_t>(), pb.as<uint64_t>(), n,
                           pass ? 1 : 64, tmp.p, tmp.n, ctx->lane().stream));
      std::swap(perm, pb);
    }
  }
}

// Multi-column ORDER BY (src/sql/executor.ts:114-141: stable, NULLs last in
// either direction, NaN above every number), then OFFSET / LIMIT.
void order_by_stage(qm_result *res, const std::vector<OrderKey> &order, int64_t limit, int64_t offset) {
  qm_ctx *ctx = res->ctx;
  std::vector<std::pair<DCol, bool>> keys;
  for (const OrderKey &k : order) {
    const int ci = find_out(res, k.col);
    if (ci < 0) fail(QM_ERR_COLUMN_NOT_FOUND, "ORDER BY column not found: %s", k.col.c_str());
    keys.push_back({dcol_of(res->cols[ci]), k.desc});
  }
  const uint64_t n = res->rows;
  DevBuf perm(ctx, std::max<uint64_t>(n, 1) * 8);
  HIPX(qmk_iota(perm.as<uint64_t>(), n, ctx->lane().stream));
  stable_order(ctx, keys, n, perm);
  const uint64_t start = std::min<uint64_t>((uint64_t)offset, n);
  uint64_t take = n - start;
  if (limit >= 0) take = std::min<uint64_t>(take, (uint64_t)limit);
  permute_result(res, perm.as<uint64_t>() + start, take);
}

// DISTINCT (src/operators.ts:1188-1227): the first row of every distinct key
// tuple, in row order; NULLs are one value, every NaN one value, -0 equals 0
// (the reference keys rows by String(value)). `names` empty: every column.
void distinct_stage(qm_result *res, const std::vector<std::string> &names) {
  qm_ctx *ctx = res->ctx;
  std::vector<std::pair<DCol, bool>> keys;
  std::vector<DCol> cols;
  if (names.empty()) {
    for (const OutCol &o : res->cols) cols.push_back(dcol_of(o));
  } else {
    for (const std::string &nm : names) {
      const int ci = find_out(res, nm);
      if (ci < 0) fail(QM_ERR_COLUMN_NOT_FOUND, "DISTINCT column not found: %s", nm.c_str());
      cols.push_back(dcol_of(res->cols[ci]));
    }
  }
  for (const DCol &c : cols) keys.push_back({c, false});
  const uint64_t n = res->rows;
  if (n < 2 || cols.empty()) return;
  DevBuf perm(ctx, n * 8);
  HIPX(qmk_iota(perm.as<uint64_t>(), n, ctx->lane().stream));
  stable_order(ctx, keys, n, perm);  // equal tuples side by side, each run in row order
  DevBuf dcols(ctx, cols.size() * sizeof(DCol));
  h2d(ctx, dcols.p, reinterpret_cast<const uint8_t *>(cols.data()), cols.size() * sizeof(DCol));
  DevBuf keep(ctx, n);
  HIPX(qmk_distinct_marks(dcols.as<DCol>(), (int)cols.size(), perm.as<uint64_t>(), n, keep.as<uint8_t>(),
                          ctx->lane().stream));
  uint64_t m = 0;
  DevBuf pos = select_kept(ctx, keep.as<uint8_t>(), n, &m);
  DevBuf first(ctx, std::max<uint64_t>(m, 1) * 8), sorted(ctx, std::max<uint64_t>(m, 1) * 8),
      sv(ctx, std::max<uint64_t>(m, 1) * 8);
  HIPX(qmk_gather_u64(perm.as<uint64_t>(), pos.as<uint64_t>(), m, first.as<uint64_t>(), ctx->lane().stream));
  if (m > 1) {  // back to row order
    DevBuf tmp(ctx, std::max<size_t>(qmk_radix_pairs_tmp_bytes(m), 256));
    HIPX(qmk_radix_pairs(first.as<uint64_t>(), sorted.as<uint64_t>(), first.as<uint64_t>(), sv.as<uint64_t>(), m, 64,
                         tmp.p, tmp.n, ctx->lane().stream));
    permute_result(res, sorted.as<uint64_t>(), m);
  } else {
    permute_result(res, first.as<uint64_t>(), m);
  }
}

// The output columns of a plan whose projections name a subset of the
// result (HAVING-only aggregates, hidden sort / DISTINCT columns): kept in
// projection order, names not in the result skipped (src/sql/executor.ts:146-160).
void project_result(qm_result *res, const std::vector<std::string> &proj) {
  if (proj.empty()) return;
  std::vector<OutCol> keep;
  for (const std::string &nm : proj) {
    const int ci = find_out(res, nm);
    if (ci >= 0 && !res->cols[ci].name.empty()) {
      keep.push_back(std::move(res->cols[ci]));
      res->cols[ci].name.clear();  // a name listed twice is taken once
    }
  }
  res->cols = std::move(keep);
}

// ORDER BY / LIMIT / OFFSET over a materialised result
void sort_limit_result(qm_result *res, const Plan &pl, const DQuery *d_noq, uint32_t *launches) {
  qm_ctx *ctx = res->ctx;
  if (!pl.having.empty()) {
    having_stage(res, pl);
    *launches += 4;
  }
  if (!pl.order_by.empty()) {
    order_by_stage(res, pl.order_by, pl.limit, pl.offset);
    *launches += 4 * (uint32_t)pl.order_by.size();
    if (!pl.having.empty() && !pl.aggs.empty()) project_result(res, pl.proj);
    return;
  }
  if (!pl.having.empty() && !pl.aggs.empty()) {
    Plan rest = pl;  // the sort / limit below, then the projection
    rest.having.clear();
    sort_limit_result(res, rest, d_noq, launches);
    project_result(res, pl.proj);
    return;
  }
  if (pl.has_sort) {
    int ci = find_out(res, pl.sort_col);
    if (ci < 0) fail(QM_ERR_COLUMN_NOT_FOUND, "sort column not found: %s", pl.sort_col.c_str());
    uint64_t n_out = 0;
    DCol c = dcol_of(res->cols[ci]);
    DevBuf rows = order_rows(ctx, c, res->rows, pl.desc, pl.limit, pl.offset, &n_out, d_noq, launches);
    permute_result(res, rows.as<uint64_t>(), n_out);
    return;
  }
  if (pl.limit >= 0 || pl.offset > 0) {
    uint64_t start = std::min<uint64_t>((uint64_t)pl.offset, res->rows);
    uint64_t take = res->rows - start;
    if (pl.limit >= 0) take = std::min<uint64_t>(take, (uint64_t)pl.limit);
    std::vector<uint64_t> idx(take);
    for (uint64_t i = 0; i < take; ++i) idx[i] = start + i;
    DevBuf rows(ctx, std::max<uint64_t>(take, 1) * 8);
    h2d(ctx, rows.p, idx.data(), take);
    permute_result(res, rows.as<uint64_t>(), take);
  }
}

}  // namespace

// ===========================================================================
// join index
// ===========================================================================
namespace {
struct JoinIndex {
  DevBuf table, rows, cursor, tmp;
  DevBuf unique;     // DIRECT with unique keys: build row + 1 per slot (empty otherwise)
  int mode = 0;      // 1 DIRECT, 2 HASH
  uint64_t cap = 0;
  int64_t kmin = 0;
  int key_float = 0;
  uint64_t build_rows = 0;
};

DevBuf select_rows(qm_ctx *ctx, const DScan &sc, const DQuery *d_q, uint64_t limit_cap, uint64_t *n_out,
                   uint32_t *launches);

// fused: the index feeds a fused aggregate, which needs only the unique-key
// DIRECT index when the build keys are unique.
// One key range of a partitioned join: the index holds the build rows with
// lo <= key <= hi only (every probe row is looked up against it; keys outside
// the range miss), so one pass per range covers the join once.
struct KeyRange {
  int64_t lo, hi;
};

void build_join(qm_ctx *ctx, ColSpace &cs, const Plan &pl_in, JoinIndex &ji, uint32_t *launches, DevBuf *keep,
                bool fused, const KeyRange *part = nullptr) {
  Plan pl_part;
  if (part) {  // the range as one more build-side filter
    pl_part = pl_in;
    Filter f;
    f.col = pl_in.rkey;
    f.op = OP_BETWEEN;
    f.vkind = LIT_PAIR;
    Lit a, b;
    a.kind = b.kind = LIT_I64;
    a.i = part->lo;
    b.i = part->hi;
    f.list = {a, b};
    pl_part.rfilters.push_back(f);
  }
  const Plan &pl = part ? pl_part : pl_in;
  const qm_table *R = cs.R;
  int rk = R->find(pl.rkey);
  if (rk < 0) fail(QM_ERR_COLUMN_NOT_FOUND, "join key not found: %s", pl.rkey.c_str());
  int lk = cs.L->find(pl.lkey);
  if (lk < 0) fail(QM_ERR_COLUMN_NOT_FOUND, "join key not found: %s", pl.lkey.c_str());
  const TCol &kc = R->cols[rk];
  const TCol &lc = cs.L->cols[lk];
  if (dt_float(kc.dtype) != dt_float(lc.dtype))
    fail(QM_ERR_UNSUPPORTED, "join keys of mixed integer/float type");
  ji.key_float = dt_float(kc.dtype) ? 1 : 0;
  // right-side filters (src/local-executor.ts:631-635)
  const uint64_t *sel = nullptr;
  uint64_t nbuild = R->rows;
  if (!pl.rfilters.empty() || !pl.rgroups.empty()) {
    ColSpace rcs;
    rcs.L = R;
    rcs.plan = &pl;
    DQuery q;
    memset(&q, 0, sizeof q);
    ListStore ls;
    bind_filters(ctx, rcs, pl.rfilters, pl.rgroups, q, ls, false);
    DevBuf dq(ctx, sizeof(DQuery));
    h2d(ctx, dq.p, reinterpret_cast<const uint8_t *>(&q), sizeof q);
    DScan sc;
    memset(&sc, 0, sizeof sc);
    sc.nrows = R->rows;
    sc.ncols = (int)rcs.cols.size();
    sc.nqueries = 1;
    for (size_t i = 0; i < rcs.cols.size(); ++i) sc.cols[i] = rcs.cols[i];
    *keep = select_rows(ctx, sc, dq.as<DQuery>(), ~0ull, &nbuild, launches);
    sel = keep->as<uint64_t>();
    stream_sync(ctx);
  }
  ji.build_rows = nbuild;
  DCol key;
  key.data = kc.d;
  key.nulls = kc.dnull;
  key.dtype = kc.dtype;
  key.side = 1;
  uint64_t dom = 0;
  int64_t klo = kc.st.imin;
  if (!ji.key_float && kc.st.has_int) {
    int64_t khi = kc.st.imax;
    if (part) {
      klo = std::max(klo, part->lo);
      khi = std::min(khi, part->hi);
    }
    dom = khi >= klo ? (uint64_t)(khi - klo) + 1 : 1;
  }
  if (!ji.key_float && kc.st.has_int && dom != 0 && dom <= std::max<uint64_t>(4 * nbuild, 1u << 16) &&
      dom <= (1ull << 31)) {
    ji.mode = 1;
    ji.cap = dom;
    ji.kmin = klo;
  } else {
    ji.mode = 2;
    ji.cap = pow2ceil(std::max<uint64_t>(1024, 2 * nbuild));
  }
  if (fused && ji.mode == 1 && ctx->tune.join_unique && ctx->tune.join_claim) {
    ji.unique = DevBuf(ctx, ji.cap * 4 + 64);
    zero(ctx, ji.unique.p, ji.cap * 4 + 64);
    DevBuf status(ctx, 64);
    zero(ctx, status.p, 64);
    HIPX(qmk_join_claim(&key, nbuild, sel, ji.kmin, ji.cap, ji.unique.as<uint32_t>(), status.as<uint32_t>(),
                        ctx->lane().stream));
    *launches += 1;
    uint32_t st[2];
    d2h(ctx, st, status.p, 2);
    if (!st[1]) return;  // unique keys: the claimed slots are the whole index
    ji.unique.reset();   // duplicate build keys: the counted build below
  }
  uint64_t nent = (ji.mode == 1) ? ji.cap : ji.cap + 1;
  uint64_t twords = (ji.mode == 1) ? ji.cap : 2 * (ji.cap + 1);
  ji.table = DevBuf(ctx, twords * 8);
  zero(ctx, ji.table.p, twords * 8);
  ji.rows = DevBuf(ctx, std::max<uint64_t>(nbuild, 1) * 4);
  ji.cursor = DevBuf(ctx, nent * 4 + 64);
  zero(ctx, ji.cursor.p, nent * 4 + 64);
  ji.tmp = DevBuf(ctx, qmk_join_build_tmp_words(nent, nbuild) * 8);
  DevBuf status(ctx, 64);
  zero(ctx, status.p, 64);
  HIPX(qmk_join_build(&key, nbuild, sel, ji.mode, ji.key_float, ji.kmin, ji.cap, ji.table.as<uint64_t>(),
                      ji.rows.as<uint32_t>(), ji.cursor.as<uint32_t>(), ji.tmp.as<uint64_t>(), status.as<uint32_t>(),
                      ctx->lane().stream));
  *launches += 8;
  if (ji.mode == 1 && ctx->tune.join_unique) {
    ji.unique = DevBuf(ctx, ji.cap * 4 + 64);
    HIPX(qmk_join_unique(ji.table.as<uint64_t>(), ji.cap, ji.rows.as<uint32_t>(), ji.unique.as<uint32_t>(),
                         status.as<uint32_t>(), ctx->lane().stream));
    *launches += 1;
  }
  uint32_t st[2];
  d2h(ctx, st, status.p, 2);
  if (st[0]) fail(QM_ERR_QUERY_FAILED, "join hash table overflow");
  if (st[1]) ji.unique.reset();  // duplicate build keys: probe through the row lists
}

// Pack the build-side columns a fused probe reads (at most JOIN_PAY_BYTES)
// into 16 bytes per DIRECT key slot, widest first, with a flags word
// (DScan::join_pay). Leaves sc.join_pay null when they do not fit.
// Device bytes of the index build_join lays out for `nbuild` build rows whose
// keys span `dom` values (DIRECT when dense enough, else open addressing). A
// fused join over unique DIRECT keys claims one slot per key and packs the
// build columns beside it, and builds nothing else.
uint64_t join_index_bytes(const qm_ctx *ctx, uint64_t nbuild, uint64_t dom, bool key_float, bool fused,
                          bool unique_keys) {
  const bool direct = !key_float && dom != 0 && dom <= std::max<uint64_t>(4 * nbuild, 1u << 16) && dom <= (1ull << 31);
  if (direct && fused && unique_keys && ctx->tune.join_unique && ctx->tune.join_claim)
    return dom * 4 + (ctx->tune.join_payload ? dom * 16 : 0) + 128;
  const uint64_t cap = direct ? dom : pow2ceil(std::max<uint64_t>(1024, 2 * nbuild));
  const uint64_t nent = direct ? cap : cap + 1;
  const uint64_t twords = direct ? cap : 2 * (cap + 1);
  uint64_t b = twords * 8 + nbuild * 4 + nent * 4 + qmk_join_build_tmp_words(nent, nbuild) * 8;
  if (direct) b += cap * 4 + (fused ? cap * 16 : 0);  // unique-row slots, packed build columns
  return b;
}

// The key ranges of a join's passes (the reference partitions a build side
// that exceeds its memory budget and joins partition by partition,
// src/operators.ts:2466-2637; here the build side stays resident and the
// index is what must fit). Empty: one pass over the whole build side. Else
// the fewest equal ranges of the build key's span (a power of two) whose
// indexes each fit half the HBM budget left, or `join_parts` ranges when that
// knob is set. Integer keys only.
std::vector<KeyRange> join_ranges(qm_ctx *ctx, const qm_table *R, const Plan &pl, bool fused) {
  std::vector<KeyRange> out;
  const int rk = R->find(pl.rkey);
  if (rk < 0) return out;
  const TCol &kc = R->cols[rk];
  if (dt_float(kc.dtype) || !kc.st.has_int) return out;
  const uint64_t dom = (uint64_t)kc.st.imax - (uint64_t)kc.st.imin + 1;  // 0: the whole 2^64 span
  const bool uniq = kc.st.dup_frac == 0 && !kc.st.nulls && kc.st.distinct >= 0.9 * (double)kc.st.nonnull;
  uint64_t P = (uint64_t)std::max(0, ctx->tune.join_parts);
  if (P == 0) {
    const uint64_t need = join_index_bytes(ctx, R->rows, dom, false, fused, uniq);
    uint64_t used;
    {
      std::lock_guard<std::mutex> g(ctx->pool_mu);
      used = ctx->used.load() - ctx->pooled;
    }
    const uint64_t room = (ctx->budget > used ? ctx->budget - used : 0) / 2;
    if (need <= room) return out;
    P = 2;
    while (P < 4096 &&
           join_index_bytes(ctx, R->rows / P + 1, dom ? (dom + P - 1) / P : (~0ull / P + 1), false, fused, uniq) > room)
      P *= 2;
  }
  if (P <= 1) return out;
  const uint64_t width = dom ? (dom + P - 1) / P : (~0ull / P + 1);
  for (uint64_t p = 0; p < P; ++p) {
    const uint64_t off = p * width;
    if (dom && off >= dom) break;
    const uint64_t last = (dom && dom - off <= width) ? dom - 1 : off + width - 1;
    out.push_back({(int64_t)((uint64_t)kc.st.imin + off), (int64_t)((uint64_t)kc.st.imin + last)});
    if (!dom && p + 1 == P) break;
  }
  return out;
}

void attach_join_payload(qm_ctx *ctx, const ColSpace &cs, const JoinIndex &ji, DScan &sc, DevBuf &jpay,
                         uint32_t *launches) {
  std::vector<int> side1;
  int bytes = 0;
  for (size_t i = 0; i < cs.cols.size(); ++i)
    if (cs.cols[i].side == 1) {
      side1.push_back((int)i);
      bytes += (int)dt_size(cs.cols[i].dtype);
    }
  if (side1.empty() || bytes > JOIN_PAY_BYTES) return;
  std::stable_sort(side1.begin(), side1.end(),
                   [&](int a, int b) { return dt_size(cs.cols[a].dtype) > dt_size(cs.cols[b].dtype); });
  DJoinPay jp;
  memset(&jp, 0, sizeof jp);
  int off = 0;
  for (int ci : side1) {
    const int j = jp.n++;
    jp.cols[j] = cs.cols[ci];
    jp.off[j] = (int8_t)off;
    jp.nbit[j] = cs.cols[ci].nulls ? (int8_t)(j + 1) : 0;
    sc.pay_off[ci] = jp.off[j];
    sc.pay_null[ci] = jp.nbit[j];
    off += (int)dt_size(cs.cols[ci].dtype);
  }
  if (!jpay.p) {
    jpay = DevBuf(ctx, ji.cap * 16);
    HIPX(qmk_join_payload(ji.unique.as<uint32_t>(), ji.cap, &jp, jpay.as<uint32_t>(), ctx->lane().stream));
    *launches += 1;
  }
  sc.join_pay = jpay.as<uint32_t>();
}

void attach_join(DScan &sc, const JoinIndex &ji, int probe_slot) {
  sc.join_table = ji.table.as<uint64_t>();
  sc.join_rows = ji.rows.as<uint32_t>();
  sc.join_cap = ji.cap;
  sc.join_min = ji.kmin;
  sc.join_col = probe_slot;
  sc.join_mode = ji.mode;
  sc.join_key_float = ji.key_float;
}

// stream compaction: ascending row ids of rows passing the query
DevBuf select_rows(qm_ctx *ctx, const DScan &sc, const DQuery *d_q, uint64_t limit_cap, uint64_t *n_out,
                   uint32_t *launches) {
  const uint64_t tile_rows = 256 * RPT;
  const uint64_t ntiles = std::max<uint64_t>(1, (sc.nrows + tile_rows - 1) / tile_rows);
  uint64_t nblocks = std::min<uint64_t>(ntiles, (uint64_t)ctx->num_cus * 8);
  uint64_t tpb = (ntiles + nblocks - 1) / nblocks;
  nblocks = (ntiles + tpb - 1) / tpb;
  DevBuf counts(ctx, nblocks * 8), offs(ctx, nblocks * 8), stmp(ctx, qmk_exscan_tmp_words(nblocks) * 8 + 64);
  HIPX(qmk_select_count(&sc, d_q, nblocks, tpb, counts.as<uint64_t>(), ctx->lane().stream));
  uint64_t *total_d = stmp.as<uint64_t>() + qmk_exscan_tmp_words(nblocks);
  HIPX(qmk_exscan_u64(counts.as<uint64_t>(), nblocks, offs.as<uint64_t>(), total_d, stmp.as<uint64_t>(), ctx->lane().stream));
  uint64_t total = 0;
  d2h(ctx, &total, total_d, 1);
  uint64_t n = std::min(total, limit_cap);
  DevBuf rows(ctx, std::max<uint64_t>(n, 1) * 8 + 64);
  if (n) HIPX(qmk_select_write(&sc, d_q, nblocks, tpb, offs.as<uint64_t>(), n, rows.as<uint64_t>(), ctx->lane().stream));
  *launches += 5;
  *n_out = n;
  return rows;
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
namespace {
// The planner's tuning knobs: name, environment override read at qm_init,
// range (values are clamped), setter and getter.
struct KnobDef {
  const char *name, *env;
  double lo, hi;
  void (*set)(qm_ctx::Tune &, double);
  double (*get)(const qm_ctx::Tune &);
};
const KnobDef kKnobs[] = {
    {"scatter_rpb", "QM_TUNE_SCATTER_RPB", 0, 64, [](qm_ctx::Tune &t, double v) { t.scatter_rpb = (int)v; },
     [](const qm_ctx::Tune &t) { return (double)t.scatter_rpb; }},
    {"part_load", "QM_TUNE_PART_LOAD", 0.1, 0.9, [](qm_ctx::Tune &t, double v) { t.part_load = v; },
     [](const qm_ctx::Tune &t) { return t.part_load; }},
    {"hot", "QM_TUNE_HOT", 0, 1, [](qm_ctx::Tune &t, double v) { t.hot = v != 0; },
     [](const qm_ctx::Tune &t) { return (double)t.hot; }},
    {"qp_min", "QM_TUNE_QP_MIN", 1, 1 << 20, [](qm_ctx::Tune &t, double v) { t.qp_min = (int)v; },
     [](const qm_ctx::Tune &t) { return (double)t.qp_min; }},
    {"runs_min_q", "QM_TUNE_RUNS_MIN_Q", 1, 1 << 20, [](qm_ctx::Tune &t, double v) { t.runs_min_q = (int)v; },
     [](const qm_ctx::Tune &t) { return (double)t.runs_min_q; }},
    {"runs_min_len", "QM_TUNE_RUNS_MIN_LEN", 0, 1e12, [](qm_ctx::Tune &t, double v) { t.runs_min_len = v; },
     [](const qm_ctx::Tune &t) { return t.runs_min_len; }},
    {"runs_chunks_per_cu", "QM_TUNE_RUNS_CHUNKS_PER_CU", 1, 1024,
     [](qm_ctx::Tune &t, double v) { t.runs_chunks_per_cu = (int)v; },
     [](const qm_ctx::Tune &t) { return (double)t.runs_chunks_per_cu; }},
    {"runs_win", "QM_TUNE_RUNS_WIN", 1024, 2048, [](qm_ctx::Tune &t, double v) { t.runs_win = v >= 2048 ? 2048 : 1024; },
     [](const qm_ctx::Tune &t) { return (double)t.runs_win; }},
    {"runs_qblock", "QM_TUNE_RUNS_QBLOCK", 0, 64,
     [](qm_ctx::Tune &t, double v) { t.runs_qblock = v >= 48 ? 64 : (v >= 24 ? 32 : (v >= 8 ? 16 : 0)); },
     [](const qm_ctx::Tune &t) { return (double)t.runs_qblock; }},
    {"shared_lean", "QM_TUNE_SHARED_LEAN", 0, 1, [](qm_ctx::Tune &t, double v) { t.shared_lean = v != 0; },
     [](const qm_ctx::Tune &t) { return (double)t.shared_lean; }},
    {"shared_waves", "QM_TUNE_SHARED_WAVES", 1, 64, [](qm_ctx::Tune &t, double v) { t.shared_waves = (int)v; },
     [](const qm_ctx::Tune &t) { return (double)t.shared_waves; }},
    {"sorted1", "QM_TUNE_SORTED1", 0, 1, [](qm_ctx::Tune &t, double v) { t.sorted1 = v != 0; },
     [](const qm_ctx::Tune &t) { return (double)t.sorted1; }},
    {"join_unique", "QM_TUNE_JOIN_UNIQUE", 0, 1, [](qm_ctx::Tune &t, double v) { t.join_unique = v != 0; },
     [](const qm_ctx::Tune &t) { return (double)t.join_unique; }},
    {"join_parts", "QM_TUNE_JOIN_PARTS", 0, 4096, [](qm_ctx::Tune &t, double v) { t.join_parts = (int)v; },
     [](const qm_ctx::Tune &t) { return (double)t.join_parts; }},
    {"sort_parts", "QM_TUNE_SORT_PARTS", 0, 4096, [](qm_ctx::Tune &t, double v) { t.sort_parts = (int)v; },
     [](const qm_ctx::Tune &t) { return (double)t.sort_parts; }},
    {"join_claim", "QM_TUNE_JOIN_CLAIM", 0, 1, [](qm_ctx::Tune &t, double v) { t.join_claim = v != 0; },
     [](const qm_ctx::Tune &t) { return (double)t.join_claim; }},
    {"join_payload", "QM_TUNE_JOIN_PAYLOAD", 0, 1, [](qm_ctx::Tune &t, double v) { t.join_payload = v != 0; },
     [](const qm_ctx::Tune &t) { return (double)t.join_payload; }},
    {"zone_maps", "QM_TUNE_ZONE_MAPS", 0, 1, [](qm_ctx::Tune &t, double v) { t.zone_maps = v != 0; },
     [](const qm_ctx::Tune &t) { return (double)t.zone_maps; }},
    {"pr_max_s", "QM_TUNE_PR_MAX_S", 1, PR_MAX_S, [](qm_ctx::Tune &t, double v) { t.pr_max_s = (int)v; },
     [](const qm_ctx::Tune &t) { return (double)t.pr_max_s; }},
    {"pr_b2", "QM_TUNE_PR_B2", 0, 10, [](qm_ctx::Tune &t, double v) { t.pr_b2 = (int)v; },
     [](const qm_ctx::Tune &t) { return (double)t.pr_b2; }},
    {"pr_scatter_bpc", "QM_TUNE_PR_SCATTER_BPC", 1, 8, [](qm_ctx::Tune &t, double v) { t.pr_scatter_bpc = (int)v; },
     [](const qm_ctx::Tune &t) { return (double)t.pr_scatter_bpc; }},
    {"topk_bpc", "QM_TUNE_TOPK_BPC", 1, 16, [](qm_ctx::Tune &t, double v) { t.topk_bpc = (int)v; },
     [](const qm_ctx::Tune &t) { return (double)t.topk_bpc; }},
    {"runs_lds_pad", "QM_TUNE_RUNS_LDS_PAD", 0, 96 * 1024, [](qm_ctx::Tune &t, double v) { t.runs_lds_pad = (int)v; },
     [](const qm_ctx::Tune &t) { return (double)t.runs_lds_pad; }},
};
const KnobDef *find_knob(const char *name) {
  if (!name) return nullptr;
  for (const KnobDef &k : kKnobs)
    if (!strcmp(k.name, name)) return &k;
  return nullptr;
}
}  // namespace

// Scratch of qm_exec_noalloc: group table, descriptor, counters and
// compaction temporaries of one call, in a file-scope device array (one copy
// per device, held until the process ends). Host code never uses the array's
// name as an address: qm_init records this device's copy's address
// (hipGetSymbolAddress) in the context.
static constexpr uint64_t QM_NOALLOC_SCRATCH_BYTES = 32ull << 20;
static __device__ __align__(256) uint8_t scratch[QM_NOALLOC_SCRATCH_BYTES];

extern "C" int qm_init(int device, uint64_t hbm_budget_bytes, qm_ctx **out) {
  API_BEGIN
  if (!out) fail(QM_ERR_INVALID_ARG, "null out");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    fail(QM_ERR_QUERY_FAILED, "no HIP device available");
  }
  if (device < 0 || device >= n) fail(QM_ERR_INVALID_ARG, "device %d out of range (%d devices)", device, n);
  HIPX(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIPX(hipGetDeviceProperties(&prop, device));
  std::unique_ptr<qm_ctx> c(new qm_ctx());
  c->device = device;
  size_t fr = 0, tot = 0;
  HIPX(hipMemGetInfo(&fr, &tot));
  c->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)(tot * 0.9);
  c->num_cus = prop.multiProcessorCount;
  c->name = prop.gcnArchName;
  for (const KnobDef &k : kKnobs)
    if (const char *e = getenv(k.env)) k.set(c->tune, std::min(k.hi, std::max(k.lo, atof(e))));
  for (qm_lane &l : c->lanes) {  // every stream at the default priority
    HIPX(hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking));
    HIPX(hipEventCreate(&l.ev0));
    HIPX(hipEventCreate(&l.ev1));
    HIPX(hipEventCreate(&l.evk0));
    HIPX(hipEventCreate(&l.evk1));
    HIPX(hipEventCreate(&l.evk2));
  }
  {
    void *sp = nullptr;
    const hipError_t e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !sp) fail(QM_ERR_QUERY_FAILED, "scratch array address: %s", hipGetErrorString(e));
    c->scratch = sp;
  }
  *out = c.release();
  API_END
}

extern "C" int qm_shutdown(qm_ctx *ctx) {
  API_BEGIN
  if (!ctx) return QM_OK;
  (void)hipSetDevice(ctx->device);
  {
    std::unique_lock<std::shared_mutex> wr(ctx->tables_mu);  // no call runs past this point
    for (qm_lane &l : ctx->lanes) (void)hipStreamSynchronize(l.stream);
    ctx->trim_pool();
  }
  for (qm_lane &l : ctx->lanes) {
    (void)hipEventDestroy(l.ev0);
    (void)hipEventDestroy(l.ev1);
    (void)hipEventDestroy(l.evk0);
    (void)hipEventDestroy(l.evk1);
    (void)hipEventDestroy(l.evk2);
    (void)hipStreamDestroy(l.stream);
  }
  delete ctx;
  API_END
}

extern "C" int qm_last_error(qm_ctx *ctx, char *buf, size_t cap) {
  (void)ctx;
  if (buf && cap) {
    size_t n = std::min(cap - 1, g_last_error.size());
    memcpy(buf, g_last_error.data(), n);
    buf[n] = 0;
  }
  return (int)g_last_error.size();
}

extern "C" int qm_device_info(qm_ctx *ctx, char *name, size_t cap, uint32_t *num_cus, uint64_t *budget,
                              uint64_t *used) {
  API_BEGIN
  if (!ctx) fail(QM_ERR_INVALID_ARG, "null context");
  if (name && cap) {
    size_t n = std::min(cap - 1, ctx->name.size());
    memcpy(name, ctx->name.data(), n);
    name[n] = 0;
  }
  if (num_cus) *num_cus = (uint32_t)ctx->num_cus;
  if (budget) *budget = ctx->budget;
  if (used) *used = ctx->used.load() - ctx->pooled;
  API_END
}

extern "C" int qm_table_create(qm_ctx *ctx, const char *name, uint64_t nrows, uint32_t ncols,
                               const qm_column_desc *cols, qm_table **out) {
  API_BEGIN
  if (!ctx || !out || (ncols && !cols)) fail(QM_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (nrows >= (1ull << 40)) fail(QM_ERR_INVALID_ARG, "table too large");
  LaneHold hold(ctx, true);
  HIPX(hipSetDevice(ctx->device));
  std::unique_ptr<qm_table> t(new qm_table());
  t->ctx = ctx;
  t->name = name ? name : "";
  t->rows = nrows;
  struct Cleanup {
    qm_table *t;
    bool armed = true;
    ~Cleanup() {
      if (!armed) return;
      for (auto &c : t->cols) {
        if (c.d) t->ctx->release(c.d, c.bytes);
        if (c.dnull) t->ctx->release(c.dnull, c.null_bytes);
      }
    }
  } cleanup{t.get()};
  const uint64_t pr = pad_rows(nrows);
  DevBuf stats(ctx, 8 * 8), hll(ctx, 4096 * 4);
  for (uint32_t i = 0; i < ncols; ++i) {
    const qm_column_desc &cd = cols[i];
    int dt = cd.dtype;
    if (dt != QM_F64 && dt != QM_I64 && dt != QM_I32 && dt != QM_F32 && dt != QM_I8 && dt != QM_I16 && dt != QM_BOOL)
      fail(QM_ERR_UNSUPPORTED, "column %s: dtype %d is not on the device path", cd.name ? cd.name : "?", dt);
    if (!cd.name) fail(QM_ERR_INVALID_ARG, "column %u has no name", i);
    if (nrows && !cd.data) fail(QM_ERR_INVALID_ARG, "column %s has no data", cd.name);
    for (auto &c : t->cols) if (c.name == cd.name) fail(QM_ERR_SCHEMA_MISMATCH, "duplicate column %s", cd.name);
    TCol c;
    c.name = cd.name;
    c.dtype = dt;
    const int es = dt_size(dt);
    c.bytes = pool_round(pr * es);
    c.d = ctx->alloc(c.bytes);
    t->cols.push_back(c);
    TCol &tc = t->cols.back();
    zero(ctx, tc.d, pr * es);
    if (nrows) HIPX(hipMemcpyAsync(tc.d, cd.data, nrows * es, hipMemcpyHostToDevice, ctx->lane().stream));
    bool any_null = false;
    if (cd.null_bitmap) {
      for (uint64_t b = 0; b < (nrows + 7) / 8 && !any_null; ++b) {
        uint8_t m = cd.null_bitmap[b];
        if (b == nrows / 8 && (nrows & 7)) m &= (uint8_t)((1u << (nrows & 7)) - 1);
        if (m) any_null = true;
      }
    }
    if (any_null) {
      tc.null_bytes = pool_round(pr / 8 + 64);
      tc.dnull = static_cast<uint8_t *>(ctx->alloc(tc.null_bytes));
      zero(ctx, tc.dnull, pr / 8 + 64);
      std::vector<uint8_t> nb(cd.null_bitmap, cd.null_bitmap + (nrows + 7) / 8);
      if (nrows & 7) nb.back() &= (uint8_t)((1u << (nrows & 7)) - 1);
      HIPX(hipMemcpyAsync(tc.dnull, nb.data(), nb.size(), hipMemcpyHostToDevice, ctx->lane().stream));
      stream_sync(ctx);
    }
    // statistics (min/max, max |x|, NaN/NULL counts, HLL)
    zero(ctx, stats.p, 64);
    zero(ctx, hll.p, 4096 * 4);
    DCol dc;
    dc.data = tc.d;
    dc.nulls = tc.dnull;
    dc.dtype = dt;
    dc.side = 0;
    HIPX(qmk_col_stats(&dc, nrows, stats.as<uint64_t>(), hll.as<uint32_t>(), ctx->lane().stream));
    uint64_t sw[8];
    std::vector<uint32_t> regs(4096);
    d2h(ctx, sw, stats.p, 8);
    d2h(ctx, regs.data(), hll.p, 4096);
    if (nrows) {
      const uint64_t pages = (nrows + ZONE_ROWS - 1) / ZONE_ROWS;
      DevBuf zb(ctx, pages * 16 + 64);
      HIPX(qmk_zone_maps(&dc, nrows, zb.as<uint64_t>(), zb.as<uint64_t>() + pages, ctx->lane().stream));
      tc.zmin.resize(pages);
      tc.zmax.resize(pages);
      d2h(ctx, tc.zmin.data(), zb.p, pages);
      d2h(ctx, tc.zmax.data(), zb.as<uint64_t>() + pages, pages);
    }
    ColStats &s = tc.st;
    s.nan = sw[3];
    s.nulls = sw[4];
    s.nonnull = sw[5];
    if (dt_float(dt)) {
      if (sw[1]) {
        s.has_f = true;
        s.fmin = bits_f64(decode_f64_code(~sw[0]));
        s.fmax = bits_f64(decode_f64_code(sw[1]));
      }
      s.max_abs = bits_f64(sw[2]);
      s.min_abs = sw[7] ? bits_f64(~sw[7]) : 0.0;
    } else if (s.nonnull) {
      s.has_int = true;
      s.imin = decode_i64(~sw[0]);
      s.imax = decode_i64(sw[1]);
      s.max_abs = std::max(std::fabs((double)s.imin), std::fabs((double)s.imax));
      s.sorted = s.nulls == 0 && sw[6] == 0;
    }
    s.distinct = s.nonnull ? std::max(1.0, hll_estimate(regs.data())) : 0.0;
    // key histogram: at most 2^20 bins of 2^g values
    if (s.has_int && s.nulls == 0 && (uint64_t)s.imax - (uint64_t)s.imin < (1ull << 32)) {
      const uint64_t D = (uint64_t)s.imax - (uint64_t)s.imin + 1;
      int kg = 0;
      while (((D + (1ull << kg) - 1) >> kg) > (1ull << 20)) ++kg;
      const uint64_t nbins = (D + (1ull << kg) - 1) >> kg;
      DevBuf hb(ctx, nbins * 4 + 64);
      zero(ctx, hb.p, nbins * 4);
      HIPX(qmk_key_hist(&dc, nrows, s.imin, kg, (uint32_t)nbins, hb.as<uint32_t>(), ctx->lane().stream));
      tc.khist.resize(nbins);
      d2h(ctx, tc.khist.data(), hb.p, nbins);
      tc.khist_g = kg;
    }
    // skew probe on a strided host sample
    if (nrows >= 2) {
      const uint64_t ns = std::min<uint64_t>(nrows, 16384);
      std::vector<uint64_t> smp(ns);
      const uint8_t *src = static_cast<const uint8_t *>(cd.data);
      for (uint64_t k = 0; k < ns; ++k) {
        uint64_t r = (ns == nrows) ? k : (k * (nrows / ns));
        uint64_t v = 0;
        memcpy(&v, src + r * es, es);
        smp[k] = v;
      }
      std::sort(smp.begin(), smp.end());
      uint64_t dup = 0;
      for (uint64_t k = 0; k < ns; ++k) {
        bool same = (k > 0 && smp[k] == smp[k - 1]) || (k + 1 < ns && smp[k] == smp[k + 1]);
        if (same) ++dup;
      }
      s.dup_frac = (double)dup / (double)ns;
      if (!dt_float(dt) && ns >= 4096) {
        std::vector<std::pair<uint64_t, int64_t>> runs;  // (count, value)
        for (uint64_t k = 0; k < ns;) {
          uint64_t e = k;
          while (e < ns && smp[e] == smp[k]) ++e;
          if (e - k >= ns / 2048) {
            int64_t v;
            // sign-extend narrow integers
            if (es == 1) v = (int8_t)smp[k];
            else if (es == 2) v = (int16_t)smp[k];
            else if (es == 4) v = (int32_t)smp[k];
            else v = (int64_t)smp[k];
            runs.push_back({e - k, v});
          }
          k = e;
        }
        std::sort(runs.begin(), runs.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
        for (size_t k = 0; k < runs.size() && k < 64; ++k) s.heavy.push_back(runs[k].second);
      }
    }
  }
  stream_sync(ctx);
  cleanup.armed = false;
  *out = t.release();
  API_END
}

extern "C" int qm_table_drop(qm_table *t) {
  API_BEGIN
  if (!t) return QM_OK;
  qm_ctx *ctx = t->ctx;
  LaneHold hold(ctx, true);
  if (t->prepared.load()) fail(QM_ERR_INVALID_ARG, "table %s: close its prepared batches first", t->name.c_str());
  stream_sync(ctx);
  for (auto &c : t->cols) {
    if (c.d) ctx->release(c.d, c.bytes);
    if (c.dnull) ctx->release(c.dnull, c.null_bytes);
  }
  delete t;
  API_END
}

extern "C" uint64_t qm_table_rows(const qm_table *t) { return t ? t->rows : 0; }

extern "C" int qm_plan_compile(qm_ctx *ctx, const uint8_t *plan, size_t len, qm_plan **out) {
  API_BEGIN
  if (!ctx || !out || (!plan && len)) fail(QM_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  std::unique_ptr<qm_plan> p(new qm_plan());
  p->ctx = ctx;
  p->plan = parse_plan(plan, len);
  if (p->plan.has_join && p->plan.jtype != 0)
    fail(QM_ERR_UNSUPPORTED, "only inner joins run on the device");
  *out = p.release();
  API_END
}

extern "C" int qm_plan_close(qm_plan *p) {
  delete p;
  return QM_OK;
}

// ---------------------------------------------------------------------------
// execution
// ---------------------------------------------------------------------------
namespace {

uint64_t referenced_bytes(const ColSpace &cs, uint64_t lrows, uint64_t rrows) {
  uint64_t b = 0;
  for (size_t i = 0; i < cs.cols.size(); ++i)
    b += (uint64_t)dt_size(cs.cols[i].dtype) * (cs.cols[i].side ? rrows : lrows);
  return b;
}

void scan_launch_shape(qm_ctx *ctx, uint64_t nrows, uint64_t lds_words, int *grid, int *block, size_t *lds_bytes) {
  *block = 256;
  const uint64_t tile = 256 * RPT;
  uint64_t tiles = std::max<uint64_t>(1, (nrows + tile - 1) / tile);
  *lds_bytes = (size_t)(lds_words + MAX_QUERIES) * 8;
  int per_cu = 4;
  if (*lds_bytes > 0) per_cu = std::max(1, std::min(4, (int)((160 * 1024) / std::max<size_t>(*lds_bytes, 1))));
  *grid = (int)std::min<uint64_t>(tiles, (uint64_t)ctx->num_cus * per_cu);
}

// Partitioned aggregation: pass 1 (filter + compact + histogram), scan of
// the partition histogram, pass 2 (scatter), pass 3 (per-partition LDS
// aggregation into the global table). Returns the number of launches.
uint32_t run_partitioned(qm_ctx *ctx, const DScan &sc, const DQuery *d_q, const AggBinding &ab, uint64_t nrows,
                         uint64_t *d_selected) {
  const DQuery &q = ab.q;
  const uint64_t tile_rows = 256 * RPT;
  const uint64_t ntiles = std::max<uint64_t>(1, (nrows + tile_rows - 1) / tile_rows);
  // register-descriptor pass 1 when the filter is one range per column
  bool fast = q.and_only && ab.srcs.size() <= 4;
  if (fast) {
    int per_slot[MAX_PSLOTS] = {0, 0, 0, 0};
    for (int i = 0; i < q.npreds; ++i) {
      const DPred &p = q.preds[i];
      if (p.kind == PK_INT_IN || p.kind == PK_F64_IN) fast = false;
      if (p.slot >= 0 && p.slot < MAX_PSLOTS) per_slot[(int)p.slot]++;
    }
    for (int s2 = 0; s2 < q.nslots; ++s2) if (per_slot[s2] != 1) fast = false;
    if (q.nslots < 1 || q.key.ncols < 1) fast = false;
  }
  const int fast_code = fast ? q.nslots * 4 + q.key.ncols : 0;
  const size_t hist_lds = (size_t)(ab.nparts + 1) * 4;
  const int bpc = qmk_part_scan_blocks_per_cu(fast_code, hist_lds);
  uint64_t grid = std::min<uint64_t>(ntiles, (uint64_t)ctx->num_cus * bpc);
  const uint64_t tpb = (ntiles + grid - 1) / grid;
  grid = (ntiles + tpb - 1) / tpb;
  const uint64_t P = (uint64_t)ab.nparts;
  DPart pp;
  memset(&pp, 0, sizeof pp);
  pp.nsrc = (int)ab.srcs.size();
  for (int j = 0; j < pp.nsrc; ++j) {
    pp.src_a[j] = (int8_t)ab.srcs[j].first;
    pp.src_b[j] = (int8_t)ab.srcs[j].second;
    pp.src_float[j] = (ab.srcs[j].second >= 0 || dt_float(sc.cols[ab.srcs[j].first].dtype)) ? 1 : 0;
  }
  for (int si = 0; si < MAX_STATES; ++si) pp.state_src[si] = ab.state_src[si];
  pp.has_mask = ab.has_mask ? 1 : 0;
  // 32- or 64-byte records (whole-sector writes); the mask is the last word
  pp.rec_words = (1 + pp.nsrc + pp.has_mask <= 4) ? 4 : 8;
  pp.pbits = ab.pbits;
  pp.sbits = ab.sbits;
  pp.nparts = (int)P;
  pp.lds_entries = ab.part_lc;
  const uint64_t R = (uint64_t)pp.rec_words;
  const uint64_t blk_cap = tpb * tile_rows;
  // partition P (one past the last) collects the heavy-hitter records
  const uint64_t P1 = P + 1;
  DevBuf recs(ctx, grid * blk_cap * R * 8), bcount(ctx, grid * 8 + 64), hist(ctx, P1 * grid * 8);
  DevBuf hot(ctx, HOT_SLOTS * 8);
  if (!ab.hot_table.empty()) {
    h2d(ctx, hot.p, ab.hot_table.data(), HOT_SLOTS);
    pp.hot_keys = hot.as<uint64_t>();
    pp.hot_n = HOT_SLOTS;
  }
  pp.recs = recs.as<uint64_t>();
  pp.blk_cap = blk_cap;
  pp.blk_count = bcount.as<uint64_t>();
  pp.hist = hist.as<uint64_t>();
  ev_record(ctx, ctx->lane().evk0);
  HIPX(qmk_part_scan(&sc, d_q, &pp, (uint32_t)grid, tpb, fast_code, 1, ctx->lane().stream));
  ev_record(ctx, ctx->lane().evk1);
  DevBuf poff(ctx, P1 * grid * 8), stmp(ctx, qmk_exscan_tmp_words(P1 * grid) * 8 + 64);
  uint64_t *total_d = stmp.as<uint64_t>() + qmk_exscan_tmp_words(P1 * grid);
  HIPX(qmk_exscan_u64(hist.as<uint64_t>(), P1 * grid, poff.as<uint64_t>(), total_d, stmp.as<uint64_t>(),
                      ctx->lane().stream));
  uint64_t total = 0;
  d2h(ctx, &total, total_d, 1);
  h2d(ctx, d_selected, &total, 1);
  DevBuf precs(ctx, std::max<uint64_t>(total, 1) * R * 8);
  pp.part_off = poff.as<uint64_t>();
  pp.precs = precs.as<uint64_t>();
  // one pass-1 region per scatter workgroup by default (every CU busy)
  const uint32_t rpb = ctx->tune.scatter_rpb ? (uint32_t)ctx->tune.scatter_rpb : 1u;
  HIPX(qmk_part_scatter(d_q, &pp, (uint32_t)grid, rpb, q.entry_words, ctx->lane().stream));
  DevBuf starts(ctx, (P1 + 1) * 8);
  HIPX(qmk_part_starts(poff.as<uint64_t>(), (uint32_t)P1, (uint32_t)grid, total, starts.as<uint64_t>(), ctx->lane().stream));
  std::vector<uint64_t> st(P1 + 1);
  d2h(ctx, st.data(), starts.p, P1 + 1);
  // segments: split partitions larger than seg_max across workgroups
  const uint64_t seg_max = std::max<uint64_t>(16384, total / ((uint64_t)ctx->num_cus * 2) + 1);
  std::vector<uint64_t> seg;
  for (uint64_t p = 0; p < P; ++p) {
    uint64_t a = st[p], b = st[p + 1];
    if (b <= a) continue;
    uint64_t n = b - a, nsplit = (n + seg_max - 1) / seg_max;
    for (uint64_t s = 0; s < nsplit; ++s) {
      uint64_t s0 = a + s * seg_max, s1 = std::min(b, s0 + seg_max);
      seg.push_back(p);
      seg.push_back(s0);
      seg.push_back(s1 - s0);
      seg.push_back(nsplit == 1 ? 1 : 0);
    }
  }
  const uint64_t nseg = seg.size() / 4;
  DevBuf segbuf(ctx, std::max<uint64_t>(seg.size(), 4) * 8);
  h2d(ctx, segbuf.p, seg.data(), seg.size());
  pp.seg = segbuf.as<uint64_t>();
  pp.inserted = q.selected + 1;
  HIPX(qmk_part_agg(d_q, &pp, (uint32_t)nseg, (uint32_t)ctx->num_cus, q.entry_words, ctx->lane().stream));
  stream_sync(ctx);
  return 6;
}

// Run one or more aggregate queries (same table) through the fused scan and
// finalise each into a result.
void run_aggregates(qm_ctx *ctx, const qm_table *T, const qm_table *R, const std::vector<const Plan *> &plans,
                    std::vector<qm_result *> &results, bool fused_join, JoinIndex *ji,
                    const std::vector<KeyRange> *parts = nullptr) {
  const size_t nq = plans.size();
  std::vector<ColSpace> spaces(1);
  ColSpace &cs = spaces[0];
  cs.L = T;
  cs.R = fused_join ? R : nullptr;
  cs.rkey = fused_join ? plans[0]->rkey : "";
  std::vector<AggBinding> ab(nq);
  std::vector<ListStore> ls(nq);
  std::vector<uint32_t> regrow(nq, 0);
  DevBuf jpay;  // fused probe: the build columns packed per DIRECT key slot
  for (uint32_t attempt = 0;; ++attempt) {
    Timer tm(ctx);
    uint32_t launches = 0;
    std::vector<DevBuf> tables(nq), status(nq), selected(nq);
    uint64_t lds_off = 0;
    for (size_t i = 0; i < nq; ++i) {
      cs.plan = plans[i];
      memset(&ab[i].q, 0, sizeof(DQuery));
      ab[i].outs.clear();
      ab[i].key_slots.clear();
      ls[i].bufs.clear();
      bind_filters(ctx, cs, plans[i]->filters, plans[i]->groups, ab[i].q, ls[i], false);
      bind_aggregate(ctx, cs, *plans[i], T->rows, fused_join, ab[i], regrow[i]);
      DQuery &q = ab[i].q;
      tables[i] = DevBuf(ctx, ab[i].table_words * 8);
      zero(ctx, tables[i].p, ab[i].table_words * 8);
      ++launches;
      status[i] = DevBuf(ctx, 64);
      selected[i] = DevBuf(ctx, 64);
      zero(ctx, status[i].p, 64);
      zero(ctx, selected[i].p, 64);
      q.table = tables[i].as<uint64_t>();
      q.status = status[i].as<uint32_t>();
      q.selected = selected[i].as<uint64_t>();
      q.lds_offset = (int)lds_off;
      lds_off += (uint64_t)q.lds_words;
    }
    // launch: one pass, or one per key range of a partitioned join (every
    // pass adds into the same group tables)
    const bool ranged = fused_join && parts && !parts->empty();
    if (ranged && ab[0].part) fail(QM_ERR_UNSUPPORTED, "a join in key-range passes with a partitioned group-by");
    std::vector<DQuery> hq(nq);
    for (size_t i = 0; i < nq; ++i) hq[i] = ab[i].q;
    DevBuf dq(ctx, nq * sizeof(DQuery));
    h2d(ctx, dq.p, reinterpret_cast<const uint8_t *>(hq.data()), nq * sizeof(DQuery));
    int grid, block;
    size_t lds_bytes;
    scan_launch_shape(ctx, T->rows, lds_off, &grid, &block, &lds_bytes);
    if (lds_bytes > 160 * 1024) fail(QM_ERR_UNSUPPORTED, "fused queries exceed the workgroup LDS");
    const size_t npass = ranged ? parts->size() : 1;
    for (size_t pi = 0; pi < npass; ++pi) {
      DScan sc;
      memset(&sc, 0, sizeof sc);
      sc.nrows = T->rows;
      sc.ncols = (int)cs.cols.size();
      sc.nqueries = (int)nq;
      for (size_t i = 0; i < cs.cols.size(); ++i) sc.cols[i] = cs.cols[i];
      if (!fused_join && !(nq == 1 && ab[0].part)) sc.page_skip = T->page_skip;  // zone maps of a single query
      JoinIndex jrange;
      DevBuf jkeep;
      if (fused_join) {
        const JoinIndex *jc = ji;
        if (ranged) {
          build_join(ctx, cs, *plans[0], jrange, &launches, &jkeep, true, &(*parts)[pi]);
          jc = &jrange;
          jpay.reset();  // the packed build columns of this range's index
        }
        int ps = cs.plain(plans[0]->lkey);
        sc.ncols = (int)cs.cols.size();
        for (size_t i = 0; i < cs.cols.size(); ++i) sc.cols[i] = cs.cols[i];
        attach_join(sc, *jc, ps);
        sc.join_unique = jc->unique.as<uint32_t>();  // fused aggregate: one gather per probe row
        if (jc->unique.p && ctx->tune.join_payload) attach_join_payload(ctx, cs, *jc, sc, jpay, &launches);
      }
      if (T->rows && nq == 1 && ab[0].part) {
        launches += run_partitioned(ctx, sc, dq.as<DQuery>(), ab[0], T->rows, selected[0].as<uint64_t>());
      } else if (T->rows) {
        if (pi == 0) ev_record(ctx, ctx->lane().evk0);
        HIPX(qmk_scan_agg(&sc, dq.as<DQuery>(), grid, block, lds_bytes, ctx->lane().stream));
        if (pi + 1 == npass) ev_record(ctx, ctx->lane().evk1);
        ++launches;
      }
    }
    // overflow check
    bool retry = false;
    for (size_t i = 0; i < nq; ++i) {
      uint32_t st[2];
      d2h(ctx, st, status[i].p, 2);
      if (st[0]) { regrow[i] += 1; retry = true; }
    }
    if (retry) {
      if (attempt >= 4) fail(QM_ERR_QUERY_FAILED, "group-by hash table kept overflowing");
      continue;
    }
    // finalise each query
    for (size_t i = 0; i < nq; ++i) {
      AggBinding &b = ab[i];
      DFinal f;
      memset(&f, 0, sizeof f);
      f.table = tables[i].as<uint64_t>();
      f.nslots = b.nslots;
      f.entry_words = b.q.entry_words;
      f.count_word = b.q.count_word;
      f.gmode = b.q.gmode;
      f.always_one_row = b.always_one_row ? 1 : 0;
      f.key = b.q.key;
      f.status = status[i].as<uint32_t>();
      uint64_t nb = qmk_compact_blocks(f.nslots);
      DevBuf cnt(ctx, nb * 8), offs(ctx, nb * 8), stmp(ctx, qmk_exscan_tmp_words(nb) * 8 + 64);
      HIPX(qmk_compact_count(&f, cnt.as<uint64_t>(), ctx->lane().stream));
      uint64_t *total_d = stmp.as<uint64_t>() + qmk_exscan_tmp_words(nb);
      HIPX(qmk_exscan_u64(cnt.as<uint64_t>(), nb, offs.as<uint64_t>(), total_d, stmp.as<uint64_t>(), ctx->lane().stream));
      uint64_t total = 0;
      d2h(ctx, &total, total_d, 1);
      qm_result *res = new qm_result();
      results[i] = res;
      res->ctx = ctx;
      res->rows = total;
      f.nouts = (int)b.outs.size();
      for (size_t k = 0; k < b.outs.size(); ++k) {
        const DOut &o = b.outs[k].o;
        bool can_null = o.kind != OK_COUNT;
        OutCol oc = make_out(ctx, b.outs[k].name, o.out_dtype, total, can_null);
        f.outs[k] = o;
        f.outs[k].data = oc.data.p;
        f.outs[k].nulls = oc.nulls.as<uint32_t>();
        res->cols.push_back(std::move(oc));
      }
      if (total) HIPX(qmk_compact_write(&f, offs.as<uint64_t>(), ctx->lane().stream));
      launches += 4;
      uint32_t st[2];
      d2h(ctx, st, status[i].p, 2);
      if (st[1]) fail(QM_ERR_OVERFLOW, "SUM over an int64 column does not fit int64");
      uint64_t sel = 0;
      d2h(ctx, &sel, selected[i].p, 1);
      res->stats.rows_selected = sel;
      res->stats.strategy = b.q.gmode == GM_DIRECT ? (b.q.lmode == LM_DIRECT ? QM_STRAT_DIRECT_LDS : QM_STRAT_DIRECT_GLB)
                                                   : (b.q.lmode == LM_HASH ? QM_STRAT_HASH_LDS : QM_STRAT_HASH_GLB);
      if (b.part) res->stats.strategy = b.q.gmode == GM_DIRECT ? QM_STRAT_PART_DIRECT : QM_STRAT_PART_HASH;
      res->stats.retries = attempt;
      DQuery noq;
      memset(&noq, 0, sizeof noq);
      DevBuf dnoq(ctx, sizeof(DQuery));
      h2d(ctx, dnoq.p, reinterpret_cast<const uint8_t *>(&noq), sizeof noq);
      sort_limit_result(res, *plans[i], dnoq.as<DQuery>(), &launches);
    }
    double ms = tm.stop();
    float scan_ms = 0;
    if (T->rows) ev_ms(ctx, &scan_ms, ctx->lane().evk0, ctx->lane().evk1);
    uint64_t rrows = fused_join ? (ji ? ji->build_rows : R->rows) : 0;
    for (size_t i = 0; i < nq; ++i) {
      qm_result *res = results[i];
      res->stats.rows_scanned = T->rows + (fused_join ? R->rows : 0);
      res->stats.groups = res->rows;
      res->stats.bytes_algorithmic = referenced_bytes(cs, T->rows, rrows);
      res->stats.device_ms = ms;
      res->stats.scan_ms = scan_ms;
      res->stats.kernel_launches = launches;
      res->stats.join_parts = fused_join ? (uint32_t)std::max<size_t>(1, ranged ? parts->size() : 1) : 0;
    }
    return;
  }
}

// projection list for row-returning queries
std::vector<std::pair<std::string, int>> projection(ColSpace &cs, const Plan &pl, bool join) {
  std::vector<std::pair<std::string, int>> out;
  if (!pl.proj.empty()) {
    for (const std::string &n : pl.proj) out.push_back({n, cs.plain(n)});
    return out;
  }
  for (const TCol &c : cs.L->cols) out.push_back({c.name, cs.slot_of(&c, 0)});
  if (join) {
    // right columns minus the right key, colliding names prefixed right_
    for (const TCol &c : cs.R->cols) {
      if (c.name == cs.rkey) continue;
      std::string n = (cs.L->find(c.name) >= 0) ? "right_" + c.name : c.name;
      out.push_back({n, cs.slot_of(&c, 1)});
    }
  }
  return out;
}

void run_rows_query(qm_ctx *ctx, const qm_table *T, const qm_table *R, const Plan &pl, qm_result *res) {
  Timer tm(ctx);
  uint32_t launches = 0;
  bool timed_scan = false;
  ColSpace cs;
  cs.L = T;
  cs.R = pl.has_join ? R : nullptr;
  cs.rkey = pl.rkey;
  cs.plan = &pl;
  DQuery q;
  memset(&q, 0, sizeof q);
  ListStore ls;
  bind_filters(ctx, cs, pl.filters, pl.groups, q, ls, false);
  auto proj = projection(cs, pl, pl.has_join);
  res->ctx = ctx;
  DevBuf left_rows, right_rows;
  uint64_t n = 0;
  DScan sc;
  memset(&sc, 0, sizeof sc);
  sc.nrows = T->rows;
  sc.nqueries = 1;
  if (!pl.has_join) sc.page_skip = T->page_skip;  // zone maps: pages no row of which can be selected
  DevBuf dq(ctx, sizeof(DQuery));
  h2d(ctx, dq.p, reinterpret_cast<const uint8_t *>(&q), sizeof q);
  if (pl.has_join) {
    // one pass, or one per key range (the pairs of each range follow the
    // previous range's, as the reference's partitioned join emits them
    // partition by partition, src/operators.ts:2466-2637)
    const std::vector<KeyRange> ranges = join_ranges(ctx, R, pl, false);
    const size_t npass = ranges.empty() ? 1 : ranges.size();
    uint64_t want = ~0ull;
    if (!pl.has_sort && pl.limit >= 0 && !pl.post()) want = (uint64_t)pl.limit + (uint64_t)pl.offset;
    std::vector<DevBuf> lparts, rparts;
    std::vector<uint64_t> nparts;
    int ps = cs.plain(pl.lkey);
    sc.ncols = (int)cs.cols.size();
    for (size_t i = 0; i < cs.cols.size(); ++i) sc.cols[i] = cs.cols[i];
    for (size_t pi = 0; pi < npass && n < want; ++pi) {
      JoinIndex jr;
      DevBuf kr;
      build_join(ctx, cs, pl, jr, &launches, &kr, false, ranges.empty() ? nullptr : &ranges[pi]);
      DScan sp = sc;
      attach_join(sp, jr, ps);
      const uint64_t tile_rows = 256 * RPT;
      const uint64_t ntiles = std::max<uint64_t>(1, (sp.nrows + tile_rows - 1) / tile_rows);
      uint64_t nblocks = std::min<uint64_t>(ntiles, (uint64_t)ctx->num_cus * 8);
      uint64_t tpb = (ntiles + nblocks - 1) / nblocks;
      nblocks = (ntiles + tpb - 1) / tpb;
      DevBuf counts(ctx, nblocks * 8), offs(ctx, nblocks * 8), stmp(ctx, qmk_exscan_tmp_words(nblocks) * 8 + 64);
      HIPX(qmk_join_count(&sp, dq.as<DQuery>(), nblocks, tpb, counts.as<uint64_t>(), ctx->lane().stream));
      uint64_t *total_d = stmp.as<uint64_t>() + qmk_exscan_tmp_words(nblocks);
      HIPX(qmk_exscan_u64(counts.as<uint64_t>(), nblocks, offs.as<uint64_t>(), total_d, stmp.as<uint64_t>(),
                          ctx->lane().stream));
      uint64_t total = 0;
      d2h(ctx, &total, total_d, 1);
      const uint64_t take = std::min<uint64_t>(total, want - n);
      DevBuf lr(ctx, std::max<uint64_t>(take, 1) * 8), rr(ctx, std::max<uint64_t>(take, 1) * 8);
      if (take)
        HIPX(qmk_join_write(&sp, dq.as<DQuery>(), nblocks, tpb, offs.as<uint64_t>(), take, lr.as<uint64_t>(),
                            rr.as<uint64_t>(), ctx->lane().stream));
      launches += 5;
      n += take;
      lparts.push_back(std::move(lr));
      rparts.push_back(std::move(rr));
      nparts.push_back(take);
    }
    if (lparts.size() == 1) {
      left_rows = std::move(lparts[0]);
      right_rows = std::move(rparts[0]);
    } else {
      left_rows = DevBuf(ctx, std::max<uint64_t>(n, 1) * 8);
      right_rows = DevBuf(ctx, std::max<uint64_t>(n, 1) * 8);
      uint64_t at = 0;
      for (size_t pi = 0; pi < lparts.size(); ++pi) {
        if (!nparts[pi]) continue;
        HIPX(hipMemcpyAsync(left_rows.as<uint64_t>() + at, lparts[pi].p, nparts[pi] * 8, hipMemcpyDeviceToDevice,
                            ctx->lane().stream));
        HIPX(hipMemcpyAsync(right_rows.as<uint64_t>() + at, rparts[pi].p, nparts[pi] * 8, hipMemcpyDeviceToDevice,
                            ctx->lane().stream));
        at += nparts[pi];
      }
    }
    res->stats.join_parts = (uint32_t)npass;
    res->stats.rows_scanned = T->rows + R->rows;
  } else {
    sc.ncols = (int)cs.cols.size();
    for (size_t i = 0; i < cs.cols.size(); ++i) sc.cols[i] = cs.cols[i];
    const bool topk = !pl.post() && pl.has_sort && pl.limit >= 0 &&
                      qmk_topk_cap((uint32_t)std::min<int64_t>(pl.limit + pl.offset, 1 << 30)) &&
                      pl.limit + pl.offset < (1 << 30);
    res->stats.rows_scanned = T->rows;
    if (topk) {
      // fused filter + top-K over the table (src/operators.ts:1716-1841)
      int sslot = cs.plain(pl.sort_col);
      sc.ncols = (int)cs.cols.size();
      for (size_t i = 0; i < cs.cols.size(); ++i) sc.cols[i] = cs.cols[i];
      h2d(ctx, dq.p, reinterpret_cast<const uint8_t *>(&q), sizeof q);
      uint64_t k = (uint64_t)(pl.limit + pl.offset);
      const uint64_t tiles = std::max<uint64_t>(1, (T->rows + 2047) / 2048);
      int grid = (int)std::min<uint64_t>(tiles, (uint64_t)ctx->num_cus * (uint64_t)ctx->tune.topk_bpc);
      uint64_t take = 0;
      if (k > 0) {
        DevBuf cand(ctx, (size_t)grid * k * sizeof(TopKey));
        DevBuf ccount(ctx, (size_t)grid * 4 + 64), gthr(ctx, 64);
        HIPX(qmk_fill_u64(gthr.as<uint64_t>(), 1, ~0ull, ctx->lane().stream));
        ev_record(ctx, ctx->lane().evk0);
        HIPX(qmk_topk_scan(&sc, dq.as<DQuery>(), sslot, pl.desc ? 1 : 0, (uint32_t)k, grid, cand.as<TopKey>(),
                           ccount.as<uint32_t>(), gthr.as<uint64_t>(), ctx->lane().stream));
        ev_record(ctx, ctx->lane().evk1);
        timed_scan = true;
        ++launches;
        const TopKey *lists = cand.as<TopKey>();
        const uint32_t *lc = ccount.as<uint32_t>();
        uint32_t nl = (uint32_t)grid;
        DevBuf m1, m1c;
        if (nl > 16) {
          uint32_t g2 = (nl + 15) / 16;
          m1 = DevBuf(ctx, (size_t)g2 * k * sizeof(TopKey));
          m1c = DevBuf(ctx, (size_t)g2 * 4 + 64);
          HIPX(qmk_topk_merge(lists, lc, nl, 16, (uint32_t)k, m1.as<TopKey>(), m1c.as<uint32_t>(), ctx->lane().stream));
          ++launches;
          lists = m1.as<TopKey>();
          lc = m1c.as<uint32_t>();
          nl = g2;
        }
        DevBuf fin(ctx, (size_t)k * sizeof(TopKey)), finc(ctx, 64);
        HIPX(qmk_topk_merge(lists, lc, nl, nl, (uint32_t)k, fin.as<TopKey>(), finc.as<uint32_t>(), ctx->lane().stream));
        uint32_t fc = 0;
        d2h(ctx, &fc, finc.p, 1);
        take = fc > (uint64_t)pl.offset ? fc - (uint64_t)pl.offset : 0;
        left_rows = DevBuf(ctx, std::max<uint64_t>(take, 1) * 8);
        HIPX(qmk_topk_rows(fin.as<TopKey>(), finc.as<uint32_t>(), (uint32_t)pl.offset, left_rows.as<uint64_t>(),
                           (uint32_t)take, ctx->lane().stream));
        launches += 2;
      } else {
        left_rows = DevBuf(ctx, 8);
      }
      n = take;
    } else {
      uint64_t cap = ~0ull;
      if (!pl.has_sort && pl.limit >= 0 && !pl.post()) cap = (uint64_t)pl.limit + (uint64_t)pl.offset;
      left_rows = select_rows(ctx, sc, dq.as<DQuery>(), cap, &n, &launches);
    }
  }
  res->rows = n;
  if (pl.post()) {
    // DISTINCT / multi-column ORDER BY over every selected row, in row order:
    // the projected columns plus hidden key columns, then the stages, OFFSET /
    // LIMIT, and the projection
    std::vector<std::pair<std::string, int>> cols = proj;
    auto want = [&](const std::string &nm) {
      for (auto &pc : cols) if (pc.first == nm) return;
      cols.push_back({nm, cs.plain(nm)});
    };
    for (const std::string &d : pl.distinct) want(d);
    for (const OrderKey &k : pl.order_by) want(k.col);
    if (pl.has_sort) want(pl.sort_col);
    for (auto &pc : cols) {
      const DCol &c = cs.cols[pc.second];
      const int odt = out_dt_of(c.dtype);
      OutCol oc = make_out(ctx, pc.first, odt, n, c.nulls != nullptr);
      const uint64_t *rows = c.side ? right_rows.as<uint64_t>() : left_rows.as<uint64_t>();
      HIPX(qmk_gather(&c, rows, nullptr, n, odt, oc.data.p, oc.nulls.as<uint32_t>(), ctx->lane().stream));
      ++launches;
      res->cols.push_back(std::move(oc));
    }
    if (pl.has_distinct) {
      std::vector<std::string> dn = pl.distinct;
      if (dn.empty()) for (auto &pc : proj) dn.push_back(pc.first);
      distinct_stage(res, dn);
      launches += 8;
    }
    if (!pl.order_by.empty()) {
      order_by_stage(res, pl.order_by, pl.limit, pl.offset);
      launches += 4 * (uint32_t)pl.order_by.size();
    } else if (pl.has_sort || pl.limit >= 0 || pl.offset > 0) {
      Plan p2 = pl;
      p2.has_distinct = false;
      p2.distinct.clear();
      DQuery noq;
      memset(&noq, 0, sizeof noq);
      DevBuf dnoq(ctx, sizeof(DQuery));
      h2d(ctx, dnoq.p, reinterpret_cast<const uint8_t *>(&noq), sizeof noq);
      sort_limit_result(res, p2, dnoq.as<DQuery>(), &launches);
    }
    std::vector<std::string> names;
    for (auto &pc : proj) names.push_back(pc.first);
    project_result(res, names);
    res->stats.device_ms = tm.stop();
    res->stats.groups = res->rows;
    res->stats.rows_selected = n;
    res->stats.kernel_launches = launches;
    res->stats.bytes_algorithmic = referenced_bytes(cs, T->rows, pl.has_join ? R->rows : 0);
    return;
  }
  // gather the projected columns
  for (auto &pc : proj) {
    const DCol &c = cs.cols[pc.second];
    int odt = out_dt_of(c.dtype);
    OutCol oc = make_out(ctx, pc.first, odt, n, c.nulls != nullptr);
    const uint64_t *rows = c.side ? right_rows.as<uint64_t>() : left_rows.as<uint64_t>();
    HIPX(qmk_gather(&c, rows, nullptr, n, odt, oc.data.p, oc.nulls.as<uint32_t>(), ctx->lane().stream));
    ++launches;
    res->cols.push_back(std::move(oc));
  }
  const bool topk_done = !pl.has_join && pl.has_sort && pl.limit >= 0 &&
                         qmk_topk_cap((uint32_t)std::min<int64_t>(pl.limit + pl.offset, 1 << 30)) &&
                         pl.limit + pl.offset < (1 << 30);
  if (!topk_done) {
    // sort / limit / offset over the materialised rows
    Plan p2 = pl;
    if (!pl.has_sort && !pl.has_join) p2.limit = -1, p2.offset = pl.offset;
    if (!pl.has_sort && pl.has_join) { /* n already capped at limit+offset */ }
    DQuery noq;
    memset(&noq, 0, sizeof noq);
    DevBuf dnoq(ctx, sizeof(DQuery));
    h2d(ctx, dnoq.p, reinterpret_cast<const uint8_t *>(&noq), sizeof noq);
    if (pl.has_sort && find_out(res, pl.sort_col) < 0) {
      // sort column not projected: gather it as a hidden column
      int s = cs.plain(pl.sort_col);
      const DCol &c = cs.cols[s];
      OutCol oc = make_out(ctx, "\x01sort", out_dt_of(c.dtype), n, c.nulls != nullptr);
      const uint64_t *rows = c.side ? right_rows.as<uint64_t>() : left_rows.as<uint64_t>();
      HIPX(qmk_gather(&c, rows, nullptr, n, oc.dtype, oc.data.p, oc.nulls.as<uint32_t>(), ctx->lane().stream));
      res->cols.push_back(std::move(oc));
      p2.sort_col = "\x01sort";
      sort_limit_result(res, p2, dnoq.as<DQuery>(), &launches);
      res->cols.pop_back();
    } else if (pl.has_sort || pl.offset > 0 || (pl.limit >= 0 && pl.has_join)) {
      sort_limit_result(res, p2, dnoq.as<DQuery>(), &launches);
    }
  }
  res->stats.device_ms = tm.stop();
  if (timed_scan) {
    float sm = 0;
    ev_ms(ctx, &sm, ctx->lane().evk0, ctx->lane().evk1);
    res->stats.scan_ms = sm;
  }
  res->stats.groups = res->rows;
  res->stats.rows_selected = n;
  res->stats.kernel_launches = launches;
  res->stats.bytes_algorithmic = referenced_bytes(cs, T->rows, pl.has_join ? R->rows : 0);
}

// ---------------------------------------------------------------------------
// batched aggregate queries over one table: one fused pass (k_batch.hip)
// ---------------------------------------------------------------------------
// BK_RUNS: grouped by a sorted key too sparse for a DIRECT table; only the
// run-sorted kernels fuse such queries (they emit one group per run and need
// no table), the rest run through the per-query path
// BK_PRUNS: grouped by an integer key in any row order with a key histogram;
// the partitioned-runs kernels (k_pruns.hip) fuse such queries
enum : int { BK_NONE = 0, BK_SHARED = 1, BK_SORTED = 2, BK_RUNS = 3, BK_PRUNS = 4 };

// one query of a fused launch
struct BItem {
  uint32_t idx = 0;                 // position in the caller's plan list
  const Plan *plan = nullptr;
  ColSpace cs;                      // the query's own column space (slots)
  AggBinding ab;
  ListStore ls;
  int kind = BK_NONE;
  int nslots = 0;
  bool never = false;               // a predicate can never hold
  const TCol *slot_t[MAX_PSLOTS] = {};
  uint64_t lo[MAX_PSLOTS] = {}, hi[MAX_PSLOTS] = {};
  const TCol *key_t = nullptr;
  const TCol *st_t[MAX_STATES] = {};
  DevBuf table;                     // the query's DIRECT table (final layout)
  uint64_t e_lo = 0, e_hi = 0;      // shared: covered intervals
  int group = -1;                   // shared group
};

// unsigned order-code range [lo, hi] of one bound predicate; false when the
// predicate is not a plain range (negated, NaN-matching, IN, NULL tests)
bool pred_code_range(const DPred &p, uint64_t &lo, uint64_t &hi, bool &never) {
  never = false;
  switch (p.kind) {
    case PK_FALSE: never = true; lo = 0; hi = 0; return true;
    case PK_IS_NOT_NULL: lo = 0; hi = ~0ull; return true;
    case PK_INT_RANGE:
      if (p.negate) return false;
      if (p.ilo > p.ihi) { never = true; return true; }
      lo = code_i64(p.ilo);
      hi = code_i64(p.ihi);
      return true;
    case PK_F64_RANGE: {
      if (p.negate || p.nan_result) return false;
      if (std::isnan(p.flo) || std::isnan(p.fhi)) { never = true; return true; }
      uint64_t a = code_f64_bits(f64_bits(p.flo)), b = code_f64_bits(f64_bits(p.fhi));
      if (p.lo_strict) { if (a == ~0ull) { never = true; return true; } ++a; }
      if (p.hi_strict) { if (b == 0) { never = true; return true; } --b; }
      if (a > b) { never = true; return true; }
      lo = a;
      hi = b;
      return true;
    }
  }
  return false;
}

// Bind one plan and decide whether the fused pass can run it.
void batch_classify(qm_ctx *ctx, const qm_table *T, const Plan &pl, BItem &it) {
  it.kind = BK_NONE;
  if (pl.has_join || pl.aggs.empty() || pl.groupby.size() > 1 || !pl.computed.empty()) return;
  ColSpace &cs = it.cs;
  cs.L = T;
  cs.plan = &pl;
  try {
    bind_filters(ctx, cs, pl.filters, pl.groups, it.ab.q, it.ls, false);
    bind_aggregate(ctx, cs, pl, T->rows, false, it.ab, 0);
  } catch (const Error &) {
    return;  // the single-query path reports the error
  }
  const DQuery &q = it.ab.q;
  if (!q.and_only || q.ngroups) return;
  // one predicate per slot (the planner merges the ranges of a column)
  it.nslots = q.nslots;
  for (int s = 0; s < q.nslots; ++s) {
    int found = 0;
    for (int i = 0; i < q.npreds; ++i) {
      if (q.preds[i].slot != s) continue;
      ++found;
      bool nv = false;
      if (!pred_code_range(q.preds[i], it.lo[s], it.hi[s], nv)) return;
      it.never = it.never || nv;
    }
    if (found != 1) return;
    it.slot_t[s] = cs.tcols[q.slot_col[s]];
  }
  for (int s = 0; s < q.nstates; ++s) {
    if (q.states[s].product || q.states[s].kind == SK_SUM_WIDE) return;
    it.st_t[s] = q.states[s].col_a >= 0 ? cs.tcols[q.states[s].col_a] : nullptr;
  }
  const int nk = q.key.ncols;
  // any row order: a key histogram sizes the key partitions
  auto try_pruns = [&]() {
    if (it.kind != BK_NONE || nk != 1 || it.nslots > 1 || q.nstates > PR_MAX_STATES) return;
    const TCol *kt = cs.tcols[q.key.col[0]];
    if (kt->st.nulls || dt_float(kt->dtype) || kt->khist_g < 0 || kt->khist.empty()) return;
    it.key_t = kt;
    it.kind = BK_PRUNS;
  };
  if (q.gmode == GM_HASH && nk == 1 && it.nslots <= 1 && q.nstates <= R_MAX_STATES) {
    const TCol *kt = cs.tcols[q.key.col[0]];
    if (kt->st.sorted && !kt->st.nulls && !dt_float(kt->dtype)) {
      it.key_t = kt;
      it.kind = BK_RUNS;
    }
  }
  if (q.gmode != GM_DIRECT) { try_pruns(); return; }
  if (nk == 1) {
    const TCol *kt = cs.tcols[q.key.col[0]];
    if (kt->st.nulls || dt_float(kt->dtype)) return;
    it.key_t = kt;
    if (q.key.dom[0] <= 64 && it.nslots == 1 && q.nstates <= B_MAX_SHARED_STATES) it.kind = BK_SHARED;
    else if (kt->st.sorted && q.nstates <= B_MAX_STATES && it.nslots <= B_MAX_SLOTS) it.kind = BK_SORTED;
  } else if (nk == 0 && it.nslots == 1 && q.nstates <= B_MAX_SHARED_STATES) {
    it.kind = BK_SHARED;
  }
  try_pruns();
}

// a shared-interval group: queries with the same range column and key
struct BGroup {
  const TCol *col = nullptr, *key = nullptr;
  std::vector<BItem *> items;
  std::vector<uint64_t> bounds;
  std::vector<DState> st;              // union of states, COUNT(*) first
  std::vector<const TCol *> st_t;
  int entry_words = 0;
  uint32_t dom = 1;
  int64_t kmin = 0;
  DevBuf table, dbounds;
  int launch = -1;
  bool lean = false;  // runs in its own lean pass (k_runs.hip shared_kernel)
};

// index of `s` (state of item `it`) in the group's state union
int group_state(BGroup &g, const BItem &it, int s) {
  const DState &a = it.ab.q.states[s];
  const TCol *t = it.st_t[s];
  for (size_t i = 0; i < g.st.size(); ++i) {
    const DState &b = g.st[i];
    if (b.kind == a.kind && b.is_min == a.is_min && b.fixed == a.fixed && b.scale == a.scale && g.st_t[i] == t)
      return (int)i;
  }
  DState d = a;
  d.word = (int16_t)g.entry_words;
  g.entry_words += (d.kind == SK_SUM_128) ? (d.fixed ? 3 : 2) : 1;
  g.st.push_back(d);
  g.st_t.push_back(t);
  return (int)g.st.size() - 1;
}

struct BLaunch {
  std::vector<const TCol *> cols;
  std::vector<BGroup *> groups;
  std::vector<BItem *> sorted;
  const TCol *sorted_key = nullptr;
  int col_of(const TCol *t) const {
    for (size_t i = 0; i < cols.size(); ++i) if (cols[i] == t) return (int)i;
    return -1;
  }
  bool fits(const std::vector<const TCol *> &more) const {
    std::vector<const TCol *> u = cols;
    for (const TCol *t : more)
      if (t && std::find(u.begin(), u.end(), t) == u.end()) u.push_back(t);
    return u.size() <= (size_t)B_MAX_COLS;
  }
  void add(const std::vector<const TCol *> &more) {
    for (const TCol *t : more) if (t && col_of(t) < 0) cols.push_back(t);
  }
};

// the aggregate list of a sorted-key query (query-parallel launches share one)
bool same_signature(const BItem &a, const BItem &b) {
  const DQuery &x = a.ab.q, &y = b.ab.q;
  if (x.nstates != y.nstates || x.entry_words != y.entry_words) return false;
  for (int s = 0; s < x.nstates; ++s) {
    const DState &p = x.states[s], &q = y.states[s];
    if (p.kind != q.kind || p.is_min != q.is_min || p.fixed != q.fixed || p.scale != q.scale || p.word != q.word ||
        a.st_t[s] != b.st_t[s])
      return false;
  }
  return true;
}

const int B_WAVES = B_BLOCK / 64;
const uint64_t B_LDS_WORDS = 156 * 1024 / 8;

uint64_t group_lds_words(const BGroup &g) {
  return g.bounds.size() + (g.bounds.size() + 1) * g.dom * (uint64_t)g.entry_words;
}

std::vector<const TCol *> item_cols(const BItem &it) {
  std::vector<const TCol *> v;
  for (int s = 0; s < it.nslots; ++s) v.push_back(it.slot_t[s]);
  if (it.key_t) v.push_back(it.key_t);
  for (int s = 0; s < it.ab.q.nstates; ++s) if (it.st_t[s]) v.push_back(it.st_t[s]);
  return v;
}

// Finalise many aggregate results with one synchronisation per phase. Small
// tables (the shared-interval queries' DIRECT tables) are compacted together
// in one launch into columns sized for every slot; larger ones take the
// count / scan / write sequence each.
void finalize_batch(qm_ctx *ctx, std::vector<BItem *> &items, std::vector<qm_result *> &out, DevBuf &status,
                    DevBuf &selected) {
  const size_t n = items.size();
  if (!n) return;
  std::vector<DFinal> f(n);
  std::vector<size_t> small, large;
  for (size_t i = 0; i < n; ++i) {
    BItem &it = *items[i];
    AggBinding &b = it.ab;
    DFinal &x = f[i];
    memset(&x, 0, sizeof x);
    x.table = it.table.as<uint64_t>();
    x.nslots = b.nslots;
    x.entry_words = b.q.entry_words;
    x.count_word = b.q.count_word;
    x.gmode = b.q.gmode;
    x.always_one_row = b.always_one_row ? 1 : 0;
    x.key = b.q.key;
    x.status = status.as<uint32_t>() + 2 * i;
    x.nouts = (int)b.outs.size();
    (b.nslots <= qmk_compact_small_slots() ? small : large).push_back(i);
  }
  // large tables: per-table block counts and their scan
  std::vector<DevBuf> cnt(n), offs(n);
  uint64_t max_nb = 1;
  for (size_t i : large) max_nb = std::max<uint64_t>(max_nb, qmk_compact_blocks(f[i].nslots));
  DevBuf stmp(ctx, qmk_exscan_tmp_words(max_nb) * 8 + 64), totals(ctx, n * 8 + 64);
  for (size_t i : large) {
    uint64_t nb = qmk_compact_blocks(f[i].nslots);
    cnt[i] = DevBuf(ctx, nb * 8);
    offs[i] = DevBuf(ctx, nb * 8);
    HIPX(qmk_compact_count(&f[i], cnt[i].as<uint64_t>(), ctx->lane().stream));
    HIPX(qmk_exscan_u64(cnt[i].as<uint64_t>(), nb, offs[i].as<uint64_t>(), totals.as<uint64_t>() + i,
                        stmp.as<uint64_t>(), ctx->lane().stream));
  }
  // small tables: outputs sized for every slot, one compaction launch
  for (size_t i : small) {
    BItem &it = *items[i];
    AggBinding &b = it.ab;
    qm_result *res = new qm_result();
    out[it.idx] = res;
    res->ctx = ctx;
    for (size_t k = 0; k < b.outs.size(); ++k) {
      const DOut &o = b.outs[k].o;
      OutCol oc = make_out(ctx, b.outs[k].name, o.out_dtype, f[i].nslots, o.kind != OK_COUNT, false);
      f[i].outs[k] = o;
      f[i].outs[k].data = oc.data.p;
      f[i].outs[k].nulls = oc.nulls.as<uint32_t>();
      res->cols.push_back(std::move(oc));
    }
  }
  DevBuf dsmall(ctx, std::max<size_t>(small.size(), 1) * sizeof(DFinal));
  if (!small.empty()) {
    std::vector<DFinal> fs;
    for (size_t i : small) fs.push_back(f[i]);
    h2d(ctx, dsmall.p, reinterpret_cast<const uint8_t *>(fs.data()), fs.size() * sizeof(DFinal));
    DevBuf scount(ctx, small.size() * 8 + 64);
    HIPX(qmk_compact_small(dsmall.as<DFinal>(), (uint32_t)small.size(), scount.as<uint64_t>(), ctx->lane().stream));
    std::vector<uint64_t> sc(small.size());
    d2h(ctx, sc.data(), scount.p, small.size());
    for (size_t j = 0; j < small.size(); ++j) out[items[small[j]]->idx]->rows = sc[j];
  }
  std::vector<uint64_t> tot(n, 0);
  if (!large.empty()) d2h(ctx, tot.data(), totals.p, n);
  for (size_t i : large) {
    BItem &it = *items[i];
    AggBinding &b = it.ab;
    qm_result *res = new qm_result();
    out[it.idx] = res;
    res->ctx = ctx;
    res->rows = tot[i];
    for (size_t k = 0; k < b.outs.size(); ++k) {
      const DOut &o = b.outs[k].o;
      OutCol oc = make_out(ctx, b.outs[k].name, o.out_dtype, tot[i], o.kind != OK_COUNT);
      f[i].outs[k] = o;
      f[i].outs[k].data = oc.data.p;
      f[i].outs[k].nulls = oc.nulls.as<uint32_t>();
      res->cols.push_back(std::move(oc));
    }
    if (tot[i]) HIPX(qmk_compact_write(&f[i], offs[i].as<uint64_t>(), ctx->lane().stream));
  }
  std::vector<uint32_t> st(2 * n);
  d2h(ctx, st.data(), status.p, 2 * n);
  std::vector<uint64_t> sel(n);
  d2h(ctx, sel.data(), selected.p, n);
  for (size_t i = 0; i < n; ++i) {
    if (st[2 * i + 1]) fail(QM_ERR_OVERFLOW, "SUM over an int64 column does not fit int64");
    qm_result *res = out[items[i]->idx];
    res->stats.rows_selected = sel[i];
    res->stats.strategy = items[i]->kind == BK_SHARED ? QM_STRAT_SHARED_INTERVAL : QM_STRAT_SORTED_RUNS;
  }
  DQuery noq;
  memset(&noq, 0, sizeof noq);
  DevBuf dnoq(ctx, sizeof(DQuery));
  h2d(ctx, dnoq.p, reinterpret_cast<const uint8_t *>(&noq), sizeof noq);
  for (size_t i = 0; i < n; ++i) {
    uint32_t l = 0;
    sort_limit_result(out[items[i]->idx], *items[i]->plan, dnoq.as<DQuery>(), &l);
  }
}

// ---- run-sorted launches: queries grouped by a sorted key, each filtered by
// at most one range on one common column (k_runs.hip) ----
struct RunsPlan {
  const TCol *pcol = nullptr;
  int perm[PR_MAX_STATES] = {};      // run state i = query state perm[i]
  int nsum = 0, nmax = 0;
  ROut outs[R_MAX_OUTS] = {};
  int nouts = 0;
  std::vector<uint64_t> bounds;      // sorted distinct bound codes
};

// any_order: the key may be stored in any row order (the partitioned-runs
// launch); otherwise it must be sorted with runs of runs_min_len rows
bool runs_plan(const qm_ctx *ctx, const qm_table *T, const BLaunch &L, RunsPlan &rp, bool any_order = false) {
  const size_t n = L.sorted.size();
  if (n == 0 || n > (size_t)R_MAX_QUERIES || !L.groups.empty()) return false;
  if (!any_order && (int)n < ctx->tune.runs_min_q) return false;
  if (T->rows >= (1ull << 32)) return false;  // 32-bit positions and per-chunk counters
  const TCol *key = L.sorted_key;
  if (!key || key->st.nulls || dt_float(key->dtype)) return false;
  if (!any_order && !key->st.sorted) return false;
  if (!any_order && (double)T->rows / std::max(1.0, key->st.distinct) < ctx->tune.runs_min_len) return false;
  for (const BItem *it : L.sorted) {
    if (it->nslots > 1) return false;
    if (it->nslots == 1 && !it->never) {
      if (rp.pcol && rp.pcol != it->slot_t[0]) return false;
      rp.pcol = it->slot_t[0];
    }
  }
  if (!rp.pcol) rp.pcol = key;
  const BItem &i0 = *L.sorted[0];
  const DQuery &q0 = i0.ab.q;
  if (q0.nstates < 1 || q0.nstates > (any_order ? PR_MAX_STATES : R_MAX_STATES)) return false;
  if (q0.states[0].kind != SK_COUNT || q0.states[0].col_a >= 0) return false;
  std::vector<int> adds, maxs;
  for (int s = 1; s < q0.nstates; ++s) {
    const DState &d = q0.states[s];
    if (d.product || !i0.st_t[s]) return false;
    if (d.kind == SK_SUM_I64) {
      if (T->rows >= (1ull << 32)) return false;  // |sum| of < 2^32 values of <= 32 bits fits int64
      adds.push_back(s);
    } else if (d.kind == SK_COUNT) {
      adds.push_back(s);
    } else if (d.kind == SK_MAX_CODE) {
      maxs.push_back(s);
    } else {
      return false;
    }
  }
  rp.nsum = (int)adds.size();
  rp.nmax = (int)maxs.size();
  if (!(any_order ? qmk_pr_supported(rp.nsum, rp.nmax) : qmk_runs_supported(rp.nsum, rp.nmax))) return false;
  int k = 0;
  rp.perm[k++] = 0;
  for (int s : adds) rp.perm[k++] = s;
  for (int s : maxs) rp.perm[k++] = s;
  auto state_of_word = [&](int w) -> int {
    for (int i = 0; i < q0.nstates; ++i) if (q0.states[rp.perm[i]].word == w) return i;
    return -1;
  };
  const auto &o0 = i0.ab.outs;
  if (o0.empty() || o0.size() > (size_t)R_MAX_OUTS) return false;
  for (const BItem *it : L.sorted) {
    const auto &o = it->ab.outs;
    if (o.size() != o0.size()) return false;
    for (size_t j = 0; j < o.size(); ++j) {
      const DOut &a = o[j].o, &b = o0[j].o;
      if (a.kind != b.kind || a.out_dtype != b.out_dtype || a.src_is_float != b.src_is_float || a.word != b.word ||
          a.cnt_word != b.cnt_word || a.key_idx != b.key_idx)
        return false;
    }
  }
  for (size_t j = 0; j < o0.size(); ++j) {
    const DOut &d = o0[j].o;
    ROut &r = rp.outs[j];
    switch (d.kind) {
      case OK_KEY: case OK_COUNT: case OK_SUM_I64: case OK_AVG_I64: case OK_MIN: case OK_MAX: break;
      default: return false;
    }
    r.kind = d.kind;
    r.src_is_float = d.src_is_float;
    r.out_dtype = d.out_dtype;
    // a COUNT output reads its count state only (its value word is unused)
    r.state = (int8_t)((d.kind == OK_KEY || d.kind == OK_COUNT) ? 0 : state_of_word(d.word));
    r.cnt_state = (int8_t)(d.cnt_word >= 0 ? state_of_word(d.cnt_word) : -1);
    if (r.state < 0 || (d.cnt_word >= 0 && r.cnt_state < 0)) return false;
  }
  rp.nouts = (int)o0.size();
  std::vector<uint64_t> &b = rp.bounds;
  for (const BItem *it : L.sorted) {
    if (it->never || it->nslots == 0) continue;
    b.push_back(it->lo[0]);
    if (it->hi[0] != ~0ull) b.push_back(it->hi[0] + 1);
  }
  std::sort(b.begin(), b.end());
  b.erase(std::unique(b.begin(), b.end()), b.end());
  return b.size() + 2 <= (size_t)R_MAX_BUCKETS;
}

// Count pass, offsets, then the write pass straight into each query's result
// columns (ascending key order). scan_ms: the longest of its launches that
// read the table so far.
void run_runs(qm_ctx *ctx, const qm_table *T, const BLaunch &L, const RunsPlan &rp, std::vector<qm_result *> &out,
              float &scan_ms) {
  const size_t n = L.sorted.size();
  const std::vector<uint64_t> &b = rp.bounds;
  const uint32_t nb = (uint32_t)b.size();
  int steps = 0;
  while (((1u << steps) - 1) < nb) ++steps;
  std::vector<uint64_t> padded((1u << steps) - 1, ~0ull);
  std::copy(b.begin(), b.end(), padded.begin());
  std::vector<uint32_t> qe(n);
  for (size_t i = 0; i < n; ++i) {
    const BItem &it = *L.sorted[i];
    uint32_t lo, hi;
    if (it.never) {
      lo = hi = 0;
    } else if (it.nslots == 0) {
      lo = 0;
      hi = nb + 2;  // every row, NULL filter values included
    } else {
      lo = (uint32_t)(std::upper_bound(b.begin(), b.end(), it.lo[0]) - b.begin());
      hi = it.hi[0] == ~0ull ? nb + 1 : (uint32_t)(std::upper_bound(b.begin(), b.end(), it.hi[0] + 1) - b.begin());
    }
    qe[i] = lo | (hi << 16);
  }
  const uint64_t rows = T->rows;
  const uint64_t target = (uint64_t)ctx->num_cus * (uint64_t)ctx->tune.runs_chunks_per_cu;
  uint64_t chunk_rows = std::max<uint64_t>(1024, (rows + target - 1) / target);
  chunk_rows = (chunk_rows + 63) / 64 * 64;
  const uint32_t nchunks = (uint32_t)std::max<uint64_t>(1, (rows + chunk_rows - 1) / chunk_rows);

  DevBuf dbounds(ctx, std::max<size_t>(padded.size(), 1) * 8), dqe(ctx, n * 4 + 64);
  DevBuf counts(ctx, (size_t)nchunks * n * 4 + 64), offs(ctx, (size_t)nchunks * n * 8 + 64);
  DevBuf totals(ctx, n * 8 + 64), selected(ctx, n * 8 + 64);
  const int qbw = ctx->tune.runs_qblock;  // 0, 16, 32 or 64
  const uint32_t nqb = qbw ? (uint32_t)((n + qbw - 1) / qbw) : 1u;
  const uint64_t nseg = (uint64_t)nchunks * nqb;  // staging ranges: (chunk, query block)
  DevBuf cgroups(ctx, nseg * 8 + 64), sbase(ctx, (nseg + 1) * 8 + 64);
  DevBuf stmp(ctx, qmk_exscan_tmp_words(nseg) * 8 + 64);
  if (!padded.empty()) h2d(ctx, dbounds.p, padded.data(), padded.size());
  h2d(ctx, dqe.p, qe.data(), n);
  zero(ctx, selected.p, n * 8 + 64);

  const BItem &i0 = *L.sorted[0];
  const DQuery &q0 = i0.ab.q;
  RScan rs;
  memset(&rs, 0, sizeof rs);
  rs.nrows = rows;
  rs.chunk_rows = chunk_rows;
  rs.nchunks = nchunks;
  rs.nq = (int)n;
  rs.key = L.sorted_key->d;
  rs.key_dt = L.sorted_key->dtype;
  rs.pcol = rp.pcol->d;
  rs.p_dt = rp.pcol->dtype;
  rs.pnull = rp.pcol->st.nulls ? rp.pcol->dnull : nullptr;
  rs.nstates = 1 + rp.nsum + rp.nmax;
  rs.nsum = rp.nsum;
  for (int i = 0; i < rs.nstates; ++i) {
    const int s = rp.perm[i];
    rs.st[i] = q0.states[s];
    const TCol *t = i0.st_t[s];
    rs.scol[i] = t ? t->d : nullptr;
    rs.snull[i] = (t && t->st.nulls) ? t->dnull : nullptr;
    rs.s_dt[i] = t ? t->dtype : 0;
  }
  rs.nbounds = (int)nb;
  rs.bsteps = steps;
  rs.bounds = dbounds.as<uint64_t>();
  rs.qe = dqe.as<uint32_t>();
  rs.counts = counts.as<uint32_t>();
  rs.offsets = offs.as<uint64_t>();
  rs.totals = totals.as<uint64_t>();
  rs.nouts = rp.nouts;
  for (int j = 0; j < rp.nouts; ++j) rs.outs[j] = rp.outs[j];
  rs.selected = selected.as<uint64_t>();
  rs.nqb = (int)nqb;
  rs.qbw = qbw;
  rs.chunk_groups = cgroups.as<uint64_t>();
  rs.stg_base = sbase.as<uint64_t>();

  std::vector<uint64_t> tot(n + 1, 0);  // per query, then all staged records
  if (rows) {
    ev_record(ctx, ctx->lane().evk0);
    HIPX(qmk_runs_count(&rs, ctx->lane().stream));
    HIPX(qmk_runs_offsets(rs.counts, nchunks, (uint32_t)n, offs.as<uint64_t>(), totals.as<uint64_t>(), ctx->lane().stream));
    HIPX(qmk_exscan_u64(cgroups.as<uint64_t>(), nseg, sbase.as<uint64_t>(), totals.as<uint64_t>() + n,
                        stmp.as<uint64_t>(), ctx->lane().stream));
    ev_record(ctx, ctx->lane().evk1);
    d2h(ctx, tot.data(), totals.p, n + 1);
    float ms = 0;
    ev_ms(ctx, &ms, ctx->lane().evk0, ctx->lane().evk1);
    scan_ms = std::max(scan_ms, ms);
  }
  // stg_base[nseg] = all records (the exclusive scan leaves it to the total)
  if (rows) h2d(ctx, sbase.as<uint64_t>() + nseg, &tot[n], 1);
  // result columns: one arena per output column shared by the launch's
  // results (query i's rows at row_off[i], padded like any result column);
  // the write pass fills them
  std::vector<uint64_t> roff(n);
  uint64_t arows = 0;
  for (size_t i = 0; i < n; ++i) {
    roff[i] = arows;
    arows += pad_rows(tot[i]);
  }
  std::vector<std::shared_ptr<DevBuf>> arenas(rp.nouts);
  for (int j = 0; j < rp.nouts; ++j)
    arenas[j] = std::make_shared<DevBuf>(ctx, std::max<uint64_t>(arows, 1) * dt_size(rp.outs[j].out_dtype) + 64);
  // NULL bitmaps likewise: one per nullable output column over the arena's
  // rows (bit g for arena row g; row_off[i] is a multiple of 8192, so query
  // i's bitmap starts at word row_off[i] / 32). An emitted group has
  // COUNT(*) >= 1: only a COUNT(col)-guarded output can be NULL.
  std::vector<uint32_t *> nbase(rp.nouts, nullptr);
  for (int j = 0; j < rp.nouts; ++j) {
    const ROut &r = rp.outs[j];
    if (r.kind == OK_KEY || r.kind == OK_COUNT || r.cnt_state <= 0) continue;
    const uint64_t nb = std::max<uint64_t>(arows, 1) / 8 + 64;
    arenas.push_back(std::make_shared<DevBuf>(ctx, nb));
    zero(ctx, arenas.back()->p, nb);
    nbase[j] = arenas.back()->as<uint32_t>();
  }
  for (size_t i = 0; i < n; ++i) {
    const BItem &it = *L.sorted[i];
    qm_result *res = new qm_result();
    out[it.idx] = res;
    res->ctx = ctx;
    res->rows = tot[i];
    res->arenas = arenas;
    for (int j = 0; j < rp.nouts; ++j) {
      const ROut &r = rp.outs[j];
      OutCol oc;
      oc.name = it.ab.outs[j].name;
      oc.dtype = r.out_dtype;
      // views: ctx stays null, nothing to release
      oc.data.p = static_cast<uint8_t *>(arenas[j]->p) + roff[i] * dt_size(r.out_dtype);
      oc.data.n = pad_rows(tot[i]) * dt_size(r.out_dtype);
      if (nbase[j]) {
        oc.nulls.p = reinterpret_cast<uint8_t *>(nbase[j]) + roff[i] / 8;
        oc.nulls.n = pad_rows(tot[i]) / 8 + 64;
      }
      res->cols.push_back(std::move(oc));
    }
  }
  const uint64_t all = tot[n];
  uint64_t check = 0;
  for (size_t i = 0; i < n; ++i) check += tot[i];
  if (check != all) fail(QM_ERR_QUERY_FAILED, "run-sorted batch: %llu staged groups, %llu result rows",
                         (unsigned long long)all, (unsigned long long)check);
  std::vector<uint64_t> sel(n, 0);
  if (all) {
    DevBuf droff(ctx, n * 8 + 64);
    h2d(ctx, droff.p, roff.data(), n);
    for (int j = 0; j < rp.nouts; ++j) {
      rs.out_base[j] = arenas[j]->p;
      rs.out_nbase[j] = nbase[j];
    }
    rs.row_off = droff.as<uint64_t>();
    // staged widths: counts are below 2^32 (rows < 2^32); the order code of a
    // value of <= 32 bits (or its complement, for MIN) lies in
    // [2^63 - 2^31, 2^63 + 2^31); a key whose range spans < 2^32 values is
    // stored as an offset from its minimum
    const TCol *kt = L.sorted_key;
    rs.key_w = 8;
    rs.kbase = 0;
    if (kt->st.has_int && (uint64_t)kt->st.imax - (uint64_t)kt->st.imin < (1ull << 32)) {
      rs.key_w = 4;
      rs.kbase = kt->st.imin;
    }
    for (int j = 0; j < rs.nstates; ++j) {
      const DState &d = rs.st[j];
      rs.val_w[j] = 8;
      rs.vbase[j] = 0;
      if (d.kind == SK_COUNT) {
        rs.val_w[j] = 4;
      } else if (d.kind == SK_MAX_CODE && (rs.s_dt[j] == DT_I8 || rs.s_dt[j] == DT_I16 || rs.s_dt[j] == DT_I32 ||
                                           rs.s_dt[j] == DT_BOOL)) {
        rs.val_w[j] = 4;
        rs.vbase[j] = (1ull << 63) - (1ull << 31);
      }
    }
    DevBuf sq(ctx, all * 2 + 64), skey(ctx, all * rs.key_w + 64), spos(ctx, all * 4 + 64);
    std::vector<DevBuf> sval;
    for (int j = 0; j < rs.nstates; ++j) sval.emplace_back(ctx, all * rs.val_w[j] + 64);
    rs.stg_q = sq.as<uint16_t>();
    rs.stg_key = skey.p;
    rs.stg_pos = spos.as<uint32_t>();
    rs.stg_total = all;
    for (int j = 0; j < rs.nstates; ++j) rs.stg_val[j] = sval[j].p;
    ev_record(ctx, ctx->lane().evk0);
    HIPX(qmk_runs_write(&rs, (uint32_t)ctx->tune.runs_lds_pad, ctx->lane().stream));
    ev_record(ctx, ctx->lane().evk1);
    HIPX(qmk_runs_place(&rs, ctx->tune.runs_win, ctx->lane().stream));
    ev_record(ctx, ctx->lane().evk2);
    d2h(ctx, sel.data(), selected.p, n);
    float mw = 0;
    ev_ms(ctx, &mw, ctx->lane().evk0, ctx->lane().evk1);
    scan_ms = std::max(scan_ms, mw);  // the place pass reads no table column
  }
  DQuery noq;
  memset(&noq, 0, sizeof noq);
  DevBuf dnoq(ctx, sizeof(DQuery));
  h2d(ctx, dnoq.p, reinterpret_cast<const uint8_t *>(&noq), sizeof noq);
  for (size_t i = 0; i < n; ++i) {
    qm_result *res = out[L.sorted[i]->idx];
    res->stats.rows_selected = sel[i];
    res->stats.strategy = QM_STRAT_SORTED_RUNS;
    uint32_t l = 0;
    sort_limit_result(res, *L.sorted[i]->plan, dnoq.as<DQuery>(), &l);
  }
}

// ---- partitioned-runs launches: queries grouped by one integer key stored in
// any row order (k_pruns.hip) ----

// Slot table of sorted, distinct bounds for bucket_slot(): at most PR_BSLOTS
// equal slots of [b.front(), b.back()]; ts[k] = the bounds below slot k's first
// code, ts[slots] = all of them. Returns the slot count (0: no bounds).
uint32_t bucket_slots(const std::vector<uint64_t> &b, std::vector<uint16_t> &ts, uint64_t &lo, int32_t &shift) {
  ts.clear();
  lo = 0;
  shift = 0;
  if (b.empty() || b.size() > 0xFFFF) return 0;
  lo = b.front();
  const uint64_t span = b.back() - b.front();
  while ((span >> shift) >= PR_BSLOTS) ++shift;
  const uint32_t slots = (uint32_t)(span >> shift) + 1;
  ts.resize(slots + 1);
  for (uint32_t k = 0; k < slots; ++k)
    ts[k] = (uint16_t)(std::lower_bound(b.begin(), b.end(), lo + ((uint64_t)k << shift)) - b.begin());
  ts[slots] = (uint16_t)b.size();
  return slots;
}

// Partition geometry for a launch of `nbk` buckets and `nsn` non-COUNT(*)
// states, from the key histogram: final partitions of 2^s keys (the largest
// s, at most the pr_max_s knob, whose fullest partition fits the write pass's
// LDS), level-1 partitions of 2^(sh1 - s) final ones (sh1 - s from the pr_b2
// knob when set). Cached on the key column per knob setting; nullptr when no
// s fits.
const TCol::PRGeo *pr_geometry(qm_ctx *ctx, const TCol *kt_c, uint64_t rows, uint32_t nbk, int nsum, int nmax,
                              uint32_t rw) {
  TCol *kt = const_cast<TCol *>(kt_c);  // the cache, guarded by ctx->cache_mu
  std::lock_guard<std::mutex> cg(ctx->cache_mu);
  const uint32_t nsn = (uint32_t)(nsum + nmax);
  const int max_s = std::min(PR_MAX_S, std::max(1, ctx->tune.pr_max_s));
  const int want_b2 = ctx->tune.pr_b2;
  for (auto &g : kt->prgeo)
    if (g.nbk == nbk && g.nsn == nsn && g.rw == rw && g.max_s == max_s && g.want_b2 == want_b2)
      return g.s >= 0 ? &g : nullptr;
  TCol::PRGeo g;
  g.nbk = nbk;
  g.nsn = nsn;
  g.rw = rw;
  g.max_s = max_s;
  g.want_b2 = want_b2;
  g.tile = qmk_pr_tile((int)rw);
  const std::vector<uint32_t> &h = kt->khist;
  const int kg = kt->khist_g;
  const uint64_t D = (uint64_t)kt->st.imax - (uint64_t)kt->st.imin + 1;
  std::vector<uint64_t> caps;
  for (int s = max_s; s >= kg && g.tile; --s) {
    const uint32_t cap = qmk_pr_max_cap(nbk << s, nsum, nmax);
    if (!cap) continue;
    const uint64_t per = 1ull << (s - kg);
    const uint64_t np = (D + (1ull << s) - 1) >> s;
    // sparse keys (few rows per partition of 64 keys): the per-query paths serve them better
    if (np > std::max<uint64_t>(4096, rows / 32)) break;
    caps.assign(np, 0);
    uint64_t mx = 0;
    for (size_t b = 0; b < h.size(); ++b) caps[b / per] += h[b];
    for (uint64_t c : caps) mx = std::max(mx, c);
    if (mx > cap) continue;
    int lg = 0;
    while ((1ull << lg) < np) ++lg;
    const int b2max = std::min(10, 16 - s);  // 2^b2 <= PR_MAX_P digits in pass 2
    int b2 = std::min(want_b2 > 0 ? want_b2 : (lg + 1) / 2, b2max);
    while (((D + (1ull << (s + b2)) - 1) >> (s + b2)) > (uint64_t)PR_MAX_P && b2 < b2max) ++b2;
    const uint64_t np1 = (D + (1ull << (s + b2)) - 1) >> (s + b2);
    if (np1 > (uint64_t)PR_MAX_P || np >= (1ull << 31)) continue;
    g.s = s;
    g.sh1 = s + b2;
    g.np = (uint32_t)np;
    g.np1 = (uint32_t)np1;
    g.cap = cap;
    break;
  }
  if (g.s >= 0) {
    std::vector<uint64_t> st2(g.np + 1, 0), st1(g.np1 + 1, 0);
    for (uint32_t P = 0; P < g.np; ++P) st2[P + 1] = st2[P] + caps[P];
    const int b2 = g.sh1 - g.s;
    std::vector<uint32_t> ch(g.np1 + 1, 0);
    for (uint32_t p = 0; p <= g.np1; ++p) st1[p] = st2[std::min<uint64_t>((uint64_t)p << b2, g.np)];
    for (uint32_t p = 0; p < g.np1; ++p) ch[p + 1] = ch[p] + (uint32_t)((st1[p + 1] - st1[p] + g.tile - 1) / g.tile);
    g.total = st2[g.np];
    g.nchunk2 = ch[g.np1];
    g.start1 = std::make_shared<DevBuf>(ctx, st1.size() * 8 + 64);
    g.start2 = std::make_shared<DevBuf>(ctx, st2.size() * 8 + 64);
    g.chunk0 = std::make_shared<DevBuf>(ctx, ch.size() * 4 + 64);
    h2d(ctx, g.start1->p, st1.data(), st1.size());
    h2d(ctx, g.start2->p, st2.data(), st2.size());
    h2d(ctx, g.chunk0->p, ch.data(), ch.size());
    // calls on other lanes read these buffers once the entry is published
    stream_sync(ctx);
  }
  kt->prgeo.push_back(g);
  return g.s >= 0 ? &kt->prgeo.back() : nullptr;
}

// One launch through the partitioned-runs passes; false (nothing run) when
// its records or partitions do not fit the kernels.
bool run_pruns(qm_ctx *ctx, const qm_table *T, const BLaunch &L, const RunsPlan &rp, std::vector<qm_result *> &out,
               float &scan_ms, double pass_ms[4], const uint32_t *page_skip) {
  const size_t n = L.sorted.size();
  const std::vector<uint64_t> &b = rp.bounds;
  const uint32_t nb = (uint32_t)b.size();
  int steps = 0;
  while (((1u << steps) - 1) < nb) ++steps;
  if (steps == 0) steps = 1;  // one padded bound: the search loop runs once
  std::vector<uint64_t> padded((1u << steps) - 1, ~0ull);
  std::copy(b.begin(), b.end(), padded.begin());
  std::vector<uint32_t> qe(n);
  uint32_t on[R_MAX_BUCKETS / 32] = {};
  for (size_t i = 0; i < n; ++i) {
    const BItem &it = *L.sorted[i];
    uint32_t lo, hi;
    if (it.never) {
      lo = hi = 0;
    } else if (it.nslots == 0) {
      lo = 0;
      hi = nb + 2;  // every row, NULL filter values included
    } else {
      lo = (uint32_t)(std::upper_bound(b.begin(), b.end(), it.lo[0]) - b.begin());
      hi = it.hi[0] == ~0ull ? nb + 1 : (uint32_t)(std::upper_bound(b.begin(), b.end(), it.hi[0] + 1) - b.begin());
    }
    qe[i] = lo | (hi << 16);
    for (uint32_t e = lo; e < hi; ++e) on[e >> 5] |= 1u << (e & 31);
  }
  // record layout: key / bucket word, the distinct state columns, NULL bits
  const BItem &i0 = *L.sorted[0];
  const DQuery &q0 = i0.ab.q;
  PRScan ps;
  memset(&ps, 0, sizeof ps);
  std::vector<const TCol *> vc;
  int rw = 1;
  bool nullable = false;
  ps.nstates = 1 + rp.nsum + rp.nmax;
  ps.nsum = rp.nsum;
  for (int i = 0; i < ps.nstates; ++i) {
    const int s = rp.perm[i];
    ps.st[i] = q0.states[s];
    const TCol *t = i0.st_t[s];
    ps.st_v[i] = -1;
    if (!t || i == 0) continue;
    auto f = std::find(vc.begin(), vc.end(), t);
    if (f == vc.end()) {
      if (vc.size() >= (size_t)PR_MAX_V) return false;
      vc.push_back(t);
      f = vc.end() - 1;
    }
    ps.st_v[i] = (int8_t)(f - vc.begin());
  }
  ps.nv = (int)vc.size();
  for (size_t j = 0; j < vc.size(); ++j) {
    ps.vcol[j] = vc[j]->d;
    ps.vnull[j] = vc[j]->st.nulls ? vc[j]->dnull : nullptr;
    ps.v_dt[j] = vc[j]->dtype;
    ps.v_word[j] = rw;
    rw += (vc[j]->dtype == DT_I64 || vc[j]->dtype == DT_F64) ? 2 : 1;
    nullable = nullable || vc[j]->st.nulls;
  }
  ps.null_word = -1;
  if (nullable) ps.null_word = rw++;
  rw = std::max(rw, 2);
  if (rw > PR_MAX_RW) return false;
  const TCol *kt = L.sorted_key;
  const TCol::PRGeo *g = pr_geometry(ctx, kt, T->rows, nb + 2, rp.nsum, rp.nmax, (uint32_t)rw);
  if (!g) return false;

  const uint64_t rows = T->rows;
  ps.nrows = rows;
  ps.key = kt->d;
  ps.key_dt = kt->dtype;
  ps.pcol = rp.pcol->d;
  ps.p_dt = rp.pcol->dtype;
  ps.pnull = rp.pcol->st.nulls ? rp.pcol->dnull : nullptr;
  ps.kmin = kt->st.imin;
  ps.rw = rw;
  ps.nbounds = (int)nb;
  ps.bsteps = steps;
  memcpy(ps.bucket_on, on, sizeof on);
  ps.s = g->s;
  ps.sh1 = g->sh1;
  ps.np1 = g->np1;
  ps.np = g->np;
  ps.tile = g->tile;
  ps.start1 = g->start1->as<uint64_t>();
  ps.start2 = g->start2->as<uint64_t>();
  ps.chunk0 = g->chunk0->as<uint32_t>();
  ps.cap = g->cap;
  ps.nq = (int)n;
  ps.nouts = rp.nouts;
  for (int j = 0; j < rp.nouts; ++j) ps.outs[j] = rp.outs[j];

  DevBuf dbounds(ctx, padded.size() * 8 + 64), dqe(ctx, n * 4 + 64);
  h2d(ctx, dbounds.p, padded.data(), padded.size());
  h2d(ctx, dqe.p, qe.data(), n);
  ps.bounds = dbounds.as<uint64_t>();
  // the row scatter's bucket lookup (bucket_slot)
  std::vector<uint16_t> bslot;
  ps.bslots = bucket_slots(b, bslot, ps.bslot_lo, ps.bslot_shift);
  DevBuf dbslot(ctx, bslot.size() * 2 + 64);
  if (ps.bslots) {
    h2d(ctx, dbslot.p, bslot.data(), bslot.size());
    ps.bslot = dbslot.as<uint16_t>();
  }
  ps.qe = dqe.as<uint32_t>();
  const uint64_t recb = g->total * (uint64_t)rw * 4 + 64;
  DevBuf rec1(ctx, recb), rec2(ctx, recb);
  DevBuf cur1(ctx, (uint64_t)g->np1 * 4 + 64), cur2(ctx, (uint64_t)g->np * 4 + 64);
  DevBuf brows(ctx, R_MAX_BUCKETS * 8 + 64), status(ctx, 64);
  const uint64_t ncnt = (uint64_t)n * g->np;
  DevBuf gcount(ctx, ncnt * 8 + 64), gtotal(ctx, 64), stmp(ctx, qmk_exscan_tmp_words(ncnt) * 8 + 64);
  zero(ctx, cur1.p, (uint64_t)g->np1 * 4 + 64);
  zero(ctx, cur2.p, (uint64_t)g->np * 4 + 64);
  zero(ctx, brows.p, R_MAX_BUCKETS * 8 + 64);
  zero(ctx, status.p, 64);
  ps.rec1 = rec1.as<uint32_t>();
  ps.rec2 = rec2.as<uint32_t>();
  ps.cur1 = cur1.as<uint32_t>();
  ps.cur2 = cur2.as<uint32_t>();
  ps.bucket_rows = brows.as<uint64_t>();
  ps.status = status.as<uint32_t>();
  ps.gcount = gcount.as<uint64_t>();
  ps.page_skip = page_skip;  // the row scatter makes no records of these pages' rows

  // persistent scatter grids: the workgroups that fit the device at once
  const uint64_t tile = g->tile;
  const uint64_t ntiles = (rows + tile - 1) / tile;
  const uint64_t pgrid = (uint64_t)ctx->num_cus * (uint64_t)std::max(1, ctx->tune.pr_scatter_bpc);
  hipEvent_t ev[6];
  for (auto &e : ev) HIPX(hipEventCreate(&e));
  struct EvFree { hipEvent_t *e; ~EvFree() { for (int i = 0; i < 6; ++i) (void)hipEventDestroy(e[i]); } } evfree{ev};
  ev_record(ctx, ev[0]);
  if (ntiles) HIPX(qmk_pr_scatter(&ps, 0, (uint32_t)std::min(ntiles, pgrid), ctx->lane().stream));
  ev_record(ctx, ev[1]);
  if (g->nchunk2) HIPX(qmk_pr_scatter(&ps, 1, (uint32_t)std::min<uint64_t>(g->nchunk2, pgrid), ctx->lane().stream));
  ev_record(ctx, ev[2]);
  HIPX(qmk_pr_count(&ps, ctx->lane().stream));
  HIPX(qmk_exscan_u64(gcount.as<uint64_t>(), ncnt, gcount.as<uint64_t>(), gtotal.as<uint64_t>(),
                      stmp.as<uint64_t>(), ctx->lane().stream));
  ev_record(ctx, ev[3]);
  std::vector<uint64_t> base(n + 1, 0);
  d2h_rows(ctx, base.data(), gcount.p, 8, (size_t)g->np * 8, n);
  d2h(ctx, &base[n], gtotal.p, 1);
  std::vector<uint64_t> brow(R_MAX_BUCKETS);
  d2h(ctx, brow.data(), brows.p, R_MAX_BUCKETS);
  uint32_t stv = 0;
  d2h(ctx, &stv, status.p, 1);
  if (stv) fail(QM_ERR_QUERY_FAILED, "partitioned-runs batch: a key partition exceeded its region (status %u)", stv);
  std::vector<uint64_t> tot(n);
  for (size_t i = 0; i < n; ++i) tot[i] = base[i + 1] - base[i];
  uint64_t records = 0;
  for (uint32_t e = 0; e < R_MAX_BUCKETS; ++e) records += brow[e];

  // result arenas as in run_runs: query i's rows at a padded offset
  std::vector<uint64_t> roff(n);
  uint64_t arows = 0;
  for (size_t i = 0; i < n; ++i) {
    roff[i] = arows;
    arows += pad_rows(tot[i]);
  }
  std::vector<std::shared_ptr<DevBuf>> arenas(rp.nouts);
  for (int j = 0; j < rp.nouts; ++j)
    arenas[j] = std::make_shared<DevBuf>(ctx, std::max<uint64_t>(arows, 1) * dt_size(rp.outs[j].out_dtype) + 64);
  std::vector<uint32_t *> nbase(rp.nouts, nullptr);
  for (int j = 0; j < rp.nouts; ++j) {
    const ROut &r = rp.outs[j];
    if (r.kind == OK_KEY || r.kind == OK_COUNT || r.cnt_state <= 0) continue;
    const uint64_t nbb = std::max<uint64_t>(arows, 1) / 8 + 64;
    arenas.push_back(std::make_shared<DevBuf>(ctx, nbb));
    zero(ctx, arenas.back()->p, nbb);
    nbase[j] = arenas.back()->as<uint32_t>();
  }
  for (size_t i = 0; i < n; ++i) {
    const BItem &it = *L.sorted[i];
    qm_result *res = new qm_result();
    out[it.idx] = res;
    res->ctx = ctx;
    res->rows = tot[i];
    res->arenas = arenas;
    for (int j = 0; j < rp.nouts; ++j) {
      const ROut &r = rp.outs[j];
      OutCol oc;
      oc.name = it.ab.outs[j].name;
      oc.dtype = r.out_dtype;
      oc.data.p = static_cast<uint8_t *>(arenas[j]->p) + roff[i] * dt_size(r.out_dtype);
      oc.data.n = pad_rows(tot[i]) * dt_size(r.out_dtype);
      if (nbase[j]) {
        oc.nulls.p = reinterpret_cast<uint8_t *>(nbase[j]) + roff[i] / 8;
        oc.nulls.n = pad_rows(tot[i]) / 8 + 64;
      }
      res->cols.push_back(std::move(oc));
    }
  }
  DevBuf droff(ctx, n * 8 + 64);
  h2d(ctx, droff.p, roff.data(), n);
  for (int j = 0; j < rp.nouts; ++j) {
    ps.out_base[j] = arenas[j]->p;
    ps.out_nbase[j] = nbase[j];
  }
  ps.row_off = droff.as<uint64_t>();
  ev_record(ctx, ev[4]);
  if (base[n]) HIPX(qmk_pr_runs(&ps, ctx->lane().stream));
  ev_record(ctx, ev[5]);
  ev_sync(ctx, ev[5]);
  d2h(ctx, &stv, status.p, 1);
  if (stv) fail(QM_ERR_QUERY_FAILED, "partitioned-runs batch: a key partition exceeded the LDS capacity (status %u)", stv);
  float ms[4] = {0, 0, 0, 0};
  ev_ms(ctx, &ms[0], ev[0], ev[1]);
  ev_ms(ctx, &ms[1], ev[1], ev[2]);
  ev_ms(ctx, &ms[2], ev[2], ev[3]);
  ev_ms(ctx, &ms[3], ev[4], ev[5]);
  for (int i = 0; i < 4; ++i) pass_ms[i] = ms[i];
  scan_ms = std::max(scan_ms, ms[0]);  // the pass reading the table
  DQuery noq;
  memset(&noq, 0, sizeof noq);
  DevBuf dnoq(ctx, sizeof(DQuery));
  h2d(ctx, dnoq.p, reinterpret_cast<const uint8_t *>(&noq), sizeof noq);
  for (size_t i = 0; i < n; ++i) {
    qm_result *res = out[L.sorted[i]->idx];
    const uint32_t lo = qe[i] & 0xFFFFu, hi = qe[i] >> 16;
    uint64_t sel = 0;
    for (uint32_t e = lo; e < hi; ++e) sel += brow[e];
    res->stats.rows_selected = sel;
    res->stats.strategy = QM_STRAT_PART_RUNS;
    for (int k = 0; k < 4; ++k) res->stats.pass_ms[k] = pass_ms[k];
    res->stats.pass_records = records;
    uint32_t l = 0;
    sort_limit_result(res, *L.sorted[i]->plan, dnoq.as<DQuery>(), &l);
  }
  return true;
}

// One shared-interval group through the lean pass: bounds padded for the
// branch-free search, the group's distinct value columns (the range column
// first), one workgroup range of rows per block.
void run_shared_lean(qm_ctx *ctx, const qm_table *T, BGroup &g, float &scan_ms, const uint32_t *page_skip) {
  SScan ss;
  memset(&ss, 0, sizeof ss);
  ss.nrows = T->rows;
  ss.page_skip = page_skip;
  std::vector<const TCol *> vc{g.col};
  for (const TCol *t : g.st_t)
    if (t && std::find(vc.begin(), vc.end(), t) == vc.end()) vc.push_back(t);
  ss.nv = (int)vc.size();
  for (size_t j = 0; j < vc.size(); ++j) {
    ss.vcol[j] = vc[j]->d;
    ss.vnull[j] = vc[j]->st.nulls ? vc[j]->dnull : nullptr;
    ss.v_dt[j] = vc[j]->dtype;
  }
  if (g.key) {
    ss.key = g.key->d;
    ss.key_dt = g.key->dtype;
  }
  ss.dom = g.dom;
  ss.kmin = g.kmin;
  ss.nstates = (int)g.st.size();
  for (size_t i = 0; i < g.st.size(); ++i) {
    ss.st[i] = g.st[i];
    ss.st_v[i] = g.st_t[i] ? (int8_t)(std::find(vc.begin(), vc.end(), g.st_t[i]) - vc.begin()) : (int8_t)-1;
  }
  ss.entry_words = g.entry_words;
  const uint32_t nb = (uint32_t)g.bounds.size();
  int steps = 0;
  while (((1u << steps) - 1) < nb) ++steps;
  std::vector<uint64_t> padded((1u << steps) - 1, ~0ull);
  std::copy(g.bounds.begin(), g.bounds.end(), padded.begin());
  DevBuf db(ctx, std::max<size_t>(padded.size(), 1) * 8);
  if (!padded.empty()) h2d(ctx, db.p, padded.data(), padded.size());
  ss.nbounds = (int)nb;
  ss.bsteps = steps;
  ss.bounds = db.as<uint64_t>();
  ss.table = g.table.as<uint64_t>();
  const int bpc = qmk_shared_blocks_per_cu(&ss);
  uint64_t grid = (uint64_t)ctx->num_cus * (uint64_t)std::max(1, bpc) * (uint64_t)ctx->tune.shared_waves;
  uint64_t rpb = (T->rows + grid - 1) / grid;
  rpb = std::max<uint64_t>(S_BLOCK, (rpb + S_BLOCK - 1) / S_BLOCK * S_BLOCK);
  grid = (T->rows + rpb - 1) / rpb;
  ss.rows_per_block = rpb;
  ev_record(ctx, ctx->lane().evk0);
  HIPX(qmk_shared_scan(&ss, (uint32_t)grid, ctx->lane().stream));
  ev_record(ctx, ctx->lane().evk1);
  ev_sync(ctx, ctx->lane().evk1);
  float ms = 0;
  ev_ms(ctx, &ms, ctx->lane().evk0, ctx->lane().evk1);
  scan_ms = std::max(scan_ms, ms);
}

// Run the fusable plans of `items` in fused passes; results by item idx.
// BK_RUNS items the run-sorted kernels cannot take are left to the caller:
// their idx go to `rejected` (their out[] entries stay null).
void run_batch(qm_ctx *ctx, const qm_table *T, std::vector<BItem *> &items, std::vector<qm_result *> &out,
               std::vector<uint32_t> *rejected, const uint32_t *page_skip = nullptr) {
  Timer tm(ctx);
  // shared groups
  std::vector<std::unique_ptr<BGroup>> groups;
  std::vector<BItem *> sorted, sparse, pruns;
  for (BItem *it : items) {
    if (it->kind == BK_SORTED) { sorted.push_back(it); continue; }
    if (it->kind == BK_RUNS) { sparse.push_back(it); continue; }
    if (it->kind == BK_PRUNS) { pruns.push_back(it); continue; }
    BGroup *g = nullptr;
    for (auto &gg : groups)
      if (gg->col == it->slot_t[0] && gg->key == it->key_t && gg->items.size() < 4 * B_MAX_BOUNDS / 8) g = gg.get();
    if (!g) {
      groups.emplace_back(new BGroup());
      g = groups.back().get();
      g->col = it->slot_t[0];
      g->key = it->key_t;
      DState c0;
      memset(&c0, 0, sizeof c0);
      c0.kind = SK_COUNT;
      c0.col_a = -1;
      c0.word = 0;
      g->st.push_back(c0);
      g->st_t.push_back(nullptr);
      g->entry_words = 1;
      if (it->key_t) { g->dom = (uint32_t)it->ab.q.key.dom[0]; g->kmin = it->ab.q.key.kmin[0]; }
    }
    for (int s = 0; s < it->ab.q.nstates; ++s) group_state(*g, *it, s);
    g->items.push_back(it);
    it->group = (int)(std::find_if(groups.begin(), groups.end(), [&](auto &p) { return p.get() == g; }) - groups.begin());
  }
  for (auto &g : groups) {
    std::vector<uint64_t> b;
    for (BItem *it : g->items) {
      if (it->never) continue;
      b.push_back(it->lo[0]);
      if (it->hi[0] != ~0ull) b.push_back(it->hi[0] + 1);
    }
    std::sort(b.begin(), b.end());
    b.erase(std::unique(b.begin(), b.end()), b.end());
    if (b.size() > (size_t)B_MAX_BOUNDS) fail(QM_ERR_UNSUPPORTED, "too many range bounds in one shared group");
    g->bounds = b;
    for (BItem *it : g->items) {
      if (it->never) { it->e_lo = it->e_hi = 0; continue; }
      it->e_lo = (uint64_t)(std::upper_bound(b.begin(), b.end(), it->lo[0]) - b.begin());
      it->e_hi = it->hi[0] == ~0ull ? b.size() + 1
                                    : (uint64_t)(std::upper_bound(b.begin(), b.end(), it->hi[0] + 1) - b.begin());
    }
  }
  // groups whose value columns fit the lean pass run it (one group per pass)
  for (auto &g : groups) {
    std::vector<const TCol *> vc{g->col};
    for (const TCol *t : g->st_t)
      if (t && std::find(vc.begin(), vc.end(), t) == vc.end()) vc.push_back(t);
    int steps = 0;
    while (((1u << steps) - 1) < g->bounds.size()) ++steps;
    const uint64_t lds = (((1ull << steps) - 1) + 1 + (g->bounds.size() + 1) * g->dom * (uint64_t)g->entry_words) * 8;
    g->lean = ctx->tune.shared_lean && vc.size() <= (size_t)S_MAX_V && g->st.size() <= (size_t)B_MAX_SHARED_STATES &&
              steps <= 10 && lds <= 96 * 1024;
  }
  // launches: shared groups, then sorted-key queries (one key column per launch)
  std::vector<BLaunch> launches;
  auto lds_of = [&](const BLaunch &L, const BGroup *extra_g, size_t extra_sorted) {
    uint64_t w = 0;
    for (const BGroup *g : L.groups) w += group_lds_words(*g);
    if (extra_g) w += group_lds_words(*extra_g);
    w += (uint64_t)B_WAVES * (L.sorted.size() + extra_sorted) * qmk_batch_carry_words();
    w += (L.sorted.size() + extra_sorted) * sizeof(BQuery) / 8;
    return w;
  };
  for (auto &g : groups) {
    if (g->lean) continue;
    std::vector<const TCol *> need{g->col, g->key};
    for (const TCol *t : g->st_t) need.push_back(t);
    BLaunch *L = nullptr;
    for (auto &l : launches)
      if (l.groups.size() < (size_t)B_MAX_SHARED && l.fits(need) && lds_of(l, g.get(), 0) <= B_LDS_WORDS) L = &l;
    if (!L) { launches.emplace_back(); L = &launches.back(); }
    if (lds_of(*L, g.get(), 0) > B_LDS_WORDS || !L->fits(need)) fail(QM_ERR_UNSUPPORTED, "shared group too large");
    L->add(need);
    L->groups.push_back(g.get());
  }
  for (BItem *it : sorted) {
    std::vector<const TCol *> need = item_cols(*it);
    BLaunch *L = nullptr;
    for (auto &l : launches)
      if (l.groups.empty() && (!l.sorted_key || l.sorted_key == it->key_t) && l.fits(need) &&
          l.sorted.size() < 256 && same_signature(*l.sorted[0], *it) && it->nslots <= B_QP_SLOTS &&
          l.sorted[0]->nslots <= B_QP_SLOTS)
        L = &l;
    if (!L) {
      launches.emplace_back();
      L = &launches.back();
      L->cols.push_back(it->key_t);  // the sorted key is batch column 0
    }
    L->add(need);
    L->sorted_key = it->key_t;
    L->sorted.push_back(it);
  }
  // sparse sorted keys: one run-sorted launch per (key, range column, state
  // signature), at most R_MAX_QUERIES queries each
  std::vector<BLaunch> sparse_launches;
  std::vector<const TCol *> sparse_pcol;
  for (BItem *it : sparse) {
    const TCol *pc = (it->nslots == 1 && !it->never) ? it->slot_t[0] : nullptr;
    BLaunch *L = nullptr;
    for (size_t l = 0; l < sparse_launches.size() && !L; ++l) {
      BLaunch &c = sparse_launches[l];
      if (c.sorted_key == it->key_t && c.sorted.size() < (size_t)R_MAX_QUERIES && same_signature(*c.sorted[0], *it) &&
          (!pc || !sparse_pcol[l] || pc == sparse_pcol[l])) {
        L = &c;
        if (pc) sparse_pcol[l] = pc;
      }
    }
    if (!L) {
      sparse_launches.emplace_back();
      sparse_pcol.push_back(pc);
      L = &sparse_launches.back();
      L->sorted_key = it->key_t;
    }
    L->sorted.push_back(it);
  }
  // run-sorted launches write their results directly
  float scan_ms_total = 0;
  std::vector<BItem *> runs_done;
  size_t nlaunches = launches.size();
  for (auto &L : sparse_launches) {
    RunsPlan rp;
    if (runs_plan(ctx, T, L, rp)) {
      run_runs(ctx, T, L, rp, out, scan_ms_total);
      for (BItem *it : L.sorted) runs_done.push_back(it);
      ++nlaunches;
    } else {
      if (!rejected) fail(QM_ERR_QUERY_FAILED, "sparse sorted-key query without a run-sorted launch");
      for (BItem *it : L.sorted) rejected->push_back(it->idx);
    }
  }
  // keys in any row order: one partitioned-runs launch per (key, range
  // column, state signature), at most R_MAX_QUERIES queries each
  {
    std::vector<BLaunch> pl;
    std::vector<const TCol *> pl_pcol;
    for (BItem *it : pruns) {
      const TCol *pc = (it->nslots == 1 && !it->never) ? it->slot_t[0] : nullptr;
      BLaunch *L = nullptr;
      for (size_t l = 0; l < pl.size() && !L; ++l) {
        BLaunch &c = pl[l];
        if (c.sorted_key == it->key_t && c.sorted.size() < (size_t)R_MAX_QUERIES && same_signature(*c.sorted[0], *it) &&
            (!pc || !pl_pcol[l] || pc == pl_pcol[l])) {
          L = &c;
          if (pc) pl_pcol[l] = pc;
        }
      }
      if (!L) {
        pl.emplace_back();
        pl_pcol.push_back(pc);
        L = &pl.back();
        L->sorted_key = it->key_t;
      }
      L->sorted.push_back(it);
    }
    for (auto &L : pl) {
      RunsPlan rp;
      double pms[4] = {0, 0, 0, 0};
      if (runs_plan(ctx, T, L, rp, true) && run_pruns(ctx, T, L, rp, out, scan_ms_total, pms, page_skip)) {
        for (BItem *it : L.sorted) runs_done.push_back(it);
        ++nlaunches;
      } else {
        if (!rejected) fail(QM_ERR_QUERY_FAILED, "key-partitioned query without a launch");
        for (BItem *it : L.sorted) rejected->push_back(it->idx);
      }
    }
  }
  {
    std::vector<BLaunch> rest;
    for (auto &L : launches) {
      RunsPlan rp;
      if (!L.sorted.empty() && runs_plan(ctx, T, L, rp)) {
        run_runs(ctx, T, L, rp, out, scan_ms_total);
        for (BItem *it : L.sorted) runs_done.push_back(it);
      } else {
        rest.push_back(std::move(L));
      }
    }
    launches = std::move(rest);
  }
  // tables, status, selected counters
  const size_t n = items.size();
  DevBuf status(ctx, n * 8 + 64), selected(ctx, n * 8 + 64);
  zero(ctx, status.p, n * 8 + 64);
  zero(ctx, selected.p, n * 8 + 64);
  std::vector<BItem *> order;  // finalisation order = status / selected index
  for (auto &g : groups)  // lean groups run outside the launches
    if (g->lean) for (BItem *it : g->items) order.push_back(it);
  for (auto &L : launches) {
    for (BGroup *g : L.groups) for (BItem *it : g->items) order.push_back(it);
    for (BItem *it : L.sorted) order.push_back(it);
  }
  for (auto &g : groups)
    for (BItem *it : g->items)
      if (std::find(order.begin(), order.end(), it) == order.end())
        fail(QM_ERR_QUERY_FAILED, "batched query %u has no pass", it->idx);
  for (BItem *it : order) {
    it->table = DevBuf(ctx, it->ab.table_words * 8);
    if (it->kind == BK_SORTED) zero(ctx, it->table.p, it->ab.table_words * 8);
  }
  for (auto &g : groups) {
    uint64_t words = (g->bounds.size() + 1) * g->dom * (uint64_t)g->entry_words;
    g->table = DevBuf(ctx, words * 8);
    zero(ctx, g->table.p, words * 8);
    g->dbounds = DevBuf(ctx, std::max<size_t>(g->bounds.size(), 1) * 8);
    h2d(ctx, g->dbounds.p, g->bounds.data(), g->bounds.size());
  }
  for (auto &g : groups)
    if (g->lean && T->rows) run_shared_lean(ctx, T, *g, scan_ms_total, page_skip);
  const uint64_t ntiles = std::max<uint64_t>(1, (T->rows + 511) / 512);
  std::vector<DevBuf> qbufs;
  for (auto &L : launches) {
    BScan bs;
    memset(&bs, 0, sizeof bs);
    bs.nrows = T->rows;
    bs.ncols = (int)L.cols.size();
    for (size_t j = 0; j < L.cols.size(); ++j) {
      bs.data[j] = L.cols[j]->d;
      bs.nulls[j] = L.cols[j]->dnull;
      bs.dtype[j] = L.cols[j]->dtype;
    }
    uint32_t off = 0;
    bs.nshared = (int)L.groups.size();
    for (size_t gi = 0; gi < L.groups.size(); ++gi) {
      BGroup &g = *L.groups[gi];
      BShared &S = bs.sh[gi];
      S.col = (uint8_t)L.col_of(g.col);
      S.key_col = g.key ? (uint8_t)L.col_of(g.key) : 0xFF;
      S.nstates = (uint8_t)g.st.size();
      S.nbounds = (int)g.bounds.size();
      S.kmin = g.kmin;
      S.dom = g.dom;
      S.entry_words = g.entry_words;
      for (size_t s = 0; s < g.st.size(); ++s) {
        S.st[s] = g.st[s];
        S.st_col[s] = g.st_t[s] ? (uint8_t)L.col_of(g.st_t[s]) : 0xFF;
      }
      S.bounds = g.dbounds.as<uint64_t>();
      S.table = g.table.as<uint64_t>();
      S.bnd_off = off;
      off += (uint32_t)g.bounds.size();
      S.lds_off = off;
      off += (uint32_t)((g.bounds.size() + 1) * g.dom * (uint64_t)g.entry_words);
    }
    bs.carry_off = off;
    bs.nq = (int)L.sorted.size();
    off += (uint32_t)(B_WAVES * L.sorted.size() * qmk_batch_carry_words());
    bs.qdesc_off = off;
    off += (uint32_t)(L.sorted.size() * sizeof(BQuery) / 8);
    std::vector<BQuery> hq(L.sorted.size());
    for (size_t i = 0; i < L.sorted.size(); ++i) {
      BItem &it = *L.sorted[i];
      BQuery &Q = hq[i];
      memset(&Q, 0, sizeof Q);
      const DQuery &q = it.ab.q;
      Q.nslots = it.never ? (uint8_t)0xFF : (uint8_t)it.nslots;
      for (int s = 0; s < it.nslots; ++s) {
        Q.slot_col[s] = (uint8_t)L.col_of(it.slot_t[s]);
        Q.lo[s] = it.lo[s];
        Q.span[s] = it.hi[s] - it.lo[s];
      }
      Q.key_col = (uint8_t)L.col_of(it.key_t);
      Q.nstates = (uint8_t)q.nstates;
      for (int s = 0; s < q.nstates; ++s) {
        Q.st[s] = q.states[s];
        Q.st_col[s] = it.st_t[s] ? (uint8_t)L.col_of(it.st_t[s]) : 0xFF;
      }
      Q.kmin = q.key.kmin[0];
      Q.entry_words = q.entry_words;
      Q.table = it.table.as<uint64_t>();
      size_t oi = std::find(order.begin(), order.end(), &it) - order.begin();
      Q.selected = selected.as<uint64_t>() + oi;
    }
    DevBuf dq(ctx, std::max<size_t>(hq.size(), 1) * sizeof(BQuery));
    h2d(ctx, dq.p, reinterpret_cast<const uint8_t *>(hq.data()), hq.size() * sizeof(BQuery));
    const size_t lds_bytes = (size_t)off * 8;
    int wide = 0;
    for (const BQuery &Q : hq)
      for (int s2 = 0; s2 < Q.nstates; ++s2) if (Q.st[s2].kind == SK_SUM_128) wide = 1;
    // one sorted-key query with long runs: the lean single-query pass
    if (L.sorted.size() == 1 && !wide && ctx->tune.sorted1 && L.groups.empty()) {
      BItem &it = *L.sorted[0];
      const DQuery &q = it.ab.q;
      bool ok = q.nstates <= R_MAX_STATES && it.nslots <= B_MAX_SLOTS &&
                (double)T->rows / std::max(1.0, it.key_t->st.distinct) >= ctx->tune.runs_min_len;
      for (int s2 = 0; s2 < q.nstates && ok; ++s2) {
        const DState &d = q.states[s2];
        ok = !d.product && (d.kind == SK_COUNT || d.kind == SK_SUM_I64 || d.kind == SK_MAX_CODE) &&
             (s2 == 0 ? (d.kind == SK_COUNT && d.col_a < 0) : it.st_t[s2] != nullptr);
      }
      if (ok) {
        S1Scan q1;
        memset(&q1, 0, sizeof q1);
        q1.nrows = T->rows;
        const uint64_t target = (uint64_t)ctx->num_cus * (uint64_t)ctx->tune.runs_chunks_per_cu;
        uint64_t cr = std::max<uint64_t>(1024, (T->rows + target - 1) / target);
        cr = (cr + 63) / 64 * 64;
        q1.chunk_rows = cr;
        q1.nchunks = (uint32_t)std::max<uint64_t>(1, (T->rows + cr - 1) / cr);
        q1.nslots = it.never ? -1 : it.nslots;
        q1.key = it.key_t->d;
        q1.key_dt = it.key_t->dtype;
        q1.kmin = q.key.kmin[0];
        for (int j = 0; j < it.nslots; ++j) {
          q1.pcol[j] = it.slot_t[j]->d;
          q1.pnull[j] = it.slot_t[j]->st.nulls ? it.slot_t[j]->dnull : nullptr;
          q1.p_dt[j] = it.slot_t[j]->dtype;
          q1.lo[j] = it.lo[j];
          q1.span[j] = it.hi[j] - it.lo[j];
        }
        q1.nstates = q.nstates;
        for (int s2 = 0; s2 < q.nstates; ++s2) {
          q1.st[s2] = q.states[s2];
          const TCol *t = it.st_t[s2];
          q1.scol[s2] = t ? t->d : nullptr;
          q1.snull[s2] = (t && t->st.nulls) ? t->dnull : nullptr;
          q1.s_dt[s2] = t ? t->dtype : 0;
        }
        q1.entry_words = q.entry_words;
        q1.table = it.table.as<uint64_t>();
        size_t oi = std::find(order.begin(), order.end(), &it) - order.begin();
        q1.selected = selected.as<uint64_t>() + oi;
        if (T->rows) {
          ev_record(ctx, ctx->lane().evk0);
          HIPX(qmk_sorted1_scan(&q1, ctx->lane().stream));
          ev_record(ctx, ctx->lane().evk1);
          ev_sync(ctx, ctx->lane().evk1);
          float ms = 0;
          ev_ms(ctx, &ms, ctx->lane().evk0, ctx->lane().evk1);
          scan_ms_total = std::max(scan_ms_total, ms);
        }
        qbufs.push_back(std::move(dq));
        continue;
      }
    }
    // lean query-parallel launch: many queries sharing one aggregate list
    // (COUNT / integer SUM / MIN / MAX), each filtered by one range on one
    // common column
    bool qp = !L.sorted.empty() && (int)L.sorted.size() >= ctx->tune.qp_min && !wide;
    int pcol = -1;
    for (BItem *it : L.sorted) {
      if (!same_signature(*L.sorted[0], *it) || it->nslots > 1) qp = false;
      if (!it->never && it->nslots == 1) {
        int c = L.col_of(it->slot_t[0]);
        if (pcol >= 0 && c != pcol) qp = false;
        pcol = c;
      }
    }
    std::vector<int> order_st;  // additive states first, then order codes
    if (qp) {
      const DQuery &q0 = L.sorted[0]->ab.q;
      if (q0.nstates < 1 || q0.states[0].kind != SK_COUNT || q0.states[0].col_a >= 0) qp = false;
      for (int s2 = 0; s2 < q0.nstates && qp; ++s2) {
        const uint8_t k = q0.states[s2].kind;
        if (k == SK_COUNT || k == SK_SUM_I64) order_st.push_back(s2);
        else if (k != SK_MAX_CODE) qp = false;
      }
      const int nadd = (int)order_st.size();
      for (int s2 = 0; s2 < q0.nstates; ++s2) if (q0.states[s2].kind == SK_MAX_CODE) order_st.push_back(s2);
      const int nmax = (int)order_st.size() - nadd;
      if (qp && qmk_batch_qpf_supported(nadd, nmax)) {
        bs.sig_nstates = q0.nstates;
        bs.sig_entry_words = q0.entry_words;
        for (size_t i = 0; i < order_st.size(); ++i) {
          const int s2 = order_st[i];
          bs.sig_st[i] = q0.states[s2];
          bs.sig_col[i] = L.sorted[0]->st_t[s2] ? (uint8_t)L.col_of(L.sorted[0]->st_t[s2]) : 0xFF;
        }
        bs.qp_pred_col = pcol < 0 ? 0 : pcol;
        const int qpl = L.sorted.size() <= 64 ? 1 : (L.sorted.size() <= 128 ? 2 : 4);
        const int wpc = qmk_batch_qpf_waves_per_cu(nadd, nmax, qpl);
        uint64_t grid = (uint64_t)ctx->num_cus * std::max(1, wpc / B_WAVES);
        const uint64_t waves = grid * B_WAVES;
        bs.tiles_per_wave = (ntiles + waves - 1) / waves;
        grid = ((ntiles + bs.tiles_per_wave - 1) / bs.tiles_per_wave + B_WAVES - 1) / B_WAVES;
        ev_record(ctx, ctx->lane().evk0);
        HIPX(qmk_batch_qpf_scan(&bs, dq.as<BQuery>(), (uint32_t)grid, nadd, nmax, qpl, ctx->lane().stream));
        ev_record(ctx, ctx->lane().evk1);
        ev_sync(ctx, ctx->lane().evk1);
        float ms = 0;
        ev_ms(ctx, &ms, ctx->lane().evk0, ctx->lane().evk1);
        scan_ms_total = std::max(scan_ms_total, ms);
        qbufs.push_back(std::move(dq));
        continue;
      }
    }
    const int wpc = qmk_batch_waves_per_cu(bs.ncols, wide, bs.nshared > 0 ? 1 : 2, lds_bytes);
    const uint64_t blocks_per_cu = std::max(1, wpc / B_WAVES);
    uint64_t grid = (uint64_t)ctx->num_cus * blocks_per_cu;
    const uint64_t waves = grid * B_WAVES;
    bs.tiles_per_wave = (ntiles + waves - 1) / waves;
    grid = ((ntiles + bs.tiles_per_wave - 1) / bs.tiles_per_wave + B_WAVES - 1) / B_WAVES;
    ev_record(ctx, ctx->lane().evk0);
    HIPX(qmk_batch_scan(&bs, dq.as<BQuery>(), (uint32_t)grid, lds_bytes, wide, ctx->lane().stream));
    ev_record(ctx, ctx->lane().evk1);
    ev_sync(ctx, ctx->lane().evk1);
    float ms = 0;
    ev_ms(ctx, &ms, ctx->lane().evk0, ctx->lane().evk1);
    scan_ms_total = std::max(scan_ms_total, ms);
    qbufs.push_back(std::move(dq));
  }
  // shared groups: each query merges its intervals into its DIRECT table
  std::vector<BCombine> cb;
  uint32_t max_dom = 1;
  for (auto &g : groups) {
    for (BItem *it : g->items) {
      BCombine c;
      memset(&c, 0, sizeof c);
      c.src = g->table.as<uint64_t>();
      c.dst = it->table.as<uint64_t>();
      if (!c.src || !c.dst) fail(QM_ERR_QUERY_FAILED, "shared-interval merge without a table");
      size_t oi = std::find(order.begin(), order.end(), it) - order.begin();
      c.selected = selected.as<uint64_t>() + oi;
      c.dom = g->dom;
      c.e_lo = (uint32_t)it->e_lo;
      c.e_hi = (uint32_t)it->e_hi;
      c.src_words = g->entry_words;
      c.dst_words = it->ab.q.entry_words;
      c.nstates = it->ab.q.nstates;
      for (int s = 0; s < it->ab.q.nstates; ++s) {
        int gs = group_state(*g, *it, s);
        c.src_word[s] = g->st[gs].word;
        c.dst_word[s] = it->ab.q.states[s].word;
        c.kind[s] = it->ab.q.states[s].kind;
        c.fixed[s] = it->ab.q.states[s].fixed;
      }
      cb.push_back(c);
      max_dom = std::max(max_dom, g->dom);
    }
  }
  DevBuf dcb(ctx, std::max<size_t>(cb.size(), 1) * sizeof(BCombine));
  if (!cb.empty()) {
    h2d(ctx, dcb.p, reinterpret_cast<const uint8_t *>(cb.data()), cb.size() * sizeof(BCombine));
    HIPX(qmk_batch_combine(dcb.as<BCombine>(), (uint32_t)cb.size(), max_dom, ctx->lane().stream));
  }
  finalize_batch(ctx, order, out, status, selected);
  double ms = tm.stop();
  for (BItem *it : runs_done) order.push_back(it);
  for (BItem *it : order) {
    qm_result *res = out[it->idx];
    res->stats.rows_scanned = T->rows;
    res->stats.groups = res->rows;
    res->stats.bytes_algorithmic = referenced_bytes(it->cs, T->rows, 0);
    res->stats.device_ms = ms;
    res->stats.scan_ms = scan_ms_total;
    res->stats.kernel_launches = (uint32_t)nlaunches;
  }
}

// Zone maps (src/decode.ts:9-100 canSkipPage): a page can hold a selected row
// only if every top-level AND predicate's code range meets the page's
// [min, max] codes and, with OR groups, some group's predicates all do (a
// predicate without a code range, such as IN, IS NULL or a negation, meets
// every page). Pages that cannot match at either end of the table are cut
// off: `v` is a view of the remaining whole pages sharing the table's device
// columns. Pages that cannot match inside that span are marked in a device
// bitmap (`v.page_skip`) that the scan, selection and top-K kernels skip tile
// by tile. Returns true when `v` is to be used; `skipped` counts the pages cut
// off and `interior` the marked ones. Joins read their whole probe table.
// Per ZONE_ROWS-row page of T: can a row of the page pass the plan's filters
// (the top-level AND and, if any, one OR group), from the columns' zone maps
// (src/decode.ts:9-100 canSkipPage)? false when the maps cannot tell (zone
// maps off, a small table, a join, computed columns, no filter, or a filter the
// plan reports as an error).
bool zone_can(qm_ctx *ctx, const qm_table *T, const Plan &pl, std::vector<uint8_t> &can) {
  if (!ctx->tune.zone_maps || T->rows <= ZONE_ROWS || pl.has_join || !pl.computed.empty() ||
      (pl.filters.empty() && pl.groups.empty()))
    return false;
  ColSpace cs;
  cs.L = T;
  cs.plan = &pl;
  DQuery q;
  memset(&q, 0, sizeof q);
  ListStore ls;
  try {
    bind_filters(ctx, cs, pl.filters, pl.groups, q, ls, false);
  } catch (const Error &) {
    return false;  // the query itself reports the error
  }
  const uint64_t pages = (T->rows + ZONE_ROWS - 1) / ZONE_ROWS;
  const int ng = q.ngroups;
  // per group (0 = the AND, g = OR group g): can a page hold a row passing it?
  std::vector<std::vector<uint8_t>> meet(ng + 1, std::vector<uint8_t>(pages, 1));
  bool any = false;
  for (int i = 0; i < q.npreds; ++i) {
    const DPred &pr = q.preds[i];
    uint64_t lo = 0, hi = 0;
    bool never = false;
    if (!pred_code_range(pr, lo, hi, never)) continue;
    const int g = pr.group;
    if (g > ng) continue;
    if (never) {
      std::fill(meet[g].begin(), meet[g].end(), (uint8_t)0);
      any = true;
      continue;
    }
    const TCol *c = cs.tcols[pr.col];
    if (!c || c->zmin.size() != pages) continue;
    any = true;
    for (uint64_t p = 0; p < pages; ++p)
      if (c->zmax[p] < lo || c->zmin[p] > hi || c->zmin[p] > c->zmax[p]) meet[g][p] = 0;
  }
  if (!any) return false;
  can.assign(pages, 0);
  for (uint64_t p = 0; p < pages; ++p) {
    bool ok = meet[0][p] != 0;
    if (ok && ng) {
      bool some = false;
      for (int g = 1; g <= ng && !some; ++g) some = meet[g][p] != 0;
      ok = some;
    }
    can[p] = ok ? 1 : 0;
  }
  return true;
}

// A view of T without the pages at either end that the plan's filters cannot
// match, and (interior != nullptr) a bitmap of the pages inside that span they
// cannot match either; false (no view) when nothing can be skipped.
bool zone_view(qm_ctx *ctx, const qm_table *T, const Plan &pl, qm_table &v, uint32_t *skipped,
               uint32_t *interior = nullptr) {
  *skipped = 0;
  if (interior) *interior = 0;
  std::vector<uint8_t> can;
  if (!zone_can(ctx, T, pl, can)) return false;
  const uint64_t pages = can.size();
  uint64_t p0 = 0, p1 = pages - 1;
  while (p0 < pages && !can[p0]) ++p0;
  if (p0 == pages) p0 = p1 = 0;  // nothing can match: one page keeps the launch shapes
  else while (!can[p1]) --p1;
  std::vector<uint32_t> bits((p1 - p0 + 1 + 31) / 32, 0);
  uint32_t inner = 0;
  if (interior && can[p0])
    for (uint64_t p = p0; p <= p1; ++p)
      if (!can[p]) {
        bits[(p - p0) >> 5] |= 1u << ((p - p0) & 31);
        ++inner;
      }
  if (p0 == 0 && p1 == pages - 1 && !inner) return false;
  const uint64_t r0 = p0 * ZONE_ROWS, r1 = std::min(T->rows, (p1 + 1) * ZONE_ROWS);
  v.ctx = T->ctx;
  v.name = T->name;
  v.rows = r1 - r0;
  v.cols.clear();
  v.cols.reserve(T->cols.size());
  for (const TCol &t : T->cols) {  // the view needs no zone maps of its own
    TCol c;
    c.name = t.name;
    c.dtype = t.dtype;
    c.d = static_cast<uint8_t *>(t.d) + r0 * (uint64_t)dt_size(t.dtype);
    c.dnull = t.dnull ? t.dnull + r0 / 8 : nullptr;
    c.st = t.st;
    v.cols.push_back(std::move(c));
  }
  if (inner) {
    v.page_skip_buf = std::make_shared<DevBuf>(ctx, bits.size() * 4 + 64);
    h2d(ctx, v.page_skip_buf->p, bits.data(), bits.size());
    v.page_skip = v.page_skip_buf->as<uint32_t>();
  }
  *skipped = (uint32_t)(pages - (p1 - p0 + 1));
  if (interior) *interior = inner;
  return true;
}

// Pages of T no plan of a batch can match: a bitmap over the whole table (bit
// p = page p) for the batch's shared table-reading passes, in `buf`; nullptr
// when every page may hold a row some plan passes (or the maps cannot tell).
const uint32_t *batch_page_skip(qm_ctx *ctx, const qm_table *T, qm_plan *const *plans, uint32_t n,
                                std::shared_ptr<DevBuf> &buf, uint32_t *nskip) {
  *nskip = 0;
  if (!ctx->tune.zone_maps || T->rows <= ZONE_ROWS || n == 0) return nullptr;
  std::vector<uint8_t> any, can;
  size_t open = 0;  // pages no plan seen so far can match
  for (uint32_t i = 0; i < n; ++i) {
    if (!plans[i] || !zone_can(ctx, T, plans[i]->plan, can)) return nullptr;
    if (any.empty()) {
      any.assign(can.size(), 0);
      open = can.size();
    }
    for (size_t p = 0; p < can.size(); ++p)
      if (can[p] && !any[p]) {
        any[p] = 1;
        --open;
      }
    if (open == 0) return nullptr;  // every page may hold a row some plan passes
  }
  std::vector<uint32_t> bits((any.size() + 31) / 32, 0);
  uint32_t cnt = 0;
  for (size_t p = 0; p < any.size(); ++p)
    if (!any[p]) {
      bits[p >> 5] |= 1u << (p & 31);
      ++cnt;
    }
  if (!cnt) return nullptr;
  buf = std::make_shared<DevBuf>(ctx, bits.size() * 4 + 64);
  h2d(ctx, buf->p, bits.data(), bits.size());
  *nskip = cnt;
  return buf->as<uint32_t>();
}

}  // namespace

extern "C" int qm_exec(qm_plan *p, qm_table *left, qm_table *right, qm_result **out) {
  API_BEGIN
  if (!p || !left || !out) fail(QM_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  qm_ctx *ctx = p->ctx;
  if (left->ctx != ctx || (right && right->ctx != ctx)) fail(QM_ERR_INVALID_ARG, "handles from different contexts");
  const Plan &pl = p->plan;
  if (pl.has_join && !right) fail(QM_ERR_TABLE_NOT_FOUND, "join needs the right (build) table");
  LaneHold hold(ctx);
  HIPX(hipSetDevice(ctx->device));
  auto t0 = std::chrono::steady_clock::now();
  std::unique_ptr<qm_result> res;
  const bool agg = !pl.aggs.empty() || !pl.groupby.empty();
  if (agg) {
    if (pl.aggs.empty()) fail(QM_ERR_INVALID_AGGREGATE, "GROUP BY without aggregates");
    std::vector<const Plan *> plans{&pl};
    std::vector<qm_result *> results(1, nullptr);
    JoinIndex ji;
    DevBuf keep;
    if (pl.has_join) {
      ColSpace cs;
      cs.L = left;
      cs.R = right;
      cs.rkey = pl.rkey;
      cs.plan = &pl;
      const std::vector<KeyRange> ranges = join_ranges(ctx, right, pl, true);
      if (ranges.empty()) {
        uint32_t l = 0;
        Timer tb(ctx);
        build_join(ctx, cs, pl, ji, &l, &keep, true);
        double bms = tb.stop();
        try {
          run_aggregates(ctx, left, right, plans, results, true, &ji);
        } catch (...) {
          delete results[0];
          throw;
        }
        results[0]->stats.device_ms += bms;
        results[0]->stats.kernel_launches += l;
      } else {
        // the index in key-range passes (each builds its range's index, then probes)
        try {
          run_aggregates(ctx, left, right, plans, results, true, nullptr, &ranges);
        } catch (...) {
          delete results[0];
          throw;
        }
      }
    } else {
      // pages the filter cannot match at either end are not read (zone maps)
      qm_table view;
      uint32_t skipped = 0, interior = 0;
      qm_table *T = zone_view(ctx, left, pl, view, &skipped, &interior) ? &view : left;
      // a key column stored in order: segmented reduction over its runs
      BItem it;
      it.idx = 0;
      it.plan = &pl;
      batch_classify(ctx, T, pl, it);
      bool inner_read = false;  // the path ran reads the pages marked inside the span
      try {
        if (it.kind == BK_SORTED || (it.kind == BK_SHARED && ctx->tune.shared_lean)) {
          std::vector<BItem *> items{&it};
          run_batch(ctx, T, items, results, nullptr, T->page_skip);
          inner_read = it.kind == BK_SORTED;  // the shared-interval pass skips the marked pages
        } else {
          run_aggregates(ctx, T, nullptr, plans, results, false, nullptr);
          const uint32_t st = results[0]->stats.strategy;
          inner_read = st == QM_STRAT_PART_HASH || st == QM_STRAT_PART_DIRECT;
        }
      } catch (...) {
        delete results[0];
        throw;
      }
      results[0]->stats.pages_skipped = skipped + (inner_read ? 0 : interior);
    }
    res.reset(results[0]);
  } else {
    res.reset(new qm_result());
    // row outputs over a view: the selected rows, their order (row id ties)
    // and the gathered values are all taken within the same view
    qm_table view;
    uint32_t skipped = 0, interior = 0;
    qm_table *T = zone_view(ctx, left, pl, view, &skipped, &interior) ? &view : left;
    run_rows_query(ctx, T, right, pl, res.get());
    res->stats.pages_skipped = skipped + interior;
  }
  stream_sync(ctx);
  res->stats.wall_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  *out = res.release();
  API_END
}

// Device memory one execution of `p` needs beyond the resident tables
// (group tables, partition records, join index, result columns), from the
// same binding the execution uses; the serving loop admits requests against
// the HBM budget with it.
// One filter -> GROUP BY -> aggregate plan without any device allocation:
// the group table, the query descriptor, the counters and the compaction
// temporaries live in the context's scratch array; the result columns go to
// the caller's device buffers. Calls that overlap on one device race (they
// share the array); every call clears what it needs.
extern "C" int qm_exec_noalloc(qm_plan *p, qm_table *t, void *const *out_data, uint32_t *const *out_nulls,
                               uint64_t cap, uint64_t *rows_out) {
  API_BEGIN
  if (!p || !t || !out_data || !rows_out) fail(QM_ERR_INVALID_ARG, "null argument");
  *rows_out = 0;
  qm_ctx *ctx = t->ctx;
  if (p->ctx != ctx) fail(QM_ERR_INVALID_ARG, "plan and table belong to different contexts");
  const Plan &pl = p->plan;
  if (pl.aggs.empty() || pl.has_join || !pl.computed.empty() || pl.has_sort || pl.limit >= 0 || pl.offset ||
      pl.post())
    fail(QM_ERR_UNSUPPORTED, "qm_exec_noalloc runs filter -> GROUP BY -> aggregate plans (no join, computed "
                             "column, sort, limit, offset, HAVING or multi-column ORDER BY)");
  auto has_list = [](const std::vector<Filter> &fs) {
    for (const Filter &f : fs) if (f.op == OP_IN || f.op == OP_NOT_IN) return true;
    return false;
  };
  bool lists = has_list(pl.filters);
  for (const auto &g : pl.groups) lists = lists || has_list(g);
  if (lists) fail(QM_ERR_UNSUPPORTED, "qm_exec_noalloc takes no IN / NOT IN lists (their values need device memory)");
  LaneHold hold(ctx);
  std::lock_guard<std::mutex> sg(ctx->scratch_mu);
  HIPX(hipSetDevice(ctx->device));
  if (!ctx->scratch) fail(QM_ERR_QUERY_FAILED, "no scratch array on this device");
  uint8_t *const base = static_cast<uint8_t *>(ctx->scratch);
  ColSpace cs;
  cs.L = t;
  cs.plan = &pl;
  AggBinding ab;
  ListStore ls;
  bind_filters(ctx, cs, pl.filters, pl.groups, ab.q, ls, false);
  const DQuery filt = ab.q;
  auto a256 = [](uint64_t x) { return (x + 255) & ~255ull; };
  for (uint32_t regrow = 0;; ++regrow) {
    ab.q = filt;
    ab.outs.clear();
    ab.key_slots.clear();
    bind_aggregate(ctx, cs, pl, t->rows, false, ab, regrow);
    ab.part = false;  // one scan into the table
    DQuery &q = ab.q;
    const uint64_t nb = qmk_compact_blocks(ab.nslots);
    const uint64_t tw = qmk_exscan_tmp_words(nb);
    // scratch layout: table | status, selected | descriptor | block counts, offsets, scan temporaries
    const uint64_t o_table = 0, tb = a256(ab.table_words * 8);
    const uint64_t o_stat = o_table + tb, o_desc = o_stat + 256;
    const uint64_t o_cnt = o_desc + a256(sizeof(DQuery)), o_offs = o_cnt + a256(nb * 8);
    const uint64_t o_tmp = o_offs + a256(nb * 8), need = o_tmp + a256((tw + 1) * 8);
    if (need > QM_NOALLOC_SCRATCH_BYTES)
      fail(QM_ERR_MEMORY_EXCEEDED, "qm_exec_noalloc: the plan needs %llu bytes of scratch, the maximum is %llu",
           (unsigned long long)need, (unsigned long long)QM_NOALLOC_SCRATCH_BYTES);
    uint64_t *table = reinterpret_cast<uint64_t *>(base + o_table);
    uint64_t *stat = reinterpret_cast<uint64_t *>(base + o_stat);  // [0] status words, [1] selected rows
    // cleared with a fill kernel (no memset nodes for callers that capture)
    HIPX(qmk_fill_u64(table, ab.table_words, 0, ctx->lane().stream));
    HIPX(qmk_fill_u64(stat, 4, 0, ctx->lane().stream));
    q.table = table;
    q.status = reinterpret_cast<uint32_t *>(stat);
    q.selected = stat + 1;
    q.lds_offset = 0;
    DQuery *dq = reinterpret_cast<DQuery *>(base + o_desc);
    h2d(ctx, dq, reinterpret_cast<const uint8_t *>(&q), sizeof(DQuery));
    DScan sc;
    memset(&sc, 0, sizeof sc);
    sc.nrows = t->rows;
    sc.ncols = (int)cs.cols.size();
    sc.nqueries = 1;
    for (size_t i = 0; i < cs.cols.size(); ++i) sc.cols[i] = cs.cols[i];
    int grid, block;
    size_t lds_bytes;
    scan_launch_shape(ctx, t->rows, (uint64_t)q.lds_words, &grid, &block, &lds_bytes);
    if (lds_bytes > 160 * 1024) fail(QM_ERR_UNSUPPORTED, "group table exceeds the workgroup LDS");
    if (t->rows) HIPX(qmk_scan_agg(&sc, dq, grid, block, lds_bytes, ctx->lane().stream));
    uint32_t st[2] = {0, 0};
    d2h(ctx, st, stat, 2);
    if (st[0]) {
      if (regrow >= 4) fail(QM_ERR_QUERY_FAILED, "group-by hash table kept overflowing");
      continue;
    }
    DFinal f;
    memset(&f, 0, sizeof f);
    f.table = table;
    f.nslots = ab.nslots;
    f.entry_words = q.entry_words;
    f.count_word = q.count_word;
    f.gmode = q.gmode;
    f.always_one_row = ab.always_one_row ? 1 : 0;
    f.key = q.key;
    f.status = q.status;
    uint64_t *cnt = reinterpret_cast<uint64_t *>(base + o_cnt), *offs = reinterpret_cast<uint64_t *>(base + o_offs);
    uint64_t *tmp = reinterpret_cast<uint64_t *>(base + o_tmp);
    HIPX(qmk_compact_count(&f, cnt, ctx->lane().stream));
    HIPX(qmk_exscan_u64(cnt, nb, offs, tmp + tw, tmp, ctx->lane().stream));
    uint64_t total = 0;
    d2h(ctx, &total, tmp + tw, 1);
    if (total > cap)
      fail(QM_ERR_INVALID_ARG, "qm_exec_noalloc: the result has %llu rows, the output columns hold %llu",
           (unsigned long long)total, (unsigned long long)cap);
    f.nouts = (int)ab.outs.size();
    for (size_t k = 0; k < ab.outs.size(); ++k) {
      f.outs[k] = ab.outs[k].o;
      if (!out_data[k]) fail(QM_ERR_INVALID_ARG, "output column %zu has no buffer", k);
      f.outs[k].data = out_data[k];
      uint32_t *nbm = (out_nulls && ab.outs[k].o.kind != OK_COUNT) ? out_nulls[k] : nullptr;
      if (nbm && (reinterpret_cast<uintptr_t>(nbm) & 7))
        fail(QM_ERR_INVALID_ARG, "NULL bitmap of output column %zu is not 8-byte aligned", k);
      f.outs[k].nulls = nbm;
      if (nbm && total) HIPX(qmk_fill_u64(reinterpret_cast<uint64_t *>(nbm), (total + 63) / 64, 0, ctx->lane().stream));
    }
    if (total) HIPX(qmk_compact_write(&f, offs, ctx->lane().stream));
    d2h(ctx, st, stat, 2);
    if (st[1]) fail(QM_ERR_OVERFLOW, "SUM over an int64 column does not fit int64");
    *rows_out = total;
    break;
  }
  API_END
}

extern "C" int qm_plan_device_bytes(qm_plan *p, qm_table *left, qm_table *right, uint64_t *bytes) {
  API_BEGIN
  if (!p || !left || !bytes) fail(QM_ERR_INVALID_ARG, "null argument");
  *bytes = 0;
  qm_ctx *ctx = p->ctx;
  const Plan &pl = p->plan;
  if (pl.has_join && !right) fail(QM_ERR_TABLE_NOT_FOUND, "join needs the right (build) table");
  LaneHold hold(ctx);
  HIPX(hipSetDevice(ctx->device));
  uint64_t b = 0;
  const uint64_t n = left->rows;
  if (pl.has_join) {
    // the join index (of one key-range pass when the budget needs several)
    const int rk = right->find(pl.rkey);
    bool kf = true, uniq = false;
    uint64_t dom = 0;
    if (rk >= 0 && !dt_float(right->cols[rk].dtype) && right->cols[rk].st.has_int) {
      kf = false;
      dom = (uint64_t)right->cols[rk].st.imax - (uint64_t)right->cols[rk].st.imin + 1;
      const ColStats &ks = right->cols[rk].st;
      uniq = ks.dup_frac == 0 && !ks.nulls && ks.distinct >= 0.9 * (double)ks.nonnull;
    }
    const uint64_t P = std::max<uint64_t>(1, join_ranges(ctx, right, pl, !pl.aggs.empty()).size());
    const uint64_t pdom = kf ? 0 : (dom ? (dom + P - 1) / P : (~0ull / P + 1));
    b += join_index_bytes(ctx, right->rows / P + 1, pdom, kf, !pl.aggs.empty(), uniq) + 4096;
  }
  if (!pl.aggs.empty()) {
    BItem it;
    it.plan = &pl;
    ColSpace &cs = it.cs;
    cs.L = left;
    cs.R = pl.has_join ? right : nullptr;
    cs.rkey = pl.rkey;
    cs.plan = &pl;
    bind_filters(ctx, cs, pl.filters, pl.groups, it.ab.q, it.ls, false);
    bind_aggregate(ctx, cs, pl, n, pl.has_join, it.ab, 0);
    b += it.ab.table_words * 8;
    if (it.ab.part) {
      const uint64_t R = (1 + it.ab.srcs.size() + (it.ab.has_mask ? 1 : 0)) <= 4 ? 4 : 8;
      b += 2 * n * R * 8;  // pass-1 regions (upper bound) + partition-ordered records
    }
    b += it.ab.nslots * 8 * (it.ab.outs.size() + 1);  // compaction and result columns
  } else {
    // rows: selection ids + gathered columns (bounded by limit when present)
    uint64_t out = (pl.limit >= 0) ? (uint64_t)(pl.limit + pl.offset) : n;
    if (pl.has_join) out = std::max(out, n);
    uint64_t width = 8;
    for (const TCol &c : left->cols) width += dt_size(c.dtype);
    if (pl.has_join) for (const TCol &c : right->cols) width += dt_size(c.dtype);
    b += std::min(out, n * (pl.has_join ? 4 : 1)) * width;
    if (pl.has_sort && pl.limit < 0) b += n * 24;  // radix sort keys + values
  }
  *bytes = b;
  API_END
}

extern "C" int qm_measure_read_bw(qm_ctx *ctx, uint64_t bytes, uint32_t reps, double *gbps) {
  API_BEGIN
  if (!ctx || !gbps || bytes < (1u << 20) || reps == 0) fail(QM_ERR_INVALID_ARG, "bad argument");
  LaneHold hold(ctx);
  HIPX(hipSetDevice(ctx->device));
  bytes &= ~(uint64_t)15;
  DevBuf buf(ctx, bytes), sink(ctx, 64);
  HIPX(qmk_fill_u64(buf.as<uint64_t>(), bytes / 8, 0x0123456789ABCDEFull, ctx->lane().stream));
  const uint32_t grid = (uint32_t)ctx->num_cus * 8;
  HIPX(qmk_stream_read(buf.p, bytes, grid, sink.as<uint64_t>(), ctx->lane().stream));  // warm
  ev_record(ctx, ctx->lane().evk0);
  for (uint32_t r = 0; r < reps; ++r) HIPX(qmk_stream_read(buf.p, bytes, grid, sink.as<uint64_t>(), ctx->lane().stream));
  ev_record(ctx, ctx->lane().evk1);
  ev_sync(ctx, ctx->lane().evk1);
  float ms = 0;
  ev_ms(ctx, &ms, ctx->lane().evk0, ctx->lane().evk1);
  *gbps = (double)bytes * reps / (ms * 1e-3) / 1e9;
  API_END
}

extern "C" int qm_tune_set(qm_ctx *ctx, const char *name, double value) {
  API_BEGIN
  if (!ctx) fail(QM_ERR_INVALID_ARG, "null context");
  const KnobDef *k = find_knob(name);
  if (!k) fail(QM_ERR_INVALID_ARG, "unknown tuning knob: %s", name ? name : "(null)");
  if (std::isnan(value)) fail(QM_ERR_INVALID_ARG, "knob %s: NaN", name);
  LaneHold hold(ctx, true);
  k->set(ctx->tune, std::min(k->hi, std::max(k->lo, value)));
  API_END
}

extern "C" int qm_tune_get(qm_ctx *ctx, const char *name, double *value) {
  API_BEGIN
  if (!ctx || !value) fail(QM_ERR_INVALID_ARG, "null argument");
  const KnobDef *k = find_knob(name);
  if (!k) fail(QM_ERR_INVALID_ARG, "unknown tuning knob: %s", name ? name : "(null)");
  std::shared_lock<std::shared_mutex> g(ctx->tables_mu);
  *value = k->get(ctx->tune);
  API_END
}

namespace {
// The batch itself, on the lane the caller holds: fused launches for the
// aggregate plans that fit them, multi-query scans for the other aggregates,
// one pass per row query. On return every result is complete (the lane's
// stream has drained) or, on error, none is left allocated.
void batch_exec_body(qm_ctx *ctx, qm_plan *const *plans, uint32_t nplans, qm_table *table,
                     std::vector<qm_result *> &all) {
  all.assign(nplans, nullptr);
  std::vector<uint32_t> agg_idx, other_idx;
  std::vector<std::unique_ptr<BItem>> fused;
  for (uint32_t i = 0; i < nplans; ++i) {
    if (!plans[i]) fail(QM_ERR_INVALID_ARG, "null plan %u", i);
    const Plan &pl = plans[i]->plan;
    bool agg = !pl.aggs.empty();
    if (agg && !pl.has_join) {
      std::unique_ptr<BItem> it(new BItem());
      it->idx = i;
      it->plan = &pl;
      batch_classify(ctx, table, pl, *it);
      if (it->kind != BK_NONE) fused.push_back(std::move(it));
      else agg_idx.push_back(i);
    } else {
      other_idx.push_back(i);
    }
  }
  try {
    if (!fused.empty()) {
      std::vector<BItem *> items;
      std::vector<qm_plan *> fplans;
      for (auto &it : fused) {
        items.push_back(it.get());
        fplans.push_back(plans[it->idx]);
      }
      // pages no fused plan can match: the partitioned-runs row scatter and
      // the shared-interval pass do not read them (zone maps)
      std::shared_ptr<DevBuf> skipbuf;
      uint32_t nskip = 0;
      const uint32_t *skip = batch_page_skip(ctx, table, fplans.data(), (uint32_t)fplans.size(), skipbuf, &nskip);
      std::vector<uint32_t> rejected;
      run_batch(ctx, table, items, all, &rejected, skip);
      for (auto &it : fused) {
        qm_result *r = all[it->idx];
        if (r && (r->stats.strategy == QM_STRAT_PART_RUNS || r->stats.strategy == QM_STRAT_SHARED_INTERVAL))
          r->stats.pages_skipped = nskip;
      }
      agg_idx.insert(agg_idx.end(), rejected.begin(), rejected.end());
      std::sort(agg_idx.begin(), agg_idx.end());
    }
    // fused launches of up to MAX_QUERIES aggregate queries sharing one scan
    size_t pos = 0;
    while (pos < agg_idx.size()) {
      size_t n = std::min<size_t>(MAX_QUERIES, agg_idx.size() - pos);
      std::vector<const Plan *> ps;
      for (size_t k = 0; k < n; ++k) ps.push_back(&plans[agg_idx[pos + k]]->plan);
      std::vector<qm_result *> rs(n, nullptr);
      try {
        run_aggregates(ctx, table, nullptr, ps, rs, false, nullptr);
      } catch (const Error &e) {
        for (auto *r : rs) delete r;
        if (e.code == QM_ERR_UNSUPPORTED && n > 1) {
          // too much LDS / columns for one launch: run them one by one
          for (size_t k = 0; k < n; ++k) {
            std::vector<const Plan *> one{ps[k]};
            std::vector<qm_result *> r1(1, nullptr);
            run_aggregates(ctx, table, nullptr, one, r1, false, nullptr);
            all[agg_idx[pos + k]] = r1[0];
          }
          pos += n;
          continue;
        }
        throw;
      }
      for (size_t k = 0; k < n; ++k) all[agg_idx[pos + k]] = rs[k];
      pos += n;
    }
    for (uint32_t i : other_idx) {
      if (plans[i]->plan.has_join) fail(QM_ERR_UNSUPPORTED, "joins are not batched; run them with qm_exec");
      qm_result *r = new qm_result();
      all[i] = r;
      run_rows_query(ctx, table, nullptr, plans[i]->plan, r);
    }
  } catch (...) {
    for (auto *&r : all) { delete r; r = nullptr; }
    throw;
  }
  stream_sync(ctx);
}
}  // namespace

extern "C" int qm_batch_exec(qm_ctx *ctx, qm_plan *const *plans, uint32_t nplans, qm_table *table,
                             qm_result **results_out) {
  API_BEGIN
  if (!ctx || !table || (nplans && (!plans || !results_out))) fail(QM_ERR_INVALID_ARG, "null argument");
  for (uint32_t i = 0; i < nplans; ++i) results_out[i] = nullptr;
  LaneHold hold(ctx);
  auto t0 = std::chrono::steady_clock::now();
  std::vector<qm_result *> all;
  batch_exec_body(ctx, plans, nplans, table, all);
  double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  for (uint32_t i = 0; i < nplans; ++i) {
    all[i]->stats.wall_ms = wall;
    results_out[i] = all[i];
  }
  API_END
}

// ---- prepared batches: one batch captured into a HIP graph, replayed ----
struct qm_prepared {
  qm_ctx *ctx = nullptr;
  qm_table *table = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  qm_capture cap;                    // tape, upload staging, the graph's private buffer pool
  std::vector<qm_result *> tmpl;     // the captured run's results: the buffers every replay writes
  uint64_t nodes = 0;
  std::mutex mu;                     // one replay at a time (they share the buffers)
  void free_all() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    for (qm_result *r : tmpl) delete r;
    tmpl.clear();
    for (auto &kv : cap.free) ctx->release(kv.second, kv.first);
    cap.free.clear();
    if (cap.staging) (void)hipHostFree(cap.staging);
    cap.staging = nullptr;
  }
};

namespace {
// Ends a failed capture without ever waiting on a stream that may still be
// capturing: if the capture cannot be ended the lane gets a new stream.
void abandon_capture(qm_lane &L) {
  hipGraph_t g = nullptr;
  (void)hipStreamEndCapture(L.stream, &g);
  if (g) (void)hipGraphDestroy(g);
  (void)hipGetLastError();
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(L.stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
    (void)hipGetLastError();
    hipStream_t ns = nullptr;
    if (hipStreamCreateWithFlags(&ns, hipStreamNonBlocking) == hipSuccess) L.stream = ns;
  }
}
}  // namespace

extern "C" int qm_batch_prepare(qm_ctx *ctx, qm_plan *const *plans, uint32_t nplans, qm_table *table,
                                qm_prepared **out) {
  API_BEGIN
  if (!out) fail(QM_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (!ctx || !table || (nplans && !plans)) fail(QM_ERR_INVALID_ARG, "null argument");
  if (table->ctx != ctx) fail(QM_ERR_INVALID_ARG, "handles from different contexts");
  for (uint32_t i = 0; i < nplans; ++i) {
    if (!plans[i]) fail(QM_ERR_INVALID_ARG, "null plan %u", i);
    if (plans[i]->plan.has_join) fail(QM_ERR_UNSUPPORTED, "joins are not batched; run them with qm_exec");
    if (plans[i]->plan.post())
      fail(QM_ERR_UNSUPPORTED, "prepared batches take no HAVING, DISTINCT or multi-column ORDER BY (plan %u)", i);
  }
  LaneHold hold(ctx);
  qm_lane &L = ctx->lane();
  std::unique_ptr<qm_prepared> P(new qm_prepared());
  P->ctx = ctx;
  P->table = table;
  struct Detach {
    qm_lane &L;
    qm_prepared *P;
    bool armed = true;
    ~Detach() {
      L.cap = nullptr;
      if (armed) P->free_all();
    }
  } detach{L, P.get()};
  // 1. an ordinary run that records every readback the host plans with and
  //    the buffers it needs
  P->cap.mode = LANE_RECORD;
  L.cap = &P->cap;
  {
    std::vector<qm_result *> rec;
    batch_exec_body(ctx, plans, nplans, table, rec);
    for (qm_result *r : rec) delete r;
  }
  // 2. the graph's private pool and the pinned upload staging, allocated now
  P->cap.mode = LANE_NORMAL;
  for (const auto &kv : P->cap.peak)
    for (int c = 0; c < kv.second; ++c) {
      P->cap.free.emplace(kv.first, ctx->alloc(kv.first));
      P->cap.private_bytes += kv.first;
    }
  P->cap.staging_cap = P->cap.h2d_bytes + 4096;
  HIPX(hipHostMalloc(reinterpret_cast<void **>(&P->cap.staging), P->cap.staging_cap, hipHostMallocDefault));
  stream_sync(ctx);  // every buffer above is idle before the graph takes it
  // 3. the same host logic again while the stream is captured
  hipStreamCaptureMode tmode = hipStreamCaptureModeRelaxed;
  HIPX(hipThreadExchangeStreamCaptureMode(&tmode));
  struct ModeBack {
    hipStreamCaptureMode m;
    ~ModeBack() { (void)hipThreadExchangeStreamCaptureMode(&m); }
  } mode_back{tmode};
  P->cap.mode = LANE_CAPTURE;
  P->cap.pos = 0;
  HIPX(hipStreamBeginCapture(L.stream, hipStreamCaptureModeRelaxed));
  try {
    batch_exec_body(ctx, plans, nplans, table, P->tmpl);
  } catch (...) {
    P->cap.mode = LANE_NORMAL;
    abandon_capture(L);
    throw;
  }
  P->cap.mode = LANE_NORMAL;
  const hipError_t ee = hipStreamEndCapture(L.stream, &P->graph);
  if (ee != hipSuccess) {
    abandon_capture(L);
    fail(QM_ERR_QUERY_FAILED, "prepared batch: ending the capture: %s", hipGetErrorString(ee));
  }
  if (P->cap.pos != P->cap.tape.size())
    fail(QM_ERR_QUERY_FAILED, "prepared batch: the captured run read back %zu of %zu recorded values", P->cap.pos,
         P->cap.tape.size());
  size_t nn = 0;
  HIPX(hipGraphGetNodes(P->graph, nullptr, &nn));
  P->nodes = nn;
  HIPX(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
  table->prepared.fetch_add(1);
  detach.armed = false;
  *out = P.release();
  API_END
}

extern "C" int qm_prepared_run(qm_prepared *p, qm_result **results_out) {
  API_BEGIN
  if (!p || (!results_out && !p->tmpl.empty())) fail(QM_ERR_INVALID_ARG, "null argument");
  for (size_t i = 0; i < p->tmpl.size(); ++i) results_out[i] = nullptr;
  qm_ctx *ctx = p->ctx;
  LaneHold hold(ctx);
  std::lock_guard<std::mutex> g(p->mu);
  auto t0 = std::chrono::steady_clock::now();
  Timer tm(ctx);
  HIPX(hipGraphLaunch(p->exec, ctx->lane().stream));
  const double ms = tm.stop();
  const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  std::vector<std::unique_ptr<qm_result>> views;
  for (const qm_result *t : p->tmpl) {
    std::unique_ptr<qm_result> r(new qm_result());
    r->ctx = ctx;
    r->rows = t->rows;
    r->stats = t->stats;
    r->stats.device_ms = ms;
    r->stats.wall_ms = wall;
    r->stats.kernel_launches = 1;
    for (const OutCol &c : t->cols) {  // views: the prepared batch owns the memory
      OutCol v;
      v.name = c.name;
      v.dtype = c.dtype;
      v.data.p = c.data.p;
      v.data.n = c.data.n;
      v.nulls.p = c.nulls.p;
      v.nulls.n = c.nulls.n;
      r->cols.push_back(std::move(v));
    }
    views.push_back(std::move(r));
  }
  for (size_t i = 0; i < views.size(); ++i) results_out[i] = views[i].release();
  API_END
}

extern "C" int qm_prepared_info(qm_prepared *p, uint32_t *nplans, uint64_t *graph_nodes, uint64_t *device_bytes) {
  API_BEGIN
  if (!p) fail(QM_ERR_INVALID_ARG, "null argument");
  if (nplans) *nplans = (uint32_t)p->tmpl.size();
  if (graph_nodes) *graph_nodes = p->nodes;
  if (device_bytes) *device_bytes = p->cap.private_bytes;
  API_END
}

extern "C" int qm_prepared_close(qm_prepared *p) {
  API_BEGIN
  if (!p) return QM_OK;
  {
    std::lock_guard<std::mutex> g(p->mu);  // no replay is in flight past this point
    p->free_all();
  }
  p->table->prepared.fetch_sub(1);
  delete p;
  API_END
}

// ---------------------------------------------------------------------------
// results
// ---------------------------------------------------------------------------
namespace {
struct QmcbCol {
  bool has_nulls;
  std::vector<uint8_t> bitmap;
};

size_t qmcb_data_bytes(int dt, uint64_t rows) {
  if (dt == DT_BOOL) return (rows + 7) / 8;
  return rows * (uint64_t)dt_size(dt);
}

std::vector<QmcbCol> scan_nulls(qm_result *r) {
  std::vector<QmcbCol> v(r->cols.size());
  const uint64_t nb = (r->rows + 7) / 8;
  for (size_t i = 0; i < r->cols.size(); ++i) {
    v[i].has_nulls = false;
    if (!r->cols[i].nulls.p || !nb) continue;
    v[i].bitmap.resize(nb);
    d2h(r->ctx, v[i].bitmap.data(), r->cols[i].nulls.p, nb);
    for (uint8_t b : v[i].bitmap) if (b) { v[i].has_nulls = true; break; }
  }
  return v;
}

uint64_t qmcb_size(qm_result *r, const std::vector<QmcbCol> &nc) {
  uint64_t s = 10;
  for (auto &c : r->cols) s += 2 + c.name.size() + 2;
  for (size_t i = 0; i < r->cols.size(); ++i) {
    if (nc[i].has_nulls) s += (r->rows + 7) / 8;
    s += qmcb_data_bytes(r->cols[i].dtype, r->rows);
  }
  return s;
}
}  // namespace

extern "C" int qm_result_rows(qm_result *r, uint64_t *rows) {
  API_BEGIN
  if (!r || !rows) fail(QM_ERR_INVALID_ARG, "null argument");
  *rows = r->rows;
  API_END
}

extern "C" int qm_result_qmcb_size(qm_result *r, uint64_t *size) {
  API_BEGIN
  if (!r || !size) fail(QM_ERR_INVALID_ARG, "null argument");
  if (r->rows > 0xFFFFFFFFull) fail(QM_ERR_UNSUPPORTED, "QMCB row count is 32-bit");
  LaneHold hold(r->ctx);
  *size = qmcb_size(r, scan_nulls(r));
  API_END
}

extern "C" int qm_result_qmcb(qm_result *r, uint8_t *out, uint64_t cap, uint64_t *written) {
  API_BEGIN
  if (!r || !out) fail(QM_ERR_INVALID_ARG, "null argument");
  if (r->rows > 0xFFFFFFFFull) fail(QM_ERR_UNSUPPORTED, "QMCB row count is 32-bit");
  if (r->cols.size() > 0xFFFF) fail(QM_ERR_UNSUPPORTED, "too many columns");
  qm_ctx *ctx = r->ctx;
  LaneHold hold(ctx);
  HIPX(hipSetDevice(ctx->device));
  auto nc = scan_nulls(r);
  uint64_t need = qmcb_size(r, nc);
  if (cap < need) fail(QM_ERR_INVALID_ARG, "QMCB buffer too small: need %llu bytes", (unsigned long long)need);
  // header (src/columnar.ts:523-613)
  uint8_t *w = out;
  auto put32 = [&](uint32_t v) { memcpy(w, &v, 4); w += 4; };
  auto put16 = [&](uint16_t v) { memcpy(w, &v, 2); w += 2; };
  put32(0x42434D51u);
  put32((uint32_t)r->rows);
  put16((uint16_t)r->cols.size());
  for (size_t i = 0; i < r->cols.size(); ++i) {
    const OutCol &c = r->cols[i];
    if (c.name.size() > 0xFFFF) fail(QM_ERR_UNSUPPORTED, "column name too long");
    put16((uint16_t)c.name.size());
    memcpy(w, c.name.data(), c.name.size());
    w += c.name.size();
    *w++ = (uint8_t)qmcb_tag(c.dtype);
    *w++ = nc[i].has_nulls ? 1 : 0;
  }
  for (size_t i = 0; i < r->cols.size(); ++i) {
    const OutCol &c = r->cols[i];
    if (nc[i].has_nulls) {
      memcpy(w, nc[i].bitmap.data(), nc[i].bitmap.size());
      w += nc[i].bitmap.size();
    }
    if (c.dtype == DT_BOOL) {
      std::vector<uint8_t> bytes(r->rows);
      d2h(ctx, bytes.data(), c.data.p, r->rows);
      uint64_t nb = (r->rows + 7) / 8;
      memset(w, 0, nb);
      for (uint64_t k = 0; k < r->rows; ++k) if (bytes[k]) w[k >> 3] |= (uint8_t)(1u << (k & 7));
      w += nb;
    } else {
      size_t b = qmcb_data_bytes(c.dtype, r->rows);
      if (b) {
        HIPX(hipMemcpyAsync(w, c.data.p, b, hipMemcpyDeviceToHost, ctx->lane().stream));
        w += b;
      }
    }
  }
  stream_sync(ctx);
  if (written) *written = (uint64_t)(w - out);
  API_END
}

// ---- QMCB written on the device into pinned host memory ----
namespace {
// Pinned host buffers handed out by qm_result_qmcb_pinned, process-wide (a
// buffer may outlive its context): live ones by address, released ones by
// size class for reuse (no host allocator call per batch of a cursor).
std::mutex g_pin_mu;
std::map<const uint8_t *, size_t> g_pin_live;
std::multimap<size_t, uint8_t *> g_pin_free;

size_t pin_class(size_t bytes) {
  size_t c = 64 << 10;
  while (c < bytes) c <<= 1;
  return c;
}

uint8_t *pin_take(size_t bytes) {
  const size_t c = pin_class(bytes);
  {
    std::lock_guard<std::mutex> g(g_pin_mu);
    auto it = g_pin_free.find(c);
    if (it != g_pin_free.end()) {
      uint8_t *p = it->second;
      g_pin_free.erase(it);
      g_pin_live[p] = c;
      return p;
    }
  }
  void *p = nullptr;
  HIPX(hipHostMalloc(&p, c, hipHostMallocDefault));
  std::lock_guard<std::mutex> g(g_pin_mu);
  g_pin_live[static_cast<uint8_t *>(p)] = c;
  return static_cast<uint8_t *>(p);
}
}  // namespace

extern "C" int qm_result_qmcb_pinned(qm_result *r, uint64_t row0, uint64_t nrows, const uint8_t **out,
                                     uint64_t *size) {
  API_BEGIN
  if (!r || !out || !size) fail(QM_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  *size = 0;
  if (r->cols.size() > 0xFFFF) fail(QM_ERR_UNSUPPORTED, "too many columns");
  row0 = std::min<uint64_t>(row0, r->rows);
  nrows = std::min<uint64_t>(nrows, r->rows - row0);
  if (nrows > 0xFFFFFFFFull) fail(QM_ERR_UNSUPPORTED, "QMCB row count is 32-bit");
  qm_ctx *ctx = r->ctx;
  LaneHold hold(ctx);
  const size_t nc = r->cols.size();
  const uint64_t nb = (nrows + 7) / 8;
  // bitmaps of the range and packed BOOL columns, on the device
  std::vector<int64_t> bits_off(nc, -1), bool_off(nc, -1);
  uint64_t stage = 0;
  for (size_t i = 0; i < nc; ++i) {
    if (r->cols[i].nulls.p) { bits_off[i] = (int64_t)stage; stage += (nb + 63) / 64 * 64; }
    if (r->cols[i].dtype == DT_BOOL) { bool_off[i] = (int64_t)stage; stage += (nb + 63) / 64 * 64; }
  }
  DevBuf dstage(ctx, std::max<uint64_t>(stage, 64)), dany(ctx, nc * 4 + 64);
  zero(ctx, dany.p, nc * 4 + 64);
  for (size_t i = 0; i < nc; ++i) {
    uint8_t *st = dstage.as<uint8_t>();
    if (bits_off[i] >= 0)
      HIPX(qmk_qmcb_bits(r->cols[i].nulls.as<uint8_t>(), 0, row0, nrows, st + bits_off[i], dany.as<uint32_t>() + i,
                         ctx->lane().stream));
    if (bool_off[i] >= 0)
      HIPX(qmk_qmcb_bits(r->cols[i].data.as<uint8_t>(), 1, row0, nrows, st + bool_off[i], nullptr,
                         ctx->lane().stream));
  }
  std::vector<uint32_t> any(nc, 0);
  if (nc) d2h(ctx, any.data(), dany.p, nc);
  // layout: the header, then per column [NULL bitmap] data (src/columnar.ts:523-613)
  uint64_t total = 10;
  for (const OutCol &c : r->cols) {
    if (c.name.size() > 0xFFFF) fail(QM_ERR_UNSUPPORTED, "column name too long");
    total += 2 + c.name.size() + 2;
  }
  std::vector<uint64_t> at_bits(nc), at_data(nc);
  for (size_t i = 0; i < nc; ++i) {
    if (any[i]) { at_bits[i] = total; total += nb; }
    at_data[i] = total;
    total += qmcb_data_bytes(r->cols[i].dtype, nrows);
  }
  uint8_t *buf = pin_take(total);
  uint8_t *w = buf;
  auto put32 = [&](uint32_t v) { memcpy(w, &v, 4); w += 4; };
  auto put16 = [&](uint16_t v) { memcpy(w, &v, 2); w += 2; };
  put32(0x42434D51u);
  put32((uint32_t)nrows);
  put16((uint16_t)nc);
  for (size_t i = 0; i < nc; ++i) {
    const OutCol &c = r->cols[i];
    put16((uint16_t)c.name.size());
    memcpy(w, c.name.data(), c.name.size());
    w += c.name.size();
    *w++ = (uint8_t)qmcb_tag(c.dtype);
    *w++ = any[i] ? 1 : 0;
  }
  // the body by DMA straight into the pinned buffer
  hipStream_t st = ctx->lane().stream;
  for (size_t i = 0; i < nc; ++i) {
    const OutCol &c = r->cols[i];
    if (any[i] && nb) HIPX(hipMemcpyAsync(buf + at_bits[i], dstage.as<uint8_t>() + bits_off[i], nb, hipMemcpyDeviceToHost, st));
    const uint64_t db = qmcb_data_bytes(c.dtype, nrows);
    if (!db) continue;
    const uint8_t *src = bool_off[i] >= 0 ? dstage.as<uint8_t>() + bool_off[i]
                                          : static_cast<const uint8_t *>(c.data.p) + row0 * (uint64_t)dt_size(c.dtype);
    HIPX(hipMemcpyAsync(buf + at_data[i], src, db, hipMemcpyDeviceToHost, st));
  }
  stream_sync(ctx);
  *out = buf;
  *size = total;
  API_END
}

extern "C" int qm_qmcb_release(const uint8_t *p) {
  API_BEGIN
  if (!p) return QM_OK;
  std::lock_guard<std::mutex> g(g_pin_mu);
  auto it = g_pin_live.find(p);
  if (it == g_pin_live.end()) fail(QM_ERR_INVALID_ARG, "not a QMCB buffer of this library (or released twice)");
  g_pin_free.emplace(it->second, const_cast<uint8_t *>(p));
  g_pin_live.erase(it);
  API_END
}

extern "C" int qm_result_stats(qm_result *r, qm_stats *out) {
  API_BEGIN
  if (!r || !out) fail(QM_ERR_INVALID_ARG, "null argument");
  *out = r->stats;
  API_END
}

extern "C" int qm_result_close(qm_result *r) {
  if (!r) return QM_OK;
  // a result is complete when handed out (its call drained its lane), so its
  // buffers go straight back to the pool
  delete r;
  return QM_OK;
}

// ---- asynchronous batches: qm_batch_exec on a worker thread ----
struct qm_batch {
  qm_ctx *ctx = nullptr;
  qm_table *table = nullptr;
  std::vector<qm_plan *> plans;
  std::vector<qm_result *> results;
  std::thread worker;
  std::atomic<int> done{0};
  int status = QM_OK;
  std::string error;
  bool taken = false;
};

extern "C" int qm_batch_submit(qm_ctx *ctx, qm_plan *const *plans, uint32_t nplans, qm_table *table,
                               qm_batch **out) {
  API_BEGIN
  if (!out) fail(QM_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (!ctx || !table || (nplans && !plans)) fail(QM_ERR_INVALID_ARG, "null argument");
  for (uint32_t i = 0; i < nplans; ++i)
    if (!plans[i]) fail(QM_ERR_INVALID_ARG, "null plan %u", i);
  std::unique_ptr<qm_batch> b(new qm_batch());
  b->ctx = ctx;
  b->table = table;
  b->plans.assign(plans, plans + nplans);
  b->results.assign(nplans, nullptr);
  qm_batch *bp = b.get();
  bp->worker = std::thread([bp] {
    const int rc = qm_batch_exec(bp->ctx, bp->plans.data(), (uint32_t)bp->plans.size(), bp->table,
                                 bp->results.data());
    bp->status = rc;
    if (rc != QM_OK) bp->error = g_last_error;  // this worker thread's message
    bp->done.store(1, std::memory_order_release);
  });
  *out = b.release();
  API_END
}

extern "C" int qm_batch_poll(qm_batch *b, int *done) {
  API_BEGIN
  if (!b || !done) fail(QM_ERR_INVALID_ARG, "null argument");
  *done = b->done.load(std::memory_order_acquire);
  API_END
}

extern "C" int qm_batch_wait(qm_batch *b, qm_result **results_out) {
  API_BEGIN
  if (!b) fail(QM_ERR_INVALID_ARG, "null argument");
  if (b->worker.joinable()) b->worker.join();
  if (b->status != QM_OK) return set_error(b->status, b->error);
  if (b->taken) fail(QM_ERR_INVALID_ARG, "the batch's results were already taken");
  if (!results_out && !b->results.empty()) fail(QM_ERR_INVALID_ARG, "null argument");
  for (size_t i = 0; i < b->results.size(); ++i) results_out[i] = b->results[i];
  b->taken = true;
  API_END
}

extern "C" int qm_batch_close(qm_batch *b) {
  API_BEGIN
  if (!b) return QM_OK;
  if (b->worker.joinable()) b->worker.join();
  if (!b->taken)
    for (qm_result *r : b->results) qm_result_close(r);
  delete b;
  API_END
}

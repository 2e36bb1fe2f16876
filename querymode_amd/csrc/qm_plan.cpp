// qm_plan.cpp — the binary plan reader and the filter-literal coercions.
//
// Plan layout (little-endian; str = u16 length + UTF-8 bytes):
//   u32 magic 0x4C504D51 ("QMPL"), u16 version (1, or 2 with the SQL post-stages)
//   u16 n, n x Filter                         filters (AND)
//   u16 g, g x { u16 n, n x Filter }          filterGroups (OR of ANDs)
//   u16 c, c x { str alias, u8 kind(1=multiply), str a, str b }
//   u16 k, k x str                            groupBy
//   u16 a, a x { u8 fn, str column, str alias }   ("" alias = fn_column)
//   u8 has_join [str leftKey, str rightKey, u8 type, u16 n x Filter, u16 g x {u16 n, n x Filter}]
//   u8 has_sort [str column, u8 desc]
//   i64 limit (-1 = none), i64 offset
//   u16 p, p x str                            projections
//   version 2 only (src/sql/executor.ts:30-170):
//   u8 has_distinct [u16 d, d x str]          DISTINCT columns (d = 0: every output column)
//   u16 o, o x { str column, u8 desc }        multi-column ORDER BY
//   u16 h, h x HavingOp                       HAVING, postfix:
//     u8 1 cmp: str col, u8 cmp, lit   2 cmp_col: str col, u8 cmp, str col2
//     3 between: str col, u8 negated, lit lo, lit hi   4 in: str col, u8 negated, u16 n, n x (u8 0 = NULL | lit)
//     5 is_null: str col, u8 negated   6 and   7 or   8 not      (cmp: 0 eq 1 ne 2 lt 3 le 4 gt 5 ge)
//   Filter = str column, u8 op, u8 vkind, value:
//     vkind 1: f64 | 2: i64 | 3: u32 n, n x (u8 kind, 8 bytes) | 4: 2 x (u8 kind, 8 bytes)
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include "qm_host.h"

namespace qm {

[[noreturn]] void fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw Error{code, std::string(buf)};
}

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) fail(QM_ERR_QUERY_FAILED, "HIP error %s in %s", hipGetErrorString(e), what);
}

namespace {
struct Reader {
  const uint8_t *p;
  size_t n, pos = 0;
  void need(size_t k) {
    if (pos + k > n) fail(QM_ERR_INVALID_FORMAT, "plan truncated at byte %zu", pos);
  }
  uint8_t u8() { need(1); return p[pos++]; }
  uint16_t u16() { need(2); uint16_t v; memcpy(&v, p + pos, 2); pos += 2; return v; }
  uint32_t u32() { need(4); uint32_t v; memcpy(&v, p + pos, 4); pos += 4; return v; }
  int64_t i64() { need(8); int64_t v; memcpy(&v, p + pos, 8); pos += 8; return v; }
  double f64() { need(8); double v; memcpy(&v, p + pos, 8); pos += 8; return v; }
  std::string str() { uint16_t l = u16(); need(l); std::string s((const char *)p + pos, l); pos += l; return s; }
  Lit lit() {
    Lit L;
    L.kind = u8();
    if (L.kind == LIT_F64) L.f = f64();
    else if (L.kind == LIT_I64) L.i = i64();
    else fail(QM_ERR_INVALID_FORMAT, "bad literal kind %d", L.kind);
    return L;
  }
  Filter filter() {
    Filter f;
    f.col = str();
    f.op = u8();
    if (f.op > OP_NOT_LIKE) fail(QM_ERR_INVALID_FILTER, "unknown filter op %d", f.op);
    f.vkind = u8();
    switch (f.vkind) {
      case LIT_NONE: break;
      case LIT_F64: f.v.kind = LIT_F64; f.v.f = f64(); break;
      case LIT_I64: f.v.kind = LIT_I64; f.v.i = i64(); break;
      case LIT_LIST: { uint32_t k = u32(); for (uint32_t i = 0; i < k; ++i) f.list.push_back(lit()); break; }
      case LIT_PAIR: f.list.push_back(lit()); f.list.push_back(lit()); break;
      default: fail(QM_ERR_INVALID_FORMAT, "bad filter value kind %d", f.vkind);
    }
    return f;
  }
  std::vector<Filter> filters() {
    std::vector<Filter> v;
    uint16_t k = u16();
    for (uint16_t i = 0; i < k; ++i) v.push_back(filter());
    return v;
  }
  std::vector<std::vector<Filter>> groups() {
    std::vector<std::vector<Filter>> g;
    uint16_t k = u16();
    for (uint16_t i = 0; i < k; ++i) g.push_back(filters());
    return g;
  }
};
}  // namespace

Plan parse_plan(const uint8_t *p, size_t n) {
  Reader r{p, n};
  if (r.u32() != 0x4C504D51u) fail(QM_ERR_INVALID_FORMAT, "bad plan magic");
  const uint16_t version = r.u16();
  if (version != 1 && version != 2) fail(QM_ERR_INVALID_FORMAT, "unsupported plan version %u", version);
  Plan pl;
  pl.filters = r.filters();
  pl.groups = r.groups();
  uint16_t nc = r.u16();
  for (uint16_t i = 0; i < nc; ++i) {
    Computed c;
    c.alias = r.str();
    c.kind = r.u8();
    c.a = r.str();
    c.b = r.str();
    if (c.kind != 1) fail(QM_ERR_UNSUPPORTED, "computed column '%s': only products run on the device", c.alias.c_str());
    pl.computed.push_back(c);
  }
  uint16_t ng = r.u16();
  for (uint16_t i = 0; i < ng; ++i) pl.groupby.push_back(r.str());
  uint16_t na = r.u16();
  for (uint16_t i = 0; i < na; ++i) {
    Agg a;
    a.fn = r.u8();
    a.col = r.str();
    a.alias = r.str();
    if (a.fn > AGG_PERCENTILE) fail(QM_ERR_INVALID_AGGREGATE, "unknown aggregate fn %d", a.fn);
    static const char *names[] = {"sum", "avg", "min", "max", "count", "count_distinct", "stddev",
                                  "variance", "median", "percentile"};
    // alias defaults to fn_column (src/partial-agg.ts:136-138)
    if (a.alias.empty()) a.alias = std::string(names[a.fn]) + "_" + a.col;
    pl.aggs.push_back(a);
  }
  pl.has_join = r.u8() != 0;
  if (pl.has_join) {
    pl.lkey = r.str();
    pl.rkey = r.str();
    pl.jtype = r.u8();
    pl.rfilters = r.filters();
    pl.rgroups = r.groups();
  }
  pl.has_sort = r.u8() != 0;
  if (pl.has_sort) {
    pl.sort_col = r.str();
    pl.desc = r.u8() != 0;
  }
  pl.limit = r.i64();
  pl.offset = r.i64();
  uint16_t np = r.u16();
  for (uint16_t i = 0; i < np; ++i) pl.proj.push_back(r.str());
  if (version >= 2) {
    // DISTINCT columns, multi-column ORDER BY, HAVING program
    pl.has_distinct = r.u8() != 0;
    if (pl.has_distinct) {
      uint16_t nd = r.u16();
      for (uint16_t i = 0; i < nd; ++i) pl.distinct.push_back(r.str());
    }
    uint16_t no = r.u16();
    for (uint16_t i = 0; i < no; ++i) {
      OrderKey k;
      k.col = r.str();
      k.desc = r.u8() != 0;
      pl.order_by.push_back(k);
    }
    uint16_t nh = r.u16();
    int depth = 0;
    for (uint16_t i = 0; i < nh; ++i) {
      HavingOp h;
      h.kind = r.u8();
      switch (h.kind) {
        case HK_CMP: h.col = r.str(); h.cmp = r.u8(); h.a = r.lit(); ++depth; break;
        case HK_CMP_COL: h.col = r.str(); h.cmp = r.u8(); h.col2 = r.str(); ++depth; break;
        case HK_BETWEEN: h.col = r.str(); h.negated = r.u8() != 0; h.a = r.lit(); h.b = r.lit(); ++depth; break;
        case HK_IN: {
          h.col = r.str();
          h.negated = r.u8() != 0;
          uint16_t k = r.u16();
          for (uint16_t j = 0; j < k; ++j) {
            r.need(1);
            if (r.p[r.pos] == LIT_NONE) { ++r.pos; h.list.push_back(Lit()); }  // a NULL in the list
            else h.list.push_back(r.lit());
          }
          ++depth;
          break;
        }
        case HK_ISNULL: h.col = r.str(); h.negated = r.u8() != 0; ++depth; break;
        case HK_AND: case HK_OR: if (depth < 2) fail(QM_ERR_INVALID_FORMAT, "HAVING program underflow"); --depth; break;
        case HK_NOT: if (depth < 1) fail(QM_ERR_INVALID_FORMAT, "HAVING program underflow"); break;
        default: fail(QM_ERR_INVALID_FORMAT, "unknown HAVING op %d", h.kind);
      }
      if (h.cmp > 5) fail(QM_ERR_INVALID_FORMAT, "unknown HAVING comparison %d", h.cmp);
      pl.having.push_back(h);
    }
    if (nh && depth != 1) fail(QM_ERR_INVALID_FORMAT, "HAVING program leaves %d values", depth);
    if (!pl.having.empty() && pl.aggs.empty()) fail(QM_ERR_INVALID_AGGREGATE, "HAVING without aggregates");
    if (pl.has_distinct && !pl.aggs.empty())
      fail(QM_ERR_UNSUPPORTED, "DISTINCT together with aggregates is not on the device path");
    if (!pl.order_by.empty() && pl.has_sort)
      fail(QM_ERR_INVALID_FORMAT, "a plan has either sortColumn or a multi-column ORDER BY");
  }
  if (r.pos != n) fail(QM_ERR_INVALID_FORMAT, "trailing bytes in plan");
  if (pl.offset < 0) fail(QM_ERR_INVALID_FORMAT, "negative offset");
  return pl;
}

// ===========================================================================
// Literal coercion
// ===========================================================================
namespace {

const double TWO63 = 9223372036854775808.0;  // 2^63

// JS ToInt32 (what a JS number becomes as a WASM i32 argument).
int64_t js_to_int32(double x) {
  if (!std::isfinite(x)) return 0;
  double t = std::trunc(x);
  double m = std::fmod(t, 4294967296.0);  // exact
  if (m < 0) m += 4294967296.0;
  uint32_t u = (uint32_t)m;
  return (int32_t)u;
}

// BigInt.asIntN(64, BigInt(trunc(x))) for finite x (what safeBigInt(x)
// becomes as a WASM i64 argument).
int64_t js_to_int64_wrap(double x) {
  double t = std::trunc(x);
  if (t >= -TWO63 && t < TWO63) return (int64_t)t;
  double m = std::fmod(t, 18446744073709551616.0);  // exact, |m| < 2^64
  if (m < 0) m += 18446744073709551616.0;
  uint64_t u;
  if (m >= TWO63) u = (uint64_t)(m - TWO63) + (1ull << 63);
  else u = (uint64_t)m;
  return (int64_t)u;
}

// Integer interval [lo, hi] with "empty" / "everything" states.
struct IRange {
  bool empty = false;
  int64_t lo = INT64_MIN, hi = INT64_MAX;
  void meet(const IRange &o) {
    if (o.empty) { empty = true; return; }
    if (o.lo > lo) lo = o.lo;
    if (o.hi < hi) hi = o.hi;
    if (lo > hi) empty = true;
  }
};

IRange all() { return IRange(); }
IRange none() { IRange r; r.empty = true; return r; }
IRange ge_i(int64_t v) { IRange r; r.lo = v; return r; }
IRange le_i(int64_t v) { IRange r; r.hi = v; return r; }

// integer v  OP  real t (exact mathematical comparison, JS number vs
// integer-valued number / bigint semantics)
IRange ge_real(double t) {  // v >= t
  if (std::isnan(t)) return none();
  if (t <= -TWO63) return all();
  if (t >= TWO63) return none();
  return ge_i((int64_t)std::ceil(t));
}
IRange gt_real(double t) {  // v > t
  if (std::isnan(t)) return none();
  if (t < -TWO63) return all();
  if (t >= TWO63) return none();
  double f = std::floor(t);
  return ge_i((int64_t)f + 1);
}
IRange le_real(double t) {  // v <= t
  if (std::isnan(t)) return none();
  if (t >= TWO63) return all();
  if (t < -TWO63) return none();
  return le_i((int64_t)std::floor(t));
}
IRange lt_real(double t) {  // v < t
  if (std::isnan(t)) return none();
  if (t >= TWO63) return all();
  if (t <= -TWO63) return none();
  return le_i((int64_t)std::ceil(t) - 1);
}
IRange eq_real(double t) {
  if (!std::isfinite(t) || std::trunc(t) != t || t < -TWO63 || t >= TWO63) return none();
  IRange r; r.lo = r.hi = (int64_t)t; return r;
}

// an integer-column bound: exact int or real
struct IBound { bool real; double f; int64_t i; };

IRange cmp_int(int op, const IBound &b) {
  if (!b.real) {
    IRange r;
    switch (op) {
      case OP_EQ: r.lo = r.hi = b.i; return r;
      case OP_LT: if (b.i == INT64_MIN) return none(); return le_i(b.i - 1);
      case OP_LTE: return le_i(b.i);
      case OP_GT: if (b.i == INT64_MAX) return none(); return ge_i(b.i + 1);
      case OP_GTE: return ge_i(b.i);
    }
  } else {
    switch (op) {
      case OP_EQ: return eq_real(b.f);
      case OP_LT: return lt_real(b.f);
      case OP_LTE: return le_real(b.f);
      case OP_GT: return gt_real(b.f);
      case OP_GTE: return ge_real(b.f);
    }
  }
  return none();
}

DPred blank() {
  DPred p;
  memset(&p, 0, sizeof p);
  return p;
}

// range (or its negation) -> device predicate
void set_int_range(DPred &p, const IRange &r, bool negate) {
  if (r.empty) {
    p.kind = negate ? PK_IS_NOT_NULL : PK_FALSE;  // !false = every non-null row
    return;
  }
  if (r.lo == INT64_MIN && r.hi == INT64_MAX) {
    p.kind = negate ? PK_FALSE : PK_IS_NOT_NULL;
    return;
  }
  p.kind = PK_INT_RANGE;
  p.ilo = r.lo;
  p.ihi = r.hi;
  p.negate = negate ? 1 : 0;
}

bool is_int_dt(int dt) { return dt == DT_I8 || dt == DT_I16 || dt == DT_I32 || dt == DT_I64; }
bool is_float_dt(int dt) { return dt == DT_F64 || dt == DT_F32; }

// one literal as an integer-column bound, per the reference's path
IBound int_bound(int dt, const Lit &L, bool wasm_path) {
  IBound b{false, 0, 0};
  if (L.kind == LIT_I64) { b.i = L.i; return b; }            // bigint: JS path, exact
  if (wasm_path && dt == DT_I32) { b.i = js_to_int32(L.f); return b; }
  if (wasm_path && dt == DT_I64 && std::isfinite(L.f)) { b.i = js_to_int64_wrap(L.f); return b; }
  if (dt == DT_I64 && std::isfinite(L.f)) {
    // JS fallback coerceCompare(bigint, number): BigInt(Math.trunc(x)), exact
    double t = std::trunc(L.f);
    b.real = true; b.f = t;  // integral real; saturating comparisons below
    return b;
  }
  b.real = true; b.f = L.f;
  return b;
}

}  // namespace

BoundPred coerce_filter(const Filter &f, int dt) {
  BoundPred bp;
  bp.p = blank();
  DPred &p = bp.p;
  if (f.op == OP_IS_NULL) { p.kind = PK_IS_NULL; return bp; }
  if (f.op == OP_IS_NOT_NULL) { p.kind = PK_IS_NOT_NULL; return bp; }
  if (f.op == OP_LIKE || f.op == OP_NOT_LIKE)
    fail(QM_ERR_UNSUPPORTED, "filter %s: LIKE needs a string column; string columns are not on the device path",
         f.col.c_str());
  if (dt == DT_BOOL)
    fail(QM_ERR_UNSUPPORTED, "filter %s: comparisons on boolean columns are not on the device path", f.col.c_str());
  const bool scalar_op = f.op <= OP_GTE;
  if (scalar_op && f.vkind != LIT_F64 && f.vkind != LIT_I64)
    fail(QM_ERR_INVALID_FILTER, "filter %s: op needs a scalar value", f.col.c_str());
  if ((f.op == OP_BETWEEN || f.op == OP_NOT_BETWEEN) && (f.vkind != LIT_PAIR || f.list.size() != 2))
    fail(QM_ERR_INVALID_FILTER, "filter %s: between needs [low, high]", f.col.c_str());
  if ((f.op == OP_IN || f.op == OP_NOT_IN) && f.vkind != LIT_LIST)
    fail(QM_ERR_INVALID_FILTER, "filter %s: in needs a list", f.col.c_str());

  if (is_int_dt(dt)) {
    // WASM SIMD path: int32/int64 column, number literal (src/operators.ts:487-499,
    // 502-515); otherwise the JS matchesFilter path (src/decode.ts:485-532).
    if (scalar_op) {
      bool wasm = (f.v.kind == LIT_F64) && (dt == DT_I32 || dt == DT_I64);
      IBound b = int_bound(dt, f.v, wasm);
      if (f.op == OP_NEQ) { set_int_range(p, cmp_int(OP_EQ, b), true); return bp; }
      set_int_range(p, cmp_int(f.op, b), false);
      return bp;
    }
    if (f.op == OP_BETWEEN || f.op == OP_NOT_BETWEEN) {
      const Lit &a = f.list[0], &c = f.list[1];
      bool wasm = a.kind == LIT_F64 && c.kind == LIT_F64 && (dt == DT_I32 || dt == DT_I64);
      if (dt == DT_I64 && wasm && !(std::isfinite(a.f) && std::isfinite(c.f))) wasm = false;
      IBound lo = int_bound(dt, a, wasm), hi = int_bound(dt, c, wasm);
      if (f.op == OP_BETWEEN) {
        IRange r = cmp_int(OP_GTE, lo);
        r.meet(cmp_int(OP_LTE, hi));
        set_int_range(p, r, false);
        return bp;
      }
      // NOT BETWEEN: v < lo || v > hi (src/decode.ts:516-522; aggregates.zig:851-911)
      IRange below = cmp_int(OP_LT, lo), above = cmp_int(OP_GT, hi);
      if (below.empty && above.empty) { p.kind = PK_FALSE; return bp; }
      if (below.empty) { set_int_range(p, above, false); return bp; }
      if (above.empty) { set_int_range(p, below, false); return bp; }
      // both non-empty: below = (-inf, x], above = [y, +inf) -> complement of (x, y)
      if (below.hi >= above.lo - 1) { p.kind = PK_IS_NOT_NULL; return bp; }
      IRange mid; mid.lo = below.hi + 1; mid.hi = above.lo - 1;
      set_int_range(p, mid, true);
      return bp;
    }
    // IN / NOT IN: Set membership with cross-type equivalents (src/decode.ts:537-556)
    for (const Lit &L : f.list) {
      if (L.kind == LIT_I64) {
        // a bigint matches a number-valued column only if it is a safe integer
        if (dt != DT_I64 && (L.i > 9007199254740991LL || L.i < -9007199254740991LL)) continue;
        bp.ilist.push_back(L.i);
      } else {
        if (std::isfinite(L.f) && std::trunc(L.f) == L.f && L.f >= -TWO63 && L.f < TWO63)
          bp.ilist.push_back((int64_t)L.f);
      }
    }
    std::sort(bp.ilist.begin(), bp.ilist.end());
    bp.ilist.erase(std::unique(bp.ilist.begin(), bp.ilist.end()), bp.ilist.end());
    p.negate = (f.op == OP_NOT_IN) ? 1 : 0;
    if (bp.ilist.empty()) { p.kind = p.negate ? PK_IS_NOT_NULL : PK_FALSE; p.negate = 0; return bp; }
    p.kind = PK_INT_IN;
    return bp;
  }

  if (is_float_dt(dt)) {
    if (scalar_op || f.op == OP_BETWEEN || f.op == OP_NOT_BETWEEN) {
      if (f.v.kind == LIT_I64 || (f.list.size() == 2 && (f.list[0].kind == LIT_I64 || f.list[1].kind == LIT_I64)))
        fail(QM_ERR_UNSUPPORTED, "filter %s: bigint literal on a float column", f.col.c_str());
    }
    p.kind = PK_F64_RANGE;
    p.flo = -INFINITY; p.fhi = INFINITY;
    const double t = f.v.f;
    if (scalar_op && std::isnan(t)) {
      // every comparison with NaN is false, except !== (src/decode.ts:494-499)
      p.kind = (f.op == OP_NEQ) ? PK_IS_NOT_NULL : PK_FALSE;
      return bp;
    }
    switch (f.op) {
      case OP_EQ: p.flo = p.fhi = t; return bp;
      case OP_NEQ: p.flo = p.fhi = t; p.negate = 1; p.nan_result = 1; return bp;
      case OP_LT: p.fhi = t; p.hi_strict = 1; return bp;
      case OP_LTE: p.fhi = t; return bp;
      case OP_GT: p.flo = t; p.lo_strict = 1; return bp;
      case OP_GTE: p.flo = t; return bp;
      case OP_BETWEEN: {
        double a = f.list[0].f, c = f.list[1].f;
        if (std::isnan(a) || std::isnan(c)) { p.kind = PK_FALSE; return bp; }
        p.flo = a; p.fhi = c;
        return bp;
      }
      case OP_NOT_BETWEEN: {
        double a = f.list[0].f, c = f.list[1].f;
        bool na = std::isnan(a), nc = std::isnan(c);
        if (na && nc) { p.kind = PK_FALSE; return bp; }
        if (na) { p.flo = c; p.lo_strict = 1; return bp; }       // v > hi
        if (nc) { p.fhi = a; p.hi_strict = 1; return bp; }       // v < lo
        p.flo = a; p.fhi = c; p.negate = 1; p.nan_result = 0;    // !(lo <= v <= hi), NaN -> false
        return bp;
      }
      default: break;
    }
    // IN / NOT IN with SameValueZero (NaN matches NaN, -0 matches +0)
    for (const Lit &L : f.list) {
      if (L.kind == LIT_I64) {
        if (L.i > 9007199254740991LL || L.i < -9007199254740991LL) continue;
        bp.flist.push_back((double)L.i);
      } else if (std::isnan(L.f)) {
        p.list_has_nan = 1;
      } else {
        bp.flist.push_back(L.f == 0.0 ? 0.0 : L.f);
      }
    }
    std::sort(bp.flist.begin(), bp.flist.end());
    bp.flist.erase(std::unique(bp.flist.begin(), bp.flist.end()), bp.flist.end());
    p.kind = PK_F64_IN;
    p.negate = (f.op == OP_NOT_IN) ? 1 : 0;
    return bp;
  }
  fail(QM_ERR_UNSUPPORTED, "filter %s: column type not on the device path", f.col.c_str());
}

}  // namespace qm

/*
 * qm_oracle.c — CPU restatement of the reference's operator hot path, used
 * ONLY as a test oracle and as the timed CPU baseline ("port"). Nothing in
 * the product links or calls it.
 *
 * It follows the reference's own algorithm, not the GPU design:
 *   - per-page filtering into ascending u32 index lists, one list per AND
 *     predicate, then sorted intersection; OR groups unioned
 *     (src/operators.ts:439-547, wasm/src/wasm/aggregates.zig:605-1065);
 *   - per-value predicate semantics of the WASM numeric path
 *     (src/operators.ts:550-666: ToInt32 / BigInt.asIntN(64) literal
 *     coercion, IEEE f64 compares) and of the JS fallback matchesFilter
 *     (src/decode.ts:476-556: coerceCompare, mixed number/bigint compares,
 *     Set membership with cross-type equivalents);
 *   - hash aggregation with partial-agg semantics (src/partial-agg.ts:38-352):
 *     first-seen group order, COUNT(*) vs COUNT(col), NaN never updating
 *     MIN/MAX, empty -> null, bigAvg for int64 columns;
 *   - the in-memory hash join, build = right side, probe order then build
 *     insertion order (src/operators.ts:2349-2436, 2673-2713);
 *   - top-K: the reference's binary heap (src/operators.ts:1716-1841) for
 *     timing, and the declared total order (value, then row id) for parity.
 *
 * Deliberate differences (the build's contract): FP64 SUM is accumulated
 * with compensated long-double summation (near exactly rounded) instead of
 * the reference's left-to-right double sum; int32 SUM is an exact integer.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { DT_F64 = 0, DT_I64 = 1, DT_BOOL = 3, DT_I32 = 6, DT_F32 = 7, DT_I8 = 16, DT_I16 = 17 };
enum { OP_EQ = 0, OP_NEQ, OP_LT, OP_LTE, OP_GT, OP_GTE, OP_BETWEEN, OP_NOT_BETWEEN, OP_IN, OP_NOT_IN,
       OP_IS_NULL, OP_IS_NOT_NULL };
enum { L_NUM = 1, L_BIG = 2 };
enum { AGG_SUM = 0, AGG_AVG, AGG_MIN, AGG_MAX, AGG_COUNT };

typedef struct { const void *data; const uint8_t *nulls; int dtype; } oc_col;

typedef struct {
  int col, op, group;          /* group 0 = AND, g >= 1 = OR group g */
  int kind;                    /* literal kind for scalar ops */
  double f; int64_t i;
  int n;                       /* list / pair length */
  const int *lkinds; const double *lf; const int64_t *li;
} oc_pred;

/* ---------------- JS values: number or bigint ---------------- */
typedef struct { int big; double f; int64_t i; } jsv;

static int is_null_at(const oc_col *c, uint64_t r) { return c->nulls && ((c->nulls[r >> 3] >> (r & 7)) & 1); }

/* the decoded JS value of a cell: int64 columns decode as bigint, the other
 * numeric columns as number */
static jsv cell(const oc_col *c, uint64_t r) {
  jsv v = {0, 0, 0};
  switch (c->dtype) {
    case DT_F64: v.f = ((const double *)c->data)[r]; break;
    case DT_F32: v.f = ((const float *)c->data)[r]; break;
    case DT_I32: v.f = ((const int32_t *)c->data)[r]; break;
    case DT_I16: v.f = ((const int16_t *)c->data)[r]; break;
    case DT_I8: v.f = ((const int8_t *)c->data)[r]; break;
    case DT_BOOL: v.f = ((const uint8_t *)c->data)[r]; break;
    case DT_I64: v.big = 1; v.i = ((const int64_t *)c->data)[r]; break;
  }
  return v;
}

/* JS comparison a <=> b: returns -1/0/1, or 2 when unordered (NaN). Mixed
 * number/bigint compares are mathematical (long double holds both exactly). */
static int js_cmp(jsv a, jsv b) {
  if (a.big && b.big) return a.i < b.i ? -1 : (a.i > b.i ? 1 : 0);
  if (!a.big && !b.big) {
    if (isnan(a.f) || isnan(b.f)) return 2;
    return a.f < b.f ? -1 : (a.f > b.f ? 1 : 0);
  }
  long double x = a.big ? (long double)a.i : (long double)a.f;
  long double y = b.big ? (long double)b.i : (long double)b.f;
  if ((!a.big && isnan(a.f)) || (!b.big && isnan(b.f))) return 2;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* coerceCompare (src/decode.ts:476-481) */
static void to_big(jsv *x) {
  /* BigInt(Math.trunc(x)); values beyond int64 stay an (integral) number,
   * which compares mathematically and equals no int64 */
  double t = trunc(x->f);
  if (t >= -9223372036854775808.0 && t < 9223372036854775808.0) { x->big = 1; x->i = (int64_t)t; }
  else x->f = t;
}
static void coerce(jsv *a, jsv *b) {
  if (a->big && !b->big) {
    if (isfinite(b->f)) to_big(b);
  } else if (!a->big && b->big) {
    if (isfinite(a->f)) to_big(a);
  }
}

/* === after coercion: only same-type values are ever equal */
static int js_strict_eq(jsv a, jsv b) {
  if (a.big != b.big) return 0;
  return a.big ? a.i == b.i : a.f == b.f;
}

static jsv lit_of(int kind, double f, int64_t i) {
  jsv v; v.big = (kind == L_BIG); v.f = f; v.i = i; return v;
}

/* matchesFilter for non-null values (src/decode.ts:485-532) */
static int js_matches(jsv val, const oc_pred *p) {
  switch (p->op) {
    case OP_EQ: case OP_NEQ: case OP_LT: case OP_LTE: case OP_GT: case OP_GTE: {
      jsv a = val, b = lit_of(p->kind, p->f, p->i);
      coerce(&a, &b);
      if (p->op == OP_EQ) return js_strict_eq(a, b);
      if (p->op == OP_NEQ) return !js_strict_eq(a, b);
      int c = js_cmp(a, b);
      if (c == 2) return 0;
      if (p->op == OP_LT) return c < 0;
      if (p->op == OP_LTE) return c <= 0;
      if (p->op == OP_GT) return c > 0;
      return c >= 0;
    }
    case OP_BETWEEN: case OP_NOT_BETWEEN: {
      jsv a0 = val, lo = lit_of(p->lkinds[0], p->lf[0], p->li[0]);
      coerce(&a0, &lo);
      jsv a = val, hi = lit_of(p->lkinds[1], p->lf[1], p->li[1]);
      coerce(&a, &hi);
      int c1 = js_cmp(a, lo), c2 = js_cmp(a, hi);
      if (p->op == OP_BETWEEN) return c1 != 2 && c2 != 2 && c1 >= 0 && c2 <= 0;
      return (c1 != 2 && c1 < 0) || (c2 != 2 && c2 > 0);
    }
    case OP_IN: case OP_NOT_IN: {
      int hit = 0;
      for (int k = 0; k < p->n && !hit; ++k) {
        jsv L = lit_of(p->lkinds[k], p->lf[k], p->li[k]);
        if (!val.big) {
          if (!L.big) hit = (isnan(val.f) && isnan(L.f)) || val.f == L.f;  /* SameValueZero */
          else hit = (L.i <= 9007199254740991LL && L.i >= -9007199254740991LL) && (double)L.i == val.f;
        } else {
          if (L.big) hit = L.i == val.i;
          else hit = isfinite(L.f) && trunc(L.f) == L.f && fabs(L.f) < 9.3e18 && (long double)L.f == (long double)val.i;
        }
      }
      return p->op == OP_IN ? hit : !hit;
    }
  }
  return 1;
}

static int32_t to_int32(double x) {
  if (!isfinite(x)) return 0;
  long double t = truncl((long double)x);
  long double m = fmodl(t, 4294967296.0L);
  if (m < 0) m += 4294967296.0L;
  uint32_t u = (uint32_t)m;
  return (int32_t)u;
}
static int64_t as_int64(double x) {  /* BigInt.asIntN(64, BigInt(trunc(x))) */
  long double t = truncl((long double)x);
  long double m = fmodl(t, 18446744073709551616.0L);
  if (m < 0) m += 18446744073709551616.0L;
  uint64_t u = (uint64_t)m;
  return (int64_t)u;
}

static int cmp_op(int op, long double a, long double b) {
  switch (op) {
    case OP_EQ: return a == b;
    case OP_NEQ: return a != b;
    case OP_LT: return a < b;
    case OP_LTE: return a <= b;
    case OP_GT: return a > b;
    default: return a >= b;
  }
}

/* one cell against one predicate, choosing the reference's path */
static int match_cell(const oc_col *c, uint64_t r, const oc_pred *p) {
  int isn = is_null_at(c, r);
  if (p->op == OP_IS_NULL) return isn;
  if (p->op == OP_IS_NOT_NULL) return !isn;
  if (isn) return 0;  /* the build contract: NULL never matches */
  const int numeric_wasm = c->dtype == DT_F64 || c->dtype == DT_F32 || c->dtype == DT_I32 || c->dtype == DT_I64;
  if (p->op <= OP_GTE && p->kind == L_NUM && numeric_wasm && !(c->dtype == DT_I64 && !isfinite(p->f))) {
    /* wasmFilterNumeric (src/operators.ts:550-608) */
    if (c->dtype == DT_I32) return cmp_op(p->op, ((const int32_t *)c->data)[r], to_int32(p->f));
    if (c->dtype == DT_I64) {
      int64_t v = ((const int64_t *)c->data)[r], t = as_int64(p->f);
      return p->op == OP_EQ ? v == t : p->op == OP_NEQ ? v != t : p->op == OP_LT ? v < t
           : p->op == OP_LTE ? v <= t : p->op == OP_GT ? v > t : v >= t;
    }
    double v = c->dtype == DT_F64 ? ((const double *)c->data)[r] : ((const float *)c->data)[r];
    double t = p->f;
    switch (p->op) {
      case OP_EQ: return v == t;
      case OP_NEQ: return v != t;
      case OP_LT: return v < t;
      case OP_LTE: return v <= t;
      case OP_GT: return v > t;
      default: return v >= t;
    }
  }
  if ((p->op == OP_BETWEEN || p->op == OP_NOT_BETWEEN) && p->lkinds[0] == L_NUM && p->lkinds[1] == L_NUM &&
      numeric_wasm && !(c->dtype == DT_I64 && !(isfinite(p->lf[0]) && isfinite(p->lf[1])))) {
    /* wasmFilterRange (src/operators.ts:613-666) */
    int in;
    if (c->dtype == DT_I32) {
      int32_t v = ((const int32_t *)c->data)[r], lo = to_int32(p->lf[0]), hi = to_int32(p->lf[1]);
      in = p->op == OP_BETWEEN ? (v >= lo && v <= hi) : (v < lo || v > hi);
    } else if (c->dtype == DT_I64) {
      int64_t v = ((const int64_t *)c->data)[r], lo = as_int64(p->lf[0]), hi = as_int64(p->lf[1]);
      in = p->op == OP_BETWEEN ? (v >= lo && v <= hi) : (v < lo || v > hi);
    } else {
      double v = c->dtype == DT_F64 ? ((const double *)c->data)[r] : ((const float *)c->data)[r];
      in = p->op == OP_BETWEEN ? (v >= p->lf[0] && v <= p->lf[1]) : (v < p->lf[0] || v > p->lf[1]);
    }
    return in;
  }
  return js_matches(cell(c, r), p);
}

/* ---------------- index lists ---------------- */
static uint64_t intersect(const uint32_t *a, uint64_t na, const uint32_t *b, uint64_t nb, uint32_t *out) {
  uint64_t i = 0, j = 0, k = 0;
  while (i < na && j < nb) {
    if (a[i] < b[j]) ++i;
    else if (a[i] > b[j]) ++j;
    else { out[k++] = a[i]; ++i; ++j; }
  }
  return k;
}
static uint64_t unite(const uint32_t *a, uint64_t na, const uint32_t *b, uint64_t nb, uint32_t *out) {
  uint64_t i = 0, j = 0, k = 0;
  while (i < na || j < nb) {
    if (j >= nb || (i < na && a[i] < b[j])) out[k++] = a[i++];
    else if (i >= na || b[j] < a[i]) out[k++] = b[j++];
    else { out[k++] = a[i]; ++i; ++j; }
  }
  return k;
}

/* applyAndFilters over one page: ascending in-page indices matching every
 * predicate of group `g` (src/operators.ts:468-547) */
static uint64_t page_and(const oc_col *cols, uint64_t base, uint32_t len, const oc_pred *preds, int np, int g,
                         uint32_t *acc, uint32_t *tmp, uint32_t *tmp2) {
  int first = 1;
  uint64_t n = len;
  for (uint32_t i = 0; i < len; ++i) acc[i] = i;
  for (int k = 0; k < np; ++k) {
    if (preds[k].group != g) continue;
    uint64_t m = 0;
    for (uint32_t i = 0; i < len; ++i)
      if (match_cell(&cols[preds[k].col], base + i, &preds[k])) tmp[m++] = i;
    if (first) { memcpy(acc, tmp, m * 4); n = m; first = 0; }
    else { n = intersect(acc, n, tmp, m, tmp2); memcpy(acc, tmp2, n * 4); }
    if (n == 0) break;
  }
  return n;
}

/* Filtered row ids (ascending) of the table; returns the count. */
uint64_t oc_filter(const oc_col *cols, uint64_t nrows, const oc_pred *preds, int npreds, int ngroups,
                   uint32_t page, uint64_t *out_rows) {
  if (page == 0) page = 10000;
  uint32_t *acc = malloc(sizeof(uint32_t) * page), *tmp = malloc(sizeof(uint32_t) * page);
  uint32_t *tmp2 = malloc(sizeof(uint32_t) * page), *orr = malloc(sizeof(uint32_t) * page);
  uint32_t *g = malloc(sizeof(uint32_t) * page), *t3 = malloc(sizeof(uint32_t) * page);
  uint64_t total = 0;
  for (uint64_t base = 0; base < nrows; base += page) {
    uint32_t len = (uint32_t)((nrows - base) < page ? (nrows - base) : page);
    uint64_t n = page_and(cols, base, len, preds, npreds, 0, acc, tmp, tmp2);
    if (ngroups > 0 && n > 0) {
      uint64_t no = 0;
      int any = 0;
      for (int gi = 1; gi <= ngroups; ++gi) {
        uint64_t m = page_and(cols, base, len, preds, npreds, gi, g, tmp, tmp2);
        if (m == 0) continue;
        if (!any) { memcpy(orr, g, m * 4); no = m; any = 1; }
        else { no = unite(orr, no, g, m, t3); memcpy(orr, t3, no * 4); }
      }
      n = any ? intersect(acc, n, orr, no, tmp2) : 0;
      memcpy(acc, tmp2, n * 4);
    }
    for (uint64_t i = 0; i < n; ++i) out_rows[total++] = base + acc[i];
  }
  free(acc); free(tmp); free(tmp2); free(orr); free(g); free(t3);
  return total;
}

/* ---------------- aggregation ---------------- */
typedef struct { int fn; int col_a; int col_b; } oc_agg;  /* col_a -1 = '*', col_b >= 0: product */

typedef struct {
  uint64_t count;
  long double s, c;      /* compensated number sum (Neumaier) */
  __int128 isum;         /* exact integer sum (int columns) */
  double fmin, fmax;     /* number path: start at +Inf / -Inf */
  int64_t bmin, bmax;    /* bigint path */
  int has_big;
  int nan, pinf, ninf;
} oc_state;

typedef struct {
  uint64_t hash;
  int used;
  uint64_t group;        /* index into group arrays (first-seen order) */
} slot_t;

static uint64_t mixh(uint64_t k) {
  k ^= k >> 31; k *= 0x7fb5d329728ea185ull; k ^= k >> 27; k *= 0x81dadef4bc2dd44dull; k ^= k >> 33; return k;
}

/* canonical key part: ints by value, f64 by String(v) identity, NULL flagged */
static uint64_t key_part(const oc_col *c, uint64_t r, int *isnull) {
  *isnull = is_null_at(c, r);
  if (*isnull) return 0;
  if (c->dtype == DT_F64 || c->dtype == DT_F32) {
    double v = c->dtype == DT_F64 ? ((const double *)c->data)[r] : ((const float *)c->data)[r];
    if (v == 0) v = 0;              /* String(-0) === "0" */
    uint64_t b; memcpy(&b, &v, 8);
    if (isnan(v)) b = 0x7FF8000000000000ull;
    return b;
  }
  jsv x = cell(c, r);
  return x.big ? (uint64_t)x.i : (uint64_t)(int64_t)x.f;
}

/* value of an aggregate input for a (left, right) row pair */
static jsv agg_value(const oc_col *cols, const int *side, const oc_agg *a, uint64_t lr, uint64_t rr, int *isnull) {
  uint64_t ra = side[a->col_a] ? rr : lr;
  *isnull = is_null_at(&cols[a->col_a], ra);
  jsv x = cell(&cols[a->col_a], ra);
  if (a->col_b >= 0) {
    uint64_t rb = side[a->col_b] ? rr : lr;
    if (is_null_at(&cols[a->col_b], rb)) *isnull = 1;
    jsv y = cell(&cols[a->col_b], rb);
    double xv = x.big ? (double)x.i : x.f, yv = y.big ? (double)y.i : y.f;
    x.big = 0; x.f = xv * yv;        /* JS number product */
  }
  return x;
}

static void feed(oc_state *s, const oc_agg *a, jsv v, int is_int_col) {
  if (v.big) {                      /* updatePartialAggBigInt */
    s->isum += v.i;
    s->count++;
    if (!s->has_big || v.i < s->bmin) s->bmin = v.i;
    if (!s->has_big || v.i > s->bmax) s->bmax = v.i;
    s->has_big = 1;
    return;
  }
  /* updatePartialAgg */
  if (is_int_col) s->isum += (int64_t)v.f;
  if (isnan(v.f)) s->nan = 1;
  else if (isinf(v.f)) { if (v.f > 0) s->pinf = 1; else s->ninf = 1; }
  else {
    long double x = v.f, t = s->s + x;
    if (fabsl(s->s) >= fabsl(x)) s->c += (s->s - t) + x; else s->c += (x - t) + s->s;
    s->s = t;
  }
  s->count++;
  if (v.f < s->fmin) s->fmin = v.f;
  if (v.f > s->fmax) s->fmax = v.f;
  (void)a;
}

/* Group-by over n (left, right) row pairs (rrows may be NULL). Outputs in
 * first-seen group order; returns the group count (or -1 if cap exceeded).
 *   out_keys[g*nkeys + k]  int value or f64 bits;  out_knull same layout
 *   out_i / out_f [g*naggs + a]; out_null [g*naggs + a]; out_is_int[a] */
int64_t oc_groupby(const oc_col *cols, const int *side, const uint64_t *lrows, const uint64_t *rrows, uint64_t n,
                   const int *keycols, int nkeys, const oc_agg *aggs, int naggs, uint64_t cap,
                   int64_t *out_keys, uint8_t *out_knull, int64_t *out_i, double *out_f, uint8_t *out_null,
                   uint8_t *out_is_int) {
  uint64_t tcap = 16;
  while (tcap < 2 * (cap + 1)) tcap <<= 1;
  slot_t *tab = calloc(tcap, sizeof(slot_t));
  oc_state *st = calloc((cap + 1) * (naggs ? naggs : 1), sizeof(oc_state));
  int64_t *keys = malloc(sizeof(int64_t) * (cap + 1) * (nkeys ? nkeys : 1));
  uint8_t *knull = malloc((cap + 1) * (nkeys ? nkeys : 1));
  uint64_t ng = 0;
  if (!tab || !st || !keys || !knull) { free(tab); free(st); free(keys); free(knull); return -2; }
  if (nkeys == 0) {
    ng = 1;  /* aggregate without GROUP BY: one row even with no input */
    for (int a = 0; a < naggs; ++a) { st[a].fmin = INFINITY; st[a].fmax = -INFINITY; }
  }
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t lr = lrows ? lrows[i] : i, rr = rrows ? rrows[i] : 0;
    uint64_t g = 0;
    if (nkeys > 0) {
      int64_t kv[4]; uint8_t kn[4];
      uint64_t h = 0x12345;
      for (int k = 0; k < nkeys; ++k) {
        int isn;
        uint64_t r = side[keycols[k]] ? rr : lr;
        kv[k] = (int64_t)key_part(&cols[keycols[k]], r, &isn);
        kn[k] = (uint8_t)isn;
        h = mixh(h ^ ((uint64_t)kv[k] + 0x9E37ull * isn));
      }
      uint64_t pos = h & (tcap - 1);
      for (;;) {
        slot_t *s = &tab[pos];
        if (!s->used) {
          if (ng >= cap) { free(tab); free(st); free(keys); free(knull); return -1; }
          s->used = 1; s->hash = h; s->group = ng;
          for (int k = 0; k < nkeys; ++k) { keys[ng * nkeys + k] = kv[k]; knull[ng * nkeys + k] = kn[k]; }
          for (int a = 0; a < naggs; ++a) { st[ng * naggs + a].fmin = INFINITY; st[ng * naggs + a].fmax = -INFINITY; }
          g = ng++;
          break;
        }
        if (s->hash == h) {
          int same = 1;
          for (int k = 0; k < nkeys && same; ++k)
            same = keys[s->group * nkeys + k] == kv[k] && knull[s->group * nkeys + k] == kn[k];
          if (same) { g = s->group; break; }
        }
        pos = (pos + 1) & (tcap - 1);
      }
    }
    for (int a = 0; a < naggs; ++a) {
      oc_state *s = &st[g * naggs + a];
      if (aggs[a].col_a < 0) { s->count++; continue; }  /* '*' */
      int isn;
      jsv v = agg_value(cols, side, &aggs[a], lr, rr, &isn);
      if (isn) continue;
      int is_int_col = aggs[a].col_b < 0 && cols[aggs[a].col_a].dtype != DT_F64 &&
                       cols[aggs[a].col_a].dtype != DT_F32 && cols[aggs[a].col_a].dtype != DT_I64;
      feed(s, &aggs[a], v, is_int_col);
    }
  }
  /* resolveValue (src/partial-agg.ts:105-134) */
  for (int a = 0; a < naggs; ++a) {
    int ci = aggs[a].col_a;
    int dt = ci >= 0 ? cols[ci].dtype : -1;
    int prod = aggs[a].col_b >= 0;
    int intcol = ci >= 0 && !prod && dt != DT_F64 && dt != DT_F32;
    int f = aggs[a].fn;
    out_is_int[a] = (f == AGG_COUNT) || (intcol && (f == AGG_SUM || f == AGG_MIN || f == AGG_MAX));
  }
  for (uint64_t g = 0; g < ng; ++g) {
    for (int k = 0; k < nkeys; ++k) {
      out_keys[g * nkeys + k] = keys[g * nkeys + k];
      out_knull[g * nkeys + k] = knull[g * nkeys + k];
    }
    for (int a = 0; a < naggs; ++a) {
      oc_state *s = &st[g * naggs + a];
      uint64_t o = g * naggs + a;
      out_null[o] = 0; out_i[o] = 0; out_f[o] = 0;
      int f = aggs[a].fn;
      if (f == AGG_COUNT) { out_i[o] = (int64_t)s->count; continue; }
      if (s->count == 0) { out_null[o] = 1; continue; }
      if (s->has_big) {
        if (f == AGG_SUM) out_i[o] = (int64_t)s->isum;
        else if (f == AGG_MIN) out_i[o] = s->bmin;
        else if (f == AGG_MAX) out_i[o] = s->bmax;
        else { /* bigAvg: Number(S / n) + Number(S % n) / n */
          __int128 q = s->isum / (__int128)s->count, r = s->isum % (__int128)s->count;
          out_f[o] = (double)q + (double)r / (double)s->count;
        }
        continue;
      }
      double sum = (s->nan || (s->pinf && s->ninf)) ? NAN : s->pinf ? INFINITY : s->ninf ? -INFINITY
                 : (double)(s->s + s->c);
      if (out_is_int[a]) {
        if (f == AGG_SUM) out_i[o] = (int64_t)s->isum;
        else if (f == AGG_MIN) out_i[o] = (int64_t)s->fmin;
        else out_i[o] = (int64_t)s->fmax;
      } else if (f == AGG_SUM) out_f[o] = sum;
      else if (f == AGG_AVG) {
        int ci = aggs[a].col_a;
        int intcol = aggs[a].col_b < 0 && cols[ci].dtype != DT_F64 && cols[ci].dtype != DT_F32;
        out_f[o] = intcol ? (double)s->isum / (double)s->count : sum / (double)s->count;
      } else if (f == AGG_MIN) out_f[o] = s->fmin;
      else out_f[o] = s->fmax;
    }
  }
  free(tab); free(st); free(keys); free(knull);
  return (int64_t)ng;
}

/* ---------------- hash join (build = right) ---------------- */
/* Pairs (left row, right row) of an inner join in probe order, then build
 * insertion order. lrows/rrows: the candidate rows of each side (after their
 * filters). Returns the pair count, or -1 if cap is exceeded. */
int64_t oc_join(const oc_col *lkey, const uint64_t *lrows, uint64_t nl, const oc_col *rkey, const uint64_t *rrows,
                uint64_t nr, uint64_t cap, uint64_t *out_l, uint64_t *out_r) {
  uint64_t tcap = 16;
  while (tcap < 2 * (nr + 1)) tcap <<= 1;
  int64_t *hk = malloc(sizeof(int64_t) * tcap);
  int64_t *head = malloc(sizeof(int64_t) * tcap), *tail = malloc(sizeof(int64_t) * tcap);
  int64_t *next = malloc(sizeof(int64_t) * (nr + 1));
  uint8_t *used = calloc(tcap, 1);
  if (!hk || !head || !tail || !next || !used) { free(hk); free(head); free(tail); free(next); free(used); return -2; }
  for (uint64_t j = 0; j < nr; ++j) {
    uint64_t r = rrows ? rrows[j] : j;
    int isn;
    uint64_t k = key_part(rkey, r, &isn);
    next[j] = -1;
    if (isn) continue;  /* NULL never matches (src/operators.ts:2424-2428) */
    uint64_t pos = mixh(k) & (tcap - 1);
    while (used[pos] && (uint64_t)hk[pos] != k) pos = (pos + 1) & (tcap - 1);
    if (!used[pos]) { used[pos] = 1; hk[pos] = (int64_t)k; head[pos] = (int64_t)j; tail[pos] = (int64_t)j; }
    else { next[tail[pos]] = (int64_t)j; tail[pos] = (int64_t)j; }
  }
  uint64_t m = 0;
  for (uint64_t i = 0; i < nl; ++i) {
    uint64_t l = lrows ? lrows[i] : i;
    int isn;
    uint64_t k = key_part(lkey, l, &isn);
    if (isn) continue;
    uint64_t pos = mixh(k) & (tcap - 1);
    while (used[pos] && (uint64_t)hk[pos] != k) pos = (pos + 1) & (tcap - 1);
    if (!used[pos]) continue;
    for (int64_t j = head[pos]; j >= 0; j = next[j]) {
      if (m >= cap) { free(hk); free(head); free(tail); free(next); free(used); return -1; }
      out_l[m] = l; out_r[m] = rrows ? rrows[j] : (uint64_t)j; ++m;
    }
  }
  free(hk); free(head); free(tail); free(next); free(used);
  return (int64_t)m;
}

/* ---------------- top-K ---------------- */
typedef struct { int cls; long double v; uint64_t row; } tk_t;
static int g_desc;
static int tk_cmp(const void *x, const void *y) {
  const tk_t *a = x, *b = y;
  if (a->cls != b->cls) return a->cls - b->cls;
  if (!a->cls && a->v != b->v) return ((a->v < b->v) ^ g_desc) ? -1 : 1;
  return a->row < b->row ? -1 : (a->row > b->row ? 1 : 0);
}

static tk_t tk_of(const oc_col *c, uint64_t r) {
  tk_t t; t.row = r; t.cls = 0; t.v = 0;
  if (is_null_at(c, r)) { t.cls = 1; return t; }
  jsv x = cell(c, r);
  if (x.big) t.v = (long double)x.i;
  else if (isnan(x.f)) t.cls = 1;
  else t.v = x.f;
  return t;
}

/* Rows of ORDER BY col [desc] under the total order (value, row id), NULL and
 * NaN last; writes min(n, k) rows. */
uint64_t oc_topk_exact(const oc_col *c, const uint64_t *rows, uint64_t n, int desc, uint64_t k, uint64_t *out) {
  tk_t *a = malloc(sizeof(tk_t) * (n ? n : 1));
  for (uint64_t i = 0; i < n; ++i) a[i] = tk_of(c, rows ? rows[i] : i);
  g_desc = desc;
  qsort(a, n, sizeof(tk_t), tk_cmp);
  uint64_t m = n < k ? n : k;
  for (uint64_t i = 0; i < m; ++i) out[i] = a[i].row;
  free(a);
  return m;
}

/* The reference TopKOperator: binary heap of k rows, root = worst kept,
 * replaced only by a strictly better value (src/operators.ts:1731-1802);
 * returns the kept rows in heap-pop order reversed (best first). */
static int heap_cmp(const tk_t *a, const tk_t *b, int desc) {  /* > 0: a sorts after b */
  if (a->cls && b->cls) return 0;
  if (a->cls) return 1;
  if (b->cls) return -1;
  int c = a->v < b->v ? -1 : (a->v > b->v ? 1 : 0);
  return desc ? -c : c;
}
static void sift_down(tk_t *h, uint64_t n, uint64_t i, int desc) {
  for (;;) {
    uint64_t t = i, l = 2 * i + 1, r = 2 * i + 2;
    if (l < n && heap_cmp(&h[l], &h[t], desc) > 0) t = l;
    if (r < n && heap_cmp(&h[r], &h[t], desc) > 0) t = r;
    if (t == i) break;
    tk_t x = h[i]; h[i] = h[t]; h[t] = x; i = t;
  }
}
uint64_t oc_topk_heap(const oc_col *c, const uint64_t *rows, uint64_t n, int desc, uint64_t k, uint64_t *out) {
  tk_t *h = malloc(sizeof(tk_t) * (k ? k : 1));
  uint64_t hn = 0;
  for (uint64_t i = 0; i < n && k; ++i) {
    tk_t x = tk_of(c, rows ? rows[i] : i);
    if (hn < k) {
      h[hn] = x;
      uint64_t j = hn++;
      while (j > 0) {
        uint64_t p = (j - 1) >> 1;
        if (heap_cmp(&h[j], &h[p], desc) > 0) { tk_t t = h[j]; h[j] = h[p]; h[p] = t; j = p; }
        else break;
      }
    } else {
      /* shouldReplace: strictly better than the root */
      const tk_t *rt = &h[0];
      int better;
      if (x.cls) better = 0;
      else if (rt->cls) better = 1;
      else better = desc ? x.v > rt->v : x.v < rt->v;
      if (better) { h[0] = x; sift_down(h, hn, 0, desc); }
    }
  }
  uint64_t m = hn;
  for (uint64_t i = 0; i < m; ++i) {  /* pop worst first, then reverse */
    out[m - 1 - i] = h[0].row;
    h[0] = h[hn - 1];
    --hn;
    if (hn) sift_down(h, hn, 0, desc);
  }
  free(h);
  return m;
}

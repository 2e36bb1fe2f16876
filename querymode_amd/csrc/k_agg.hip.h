// k_agg.hip.h — group-by building blocks shared by the fused scan and the
// partitioned aggregation: key codes, the global open-addressing table,
// per-state contributions and their LDS / global atomic application.
#pragma once
#include "k_common.hip.h"

namespace qm {

static constexpr int LDS_HASH_PROBES = 8;
static constexpr uint32_t GLB_PROBE_LIMIT = 4096;

// Stored (salted) 64-bit key of a row for GM_HASH. Returns false when the
// row's key must go to a reserved side entry (*side = 0: unrepresentable
// key, 1: NULL key).
__device__ __forceinline__ uint64_t key_code_hash(const DKey &k, const int64_t kv[MAX_KEYS],
                                                  const uint32_t knull[MAX_KEYS], int *side) {
  *side = -1;
  if (k.ncols == 1) {
    if (knull[0]) { *side = 1; return 0; }
    uint64_t raw = (uint64_t)kv[0];
    uint64_t s = raw + KEY_SALT;
    if (s == 0) { *side = 0; return 0; }
    return s;
  }
  // packed codes: code_i = v - kmin (null -> dom - 1)
  uint64_t packed = 0;
  int shift = 0;
#pragma unroll
  for (int i = 0; i < MAX_KEYS; ++i) {  // fixed bound: the key arrays stay in registers
    if (i >= k.ncols) break;
    uint64_t code = knull[i] ? (k.dom[i] - 1) : (uint64_t)(kv[i] - k.kmin[i]);
    packed |= code << shift;
    shift += k.bits[i];
  }
  return packed + KEY_SALT;  // packed < 2^63 so never 0
}

__device__ __forceinline__ uint64_t key_index_direct(const DKey &k, const int64_t kv[MAX_KEYS],
                                                     const uint32_t knull[MAX_KEYS]) {
  uint64_t idx = 0;
#pragma unroll
  for (int i = 0; i < MAX_KEYS; ++i) {  // fixed bound: the key arrays stay in registers
    if (i >= k.ncols) break;
    uint64_t code = knull[i] ? (k.dom[i] - 1) : (uint64_t)(kv[i] - k.kmin[i]);
    idx += code * k.stride[i];
  }
  return idx;
}

// Global hash-table lookup/insert. Returns the entry index or UINT64_MAX on
// overflow (probe limit or fill limit exceeded; the host regrows and reruns).
__device__ __forceinline__ uint64_t glb_find(const DQuery &q, uint64_t stored) {
  const uint64_t mask = q.gcap - 1;
  uint64_t h = mix64(stored) & mask;
  const int W = q.entry_words;
  for (uint32_t p = 0; p < GLB_PROBE_LIMIT; ++p) {
    uint64_t slot = (h + p) & mask;
    uint64_t *e = q.table + slot * W;
    uint64_t k = g_load(e);
    if (k == stored) return slot;
    if (k == 0) {
      uint64_t old = g_atomic_cas(e, 0, stored);
      if (old == 0) {
        // new group: bound the fill so probe sequences stay short
        uint64_t n = g_atomic_add(q.selected + 1, 1);
        if (n + 1 > (q.gcap >> 1) + (q.gcap >> 2)) *gptr(q.status) = 1;
        return slot;
      }
      if (old == stored) return slot;
    }
  }
  *gptr(q.status) = 1;
  return ~0ull;
}

// Variant for callers that count their own inserts: *inserted = this call
// claimed the slot. No fill accounting here.
__device__ __forceinline__ uint64_t glb_find_ins(const DQuery &q, uint64_t stored, bool *inserted) {
  const uint64_t mask = q.gcap - 1;
  uint64_t h = mix64(stored) & mask;
  const int W = q.entry_words;
  *inserted = false;
  for (uint32_t p = 0; p < GLB_PROBE_LIMIT; ++p) {
    uint64_t slot = (h + p) & mask;
    uint64_t *e = q.table + slot * W;
    uint64_t k = g_load(e);
    if (k == stored) return slot;
    if (k == 0) {
      uint64_t old = g_atomic_cas(e, 0, stored);
      if (old == 0) { *inserted = true; return slot; }
      if (old == stored) return slot;
    }
  }
  *gptr(q.status) = 1;
  return ~0ull;
}

// ---- state updates ----
// value fetch for a state: integer or f64, with nulls
struct RowVals {
  int64_t iv[RPT];
  double fv[RPT];
  uint32_t nm;
};

__device__ __forceinline__ void fetch_state_vals(const DState &s, const DCol *cols, uint64_t row0,
                                                 RowVals &rv) {
  const DCol &a = cols[s.col_a];
  rv.nm = load8_nulls(a, row0);
  if (s.product) {
    const DCol &b = cols[s.col_b];
    double x[RPT], y[RPT];
    if (dt_is_float(a.dtype)) load8_f(a, row0, x);
    else { int64_t t[RPT]; load8_i(a, row0, t); for (int r = 0; r < RPT; ++r) x[r] = (double)t[r]; }
    if (dt_is_float(b.dtype)) load8_f(b, row0, y);
    else { int64_t t[RPT]; load8_i(b, row0, t); for (int r = 0; r < RPT; ++r) y[r] = (double)t[r]; }
#pragma unroll
    for (int r = 0; r < RPT; ++r) rv.fv[r] = x[r] * y[r];
    rv.nm |= load8_nulls(b, row0);
  } else if (dt_is_float(a.dtype)) {
    load8_f(a, row0, rv.fv);
  } else {
    load8_i(a, row0, rv.iv);
  }
}

// Per-value contribution of a state, as the words to atomically combine.
struct Contrib {
  uint64_t a, b, c;  // meaning depends on kind (SK_SUM_WIDE: the three signed digits)
  uint64_t f;        // SK_SUM_WIDE: non-finite flags
  int32_t wi;        // SK_SUM_WIDE: first digit word (-1: zero)
  bool skip;
};

__device__ __forceinline__ Contrib contrib_of(const DState &s, bool is_float, int64_t iv, double fv) {
  Contrib k; k.a = k.b = k.c = 0; k.f = 0; k.wi = -1; k.skip = false;
  switch (s.kind) {
    case SK_COUNT: k.a = 1; break;
    case SK_SUM_I64: k.a = (uint64_t)iv; break;
    case SK_SUM_128:
      if (s.fixed) {
        uint64_t bits = f64_bits(fv);
        if (((bits >> 52) & 0x7FF) == 0x7FF) {
          k.c = f64_is_nan_bits(bits) ? SUMF_NAN : ((bits >> 63) ? SUMF_NINF : SUMF_PINF);
        } else {
          u128 q = f64_to_fixed(fv, s.scale);
          k.a = q.lo; k.b = q.hi;
        }
      } else {
        k.a = (uint64_t)iv; k.b = (iv < 0) ? ~0ull : 0ull;
      }
      break;
    case SK_SUM_WIDE: {
      const uint64_t bits = f64_bits(fv);
      if (((bits >> 52) & 0x7FF) == 0x7FF) {
        k.f = f64_is_nan_bits(bits) ? SUMF_NAN : ((bits >> 63) ? SUMF_NINF : SUMF_PINF);
      } else {
        int64_t d[3];
        f64_wide_digits(fv, s.scale, &k.wi, d);
        k.a = (uint64_t)d[0]; k.b = (uint64_t)d[1]; k.c = (uint64_t)d[2];
      }
      break;
    }
    default: {  // SK_MAX_CODE
      uint64_t code;
      if (is_float) {
        if (fv != fv) { k.skip = true; break; }  // NaN never updates MIN/MAX
        code = code_f64_bits(f64_bits(fv));
      } else {
        code = code_i64(iv);
      }
      k.a = s.is_min ? ~code : code;
      break;
    }
  }
  return k;
}

__device__ __forceinline__ void apply_glb(const DState &s, uint64_t *e, const Contrib &k) {
  uint64_t *w = e + s.word;
  switch (s.kind) {
    case SK_COUNT: case SK_SUM_I64: g_atomic_add(w, k.a); break;
    case SK_SUM_128:
      g_add128(w, k.a, k.b);
      if (s.fixed && k.c) g_atomic_or(w + 2, k.c);
      break;
    case SK_SUM_WIDE:
      if (k.f) g_atomic_or(w, k.f);
      if (k.wi >= 0) {
        if (k.a) g_atomic_add(w + 1 + k.wi, k.a);
        if (k.b) g_atomic_add(w + 2 + k.wi, k.b);
        if (k.c) g_atomic_add(w + 3 + k.wi, k.c);
      }
      break;
    default: g_atomic_max(w, k.a); break;
  }
}

__device__ __forceinline__ void apply_lds(const DState &s, lds_u64 *e, const Contrib &k) {
  lds_u64 *w = e + s.word;
  switch (s.kind) {
    case SK_COUNT: case SK_SUM_I64: l_atomic_add(w, k.a); break;
    case SK_SUM_128:
      l_add128(w, k.a, k.b);
      if (s.fixed && k.c) l_atomic_or(w + 2, k.c);
      break;
    case SK_SUM_WIDE:
      if (k.f) l_atomic_or(w, k.f);
      if (k.wi >= 0) {
        if (k.a) l_atomic_add(w + 1 + k.wi, k.a);
        if (k.b) l_atomic_add(w + 2 + k.wi, k.b);
        if (k.c) l_atomic_add(w + 3 + k.wi, k.c);
      }
      break;
    default: l_atomic_max(w, k.a); break;
  }
}

// merge one entry's states (non-atomic source) into a global entry
__device__ __forceinline__ void merge_entry_glb(const DQuery &q, uint64_t *dst, const uint64_t *src) {
  for (int i = 0; i < q.nstates; ++i) {
    const DState &s = q.states[i];
    const uint64_t *w = src + s.word;
    uint64_t *d = dst + s.word;
    switch (s.kind) {
      case SK_COUNT: case SK_SUM_I64: if (w[0]) g_atomic_add(d, w[0]); break;
      case SK_SUM_128:
        g_add128(d, w[0], w[1]);
        if (s.fixed && w[2]) g_atomic_or(d + 2, w[2]);
        break;
      case SK_SUM_WIDE:
        if (w[0]) g_atomic_or(d, w[0]);
        for (int j = 1; j <= s.nw; ++j) if (w[j]) g_atomic_add(d + j, w[j]);
        break;
      default: if (w[0]) g_atomic_max(d, w[0]); break;
    }
  }
}

}  // namespace qm

// k_scan.hip — fused scan: predicate filter -> (optional hash-join probe) ->
// hash group-by aggregate, for one or several queries over one resident
// table in a single persistent launch.
//
// Replaces, for the local executor's hot path:
//   ScanOperator + scanFilterIndices/applyAndFilters (src/operators.ts:282-547)
//   the WASM filters + intersect (wasm/src/wasm/aggregates.zig:605-1065)
//   AggregateOperator + computePartialAggColumnar + mergeStates
//     (src/operators.ts:1646-1710, src/partial-agg.ts:141-267)
//   HashJoinOperator's in-memory probe when it feeds an aggregate
//     (src/operators.ts:2673-2713)
//
// Layout: each thread owns 8 consecutive rows of a tile (blockDim*8 rows);
// column loads are 64/128-bit vector loads; predicates AND into an 8-bit
// mask in registers (no index lists, no intersections). Selected rows update
// a per-workgroup LDS table (dense replicated array for small key domains,
// or an open-addressing cache for large ones) that is merged into the global
// table with atomics when the workgroup finishes; rows whose key misses the
// LDS cache go straight to the global table.
#include <map>
#include <mutex>
#include "k_agg.hip.h"
#include "qm_kernels.h"

namespace qm {

// ---- one query over the thread's 8 rows ----
__device__ __forceinline__ void agg_rows(const DScan &sc, const DQuery &q, const DCol *cols, uint64_t row0,
                         uint32_t mask, lds_u64 *lds) {
  const int W = q.entry_words;
  // Resolve each selected row to an entry: bit 63 set = global index.
  uint64_t ent[RPT];
  if (q.key.ncols == 0) {
    for (int r = 0; r < RPT; ++r) ent[r] = 0;  // LDS replica slot 0 / global entry 0
    if (q.lmode == LM_NONE) for (int r = 0; r < RPT; ++r) ent[r] = (1ull << 63);
  } else {
    int64_t kv[MAX_KEYS][RPT];
    uint32_t kn[MAX_KEYS] = {0, 0};
#pragma unroll
    for (int i = 0; i < MAX_KEYS; ++i) {
      if (i >= q.key.ncols) break;
      const DCol &c = cols[q.key.col[i]];
      kn[i] = load8_nulls(c, row0);
      if (q.key.is_float[i]) {
        double f[RPT];
        load8_f(c, row0, f);
        for (int r = 0; r < RPT; ++r) kv[i][r] = (int64_t)canon_f64_key_bits(f64_bits(f[r]));
      } else {
        load8_i(c, row0, kv[i]);
      }
    }
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      ent[r] = ~0ull;
      if (!((mask >> r) & 1)) continue;
      int64_t v[MAX_KEYS] = {0, 0};
      uint32_t n[MAX_KEYS] = {0, 0};
#pragma unroll
      for (int i = 0; i < MAX_KEYS; ++i) {
        if (i < q.key.ncols) { v[i] = kv[i][r]; n[i] = (kn[i] >> r) & 1; }
      }
      if (q.gmode == GM_DIRECT) {
        uint64_t idx = key_index_direct(q.key, v, n);
        ent[r] = (q.lmode == LM_DIRECT) ? idx : (idx | (1ull << 63));
      } else {
        int side;
        uint64_t stored = key_code_hash(q.key, v, n, &side);
        if (side >= 0) { ent[r] = (q.gcap + (uint64_t)side) | (1ull << 63); continue; }
        if (q.lmode == LM_HASH) {
          const uint32_t lmask = (uint32_t)q.lds_entries - 1;
          uint32_t h = (uint32_t)mix64(stored) & lmask;
          for (int p = 0; p < LDS_HASH_PROBES; ++p) {
            uint32_t slot = (h + p) & lmask;
            lds_u64 *e = lds + (uint64_t)slot * W;
            uint64_t k = *e;
            if (k == 0) k = l_atomic_cas(e, 0, stored);
            if (k == 0 || k == stored) { ent[r] = slot; break; }
          }
          if (ent[r] != ~0ull) continue;
        }
        uint64_t g = glb_find(q, stored);
        ent[r] = (g == ~0ull) ? ~0ull : (g | (1ull << 63));
      }
    }
  }
  // LDS replica for the direct LDS table
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t rep_base = (q.lmode == LM_DIRECT)
                                ? (uint64_t)(lane & (uint32_t)(q.lds_replicas - 1)) * q.lds_stride
                                : 0;
  for (int si = 0; si < q.nstates; ++si) {
    const DState &s = q.states[si];
    RowVals rv;
    bool is_float = false;
    if (s.col_a >= 0) {
      fetch_state_vals(s, cols, row0, rv);
      is_float = s.product || dt_is_float(cols[s.col_a].dtype);
    } else {
      rv.nm = 0;
    }
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      if (!((mask >> r) & 1)) continue;
      if ((rv.nm >> r) & 1) continue;  // null values do not feed aggregates
      uint64_t e = ent[r];
      if (e == ~0ull) continue;        // dropped on overflow (rerun follows)
      Contrib k = contrib_of(s, is_float, (s.col_a >= 0 && !is_float) ? rv.iv[r] : 0,
                             is_float ? rv.fv[r] : 0.0);
      if (k.skip) continue;
      if (e >> 63) {
        apply_glb(s, q.table + (e & ~(1ull << 63)) * W, k);
      } else {
        apply_lds(s, lds + rep_base + e * W, k);
      }
    }
  }
}

__device__ __forceinline__ uint64_t lds_words_total(const DQuery *qs, int nq) {
  uint64_t t = 0;
  for (int i = 0; i < nq; ++i) {
    uint64_t e = (uint64_t)qs[i].lds_offset + qs[i].lds_words;
    if (e > t) t = e;
  }
  return t;
}

// Merge a workgroup's LDS table into the global table.
__device__ __forceinline__ void flush_lds(const DQuery &q, lds_u64 *lds) {
  const int W = q.entry_words;
  if (q.lmode == LM_DIRECT) {
    const uint64_t nslots = (uint64_t)q.lds_entries;
    for (uint64_t slot = threadIdx.x; slot < nslots; slot += blockDim.x) {
      uint64_t acc[24];
      for (int w = 0; w < W; ++w) acc[w] = 0;
      for (int rep = 0; rep < q.lds_replicas; ++rep) {
        lds_u64 *e = lds + (uint64_t)rep * q.lds_stride + slot * W;
        for (int i = 0; i < q.nstates; ++i) {
          const DState &s = q.states[i];
          int w = s.word;
          switch (s.kind) {
            case SK_COUNT: case SK_SUM_I64: acc[w] += e[w]; break;
            case SK_SUM_128: {
              uint64_t lo = acc[w] + e[w];
              acc[w + 1] += e[w + 1] + (lo < acc[w] ? 1 : 0);
              acc[w] = lo;
              if (s.fixed) acc[w + 2] |= e[w + 2];
              break;
            }
            case SK_SUM_WIDE:  // flags word, then independent digit words
              acc[w] |= e[w];
              for (int j = 1; j <= s.nw; ++j) acc[w + j] += e[w + j];
              break;
            default: if (e[w] > acc[w]) acc[w] = e[w]; break;
          }
        }
      }
      if (acc[q.count_word] == 0) continue;  // no row reached this slot
      merge_entry_glb(q, q.table + slot * W, acc);
    }
  } else if (q.lmode == LM_HASH) {
    for (uint64_t slot = threadIdx.x; slot < (uint64_t)q.lds_entries; slot += blockDim.x) {
      lds_u64 *e = lds + slot * W;
      uint64_t stored = e[0];
      if (stored == 0) continue;
      uint64_t acc[24];
      for (int w = 0; w < W; ++w) acc[w] = e[w];
      if (acc[q.count_word] == 0) continue;
      uint64_t g = glb_find(q, stored);
      if (g == ~0ull) continue;
      merge_entry_glb(q, q.table + g * W, acc);
    }
  }
}

// One joined row (probe row prow, build row brow) into the query's group
// table: key and state values through single-row gathers (build columns are
// the side-1 columns).
__device__ __forceinline__ void join_apply(const DScan &sc, const DQuery &q, uint64_t prow, uint32_t brow, lds_u64 *lds) {
  // One joined row: key and state values through single-row gathers.
  const int W = q.entry_words;
  uint64_t ent = ~0ull;
  {
    int64_t v[MAX_KEYS]; uint32_t n[MAX_KEYS];
    for (int i = 0; i < q.key.ncols; ++i) {
      const DCol &c = sc.cols[q.key.col[i]];
      uint64_t row = c.side ? brow : (prow);
      n[i] = load1_null(c, row);
      v[i] = q.key.is_float[i] ? (int64_t)canon_f64_key_bits(f64_bits(load1_f(c, row)))
                               : load1_i(c, row);
    }
    if (q.key.ncols == 0) {
      ent = (q.lmode == LM_NONE) ? (1ull << 63) : 0;
    } else if (q.gmode == GM_DIRECT) {
      uint64_t idx = key_index_direct(q.key, v, n);
      ent = (q.lmode == LM_DIRECT) ? idx : (idx | (1ull << 63));
    } else {
      int side;
      uint64_t stored = key_code_hash(q.key, v, n, &side);
      if (side >= 0) ent = (q.gcap + (uint64_t)side) | (1ull << 63);
      else {
        uint64_t g = glb_find(q, stored);
        ent = (g == ~0ull) ? ~0ull : (g | (1ull << 63));
      }
    }
  }
  if (ent == ~0ull) return;
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t rep_base = (q.lmode == LM_DIRECT)
      ? (uint64_t)(lane & (uint32_t)(q.lds_replicas - 1)) * q.lds_stride : 0;
  for (int si = 0; si < q.nstates; ++si) {
    const DState &s = q.states[si];
    int64_t iv = 0; double fv = 0.0; bool isf = false; bool isnull = false;
    if (s.col_a >= 0) {
      const DCol &a = sc.cols[s.col_a];
      uint64_t ra = a.side ? brow : (prow);
      isnull = load1_null(a, ra);
      if (s.product) {
        const DCol &b = sc.cols[s.col_b];
        uint64_t rb = b.side ? brow : (prow);
        isnull = isnull || load1_null(b, rb);
        double x = dt_is_float(a.dtype) ? load1_f(a, ra) : (double)load1_i(a, ra);
        double y = dt_is_float(b.dtype) ? load1_f(b, rb) : (double)load1_i(b, rb);
        fv = x * y; isf = true;
      } else if (dt_is_float(a.dtype)) {
        fv = load1_f(a, ra); isf = true;
      } else {
        iv = load1_i(a, ra);
      }
    }
    if (isnull) continue;  // this state skips a NULL value
    Contrib k = contrib_of(s, isf, iv, fv);
    if (k.skip) continue;
    if (ent >> 63) apply_glb(s, q.table + (ent & ~(1ull << 63)) * W, k);
    else apply_lds(s, lds + q.lds_offset + rep_base + ent * W, k);
  }

}

// The joined rows of one lane's 8 probe rows (unique build keys: b1[r] =
// build row + 1, 0 = no match), with the loads of all 8 issued together:
// first every key gather, then per state every value load, then the 8
// updates. The values and updates equal join_apply's row by row (the updates
// are commutative: exact sums, counts, MIN / MAX codes).
__device__ __forceinline__ void join_apply8(const DScan &sc, const DQuery &q, uint64_t row0, const uint32_t b1[RPT],
                                            lds_u64 *lds) {
  const int W = q.entry_words;
  uint64_t ent[RPT];
  {
    int64_t v[MAX_KEYS][RPT];
    uint32_t n[MAX_KEYS][RPT];
#pragma unroll
    for (int i = 0; i < MAX_KEYS; ++i) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) { v[i][r] = 0; n[i][r] = 0; }
      if (i >= q.key.ncols) continue;
      const DCol &c = sc.cols[q.key.col[i]];
      const bool isf = q.key.is_float[i] != 0;
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        if (!b1[r]) continue;
        const uint64_t row = c.side ? (uint64_t)(b1[r] - 1u) : row0 + r;
        n[i][r] = load1_null(c, row);
        v[i][r] = isf ? (int64_t)canon_f64_key_bits(f64_bits(load1_f(c, row))) : load1_i(c, row);
      }
    }
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      ent[r] = ~0ull;
      if (!b1[r]) continue;
      int64_t kv[MAX_KEYS];
      uint32_t kn[MAX_KEYS];
#pragma unroll
      for (int i = 0; i < MAX_KEYS; ++i) { kv[i] = v[i][r]; kn[i] = n[i][r]; }
      if (q.key.ncols == 0) {
        ent[r] = (q.lmode == LM_NONE) ? (1ull << 63) : 0;
      } else if (q.gmode == GM_DIRECT) {
        const uint64_t idx = key_index_direct(q.key, kv, kn);
        ent[r] = (q.lmode == LM_DIRECT) ? idx : (idx | (1ull << 63));
      } else {
        int side;
        const uint64_t stored = key_code_hash(q.key, kv, kn, &side);
        if (side >= 0) {
          ent[r] = (q.gcap + (uint64_t)side) | (1ull << 63);
        } else {
          const uint64_t g = glb_find(q, stored);
          ent[r] = (g == ~0ull) ? ~0ull : (g | (1ull << 63));
        }
      }
    }
  }
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t rep_base = (q.lmode == LM_DIRECT)
      ? (uint64_t)(lane & (uint32_t)(q.lds_replicas - 1)) * q.lds_stride : 0;
  for (int si = 0; si < q.nstates; ++si) {
    const DState &s = q.states[si];
    int64_t iv[RPT];
    double fv[RPT];
    bool isnull[RPT];
    bool isf = false;
#pragma unroll
    for (int r = 0; r < RPT; ++r) { iv[r] = 0; fv[r] = 0.0; isnull[r] = false; }
    if (s.col_a >= 0) {
      const DCol &a = sc.cols[s.col_a];
      if (s.product) {
        const DCol &b = sc.cols[s.col_b];
        isf = true;
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          if (ent[r] == ~0ull) continue;
          const uint64_t ra = a.side ? (uint64_t)(b1[r] - 1u) : row0 + r;
          const uint64_t rb = b.side ? (uint64_t)(b1[r] - 1u) : row0 + r;
          isnull[r] = load1_null(a, ra) || load1_null(b, rb);
          const double x = dt_is_float(a.dtype) ? load1_f(a, ra) : (double)load1_i(a, ra);
          const double y = dt_is_float(b.dtype) ? load1_f(b, rb) : (double)load1_i(b, rb);
          fv[r] = x * y;
        }
      } else {
        isf = dt_is_float(a.dtype);
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          if (ent[r] == ~0ull) continue;
          const uint64_t ra = a.side ? (uint64_t)(b1[r] - 1u) : row0 + r;
          isnull[r] = load1_null(a, ra);
          if (isf) fv[r] = load1_f(a, ra);
          else iv[r] = load1_i(a, ra);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      if (ent[r] == ~0ull || isnull[r]) continue;  // no match / this state skips a NULL value
      const Contrib k = contrib_of(s, isf, iv[r], fv[r]);
      if (k.skip) continue;
      if (ent[r] >> 63) apply_glb(s, q.table + (ent[r] & ~(1ull << 63)) * W, k);
      else apply_lds(s, lds + q.lds_offset + rep_base + ent[r] * W, k);
    }
  }
}

// ---- fused probe through the join payload (DScan::join_pay) ----
__device__ __forceinline__ uint64_t pay_bits(const gv4u &p, int off) {
  const uint64_t b0 = p.x | ((uint64_t)p.y << 32), b1 = p.z | ((uint64_t)p.w << 32);
  return off < 8 ? (b0 >> (8 * off)) : (b1 >> (8 * (off - 8)));
}
__device__ __forceinline__ int64_t pay_i(const gv4u &p, int off, int32_t dt) {
  const uint64_t v = pay_bits(p, off);
  switch (dt) {
    case DT_I8: return (int8_t)v;
    case DT_BOOL: return (uint8_t)v;
    case DT_I16: return (int16_t)v;
    case DT_I32: return (int32_t)v;
    default: return (int64_t)v;
  }
}
__device__ __forceinline__ double pay_f(const gv4u &p, int off, int32_t dt) {
  const uint64_t v = pay_bits(p, off);
  return dt == DT_F32 ? (double)__uint_as_float((uint32_t)v) : __longlong_as_double((long long)v);
}

// The 8 values (and NULL bits) of column ci for one lane's probe rows: a
// probe-side column by vector loads, a build-side one out of the payload.
__device__ __forceinline__ void join_vals8(const DScan &sc, int ci, uint64_t row0, const gv4u p[RPT], bool as_f,
                                           double f[RPT], int64_t iv[RPT], uint32_t *nm) {
  const DCol &c = sc.cols[ci];
  uint32_t n = 0;
  if (c.side) {
    const int off = sc.pay_off[ci], nb = sc.pay_null[ci];
    const bool isf = dt_is_float(c.dtype);
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      if (nb) n |= ((p[r].w >> nb) & 1u) << r;
      if (as_f) f[r] = isf ? pay_f(p[r], off, c.dtype) : (double)pay_i(p[r], off, c.dtype);
      else iv[r] = pay_i(p[r], off, c.dtype);
    }
  } else {
    n = load8_nulls(c, row0);
    if (as_f && dt_is_float(c.dtype)) {
      load8_f(c, row0, f);
    } else {
      int64_t t[RPT];
      load8_i(c, row0, t);
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        if (as_f) f[r] = (double)t[r];
        else iv[r] = t[r];
      }
    }
  }
  *nm = n;
}

// join_apply8 with the build rows' columns taken from their payload slots
// (p[r].w bit 0: probe row r has a build row). Same values and updates.
__device__ __forceinline__ void join_apply8p(const DScan &sc, const DQuery &q, uint64_t row0, const gv4u p[RPT],
                                             lds_u64 *lds) {
  const int W = q.entry_words;
  uint64_t ent[RPT];
  uint32_t has = 0;
#pragma unroll
  for (int r = 0; r < RPT; ++r) has |= (p[r].w & 1u) << r;
  if (!has) return;
  {
    int64_t v[MAX_KEYS][RPT];
    uint32_t n[MAX_KEYS];
#pragma unroll
    for (int i = 0; i < MAX_KEYS; ++i) {
      n[i] = 0;
#pragma unroll
      for (int r = 0; r < RPT; ++r) v[i][r] = 0;
      if (i >= q.key.ncols) continue;
      const int ci = q.key.col[i];
      const bool isf = q.key.is_float[i] != 0;
      double f[RPT];
      join_vals8(sc, ci, row0, p, isf, f, v[i], &n[i]);
      if (isf) {
#pragma unroll
        for (int r = 0; r < RPT; ++r) v[i][r] = (int64_t)canon_f64_key_bits(f64_bits(f[r]));
      }
    }
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      ent[r] = ~0ull;
      if (!((has >> r) & 1)) continue;
      int64_t kv[MAX_KEYS];
      uint32_t kn[MAX_KEYS];
#pragma unroll
      for (int i = 0; i < MAX_KEYS; ++i) { kv[i] = v[i][r]; kn[i] = (n[i] >> r) & 1; }
      if (q.key.ncols == 0) {
        ent[r] = (q.lmode == LM_NONE) ? (1ull << 63) : 0;
      } else if (q.gmode == GM_DIRECT) {
        const uint64_t idx = key_index_direct(q.key, kv, kn);
        ent[r] = (q.lmode == LM_DIRECT) ? idx : (idx | (1ull << 63));
      } else {
        int side;
        const uint64_t stored = key_code_hash(q.key, kv, kn, &side);
        if (side >= 0) {
          ent[r] = (q.gcap + (uint64_t)side) | (1ull << 63);
        } else {
          const uint64_t g = glb_find(q, stored);
          ent[r] = (g == ~0ull) ? ~0ull : (g | (1ull << 63));
        }
      }
    }
  }
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t rep_base = (q.lmode == LM_DIRECT)
      ? (uint64_t)(lane & (uint32_t)(q.lds_replicas - 1)) * q.lds_stride : 0;
  for (int si = 0; si < q.nstates; ++si) {
    const DState &s = q.states[si];
    int64_t iv[RPT];
    double fv[RPT];
    uint32_t nm = 0;
    bool isf = false;
#pragma unroll
    for (int r = 0; r < RPT; ++r) { iv[r] = 0; fv[r] = 0.0; }
    if (s.col_a >= 0) {
      if (s.product) {
        double x[RPT], y[RPT];
        int64_t unused[RPT];
        uint32_t na, nb;
        join_vals8(sc, s.col_a, row0, p, true, x, unused, &na);
        join_vals8(sc, s.col_b, row0, p, true, y, unused, &nb);
#pragma unroll
        for (int r = 0; r < RPT; ++r) fv[r] = x[r] * y[r];
        nm = na | nb;
        isf = true;
      } else {
        isf = dt_is_float(sc.cols[s.col_a].dtype);
        join_vals8(sc, s.col_a, row0, p, isf, fv, iv, &nm);
      }
    }
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      if (ent[r] == ~0ull || ((nm >> r) & 1)) continue;  // no match / this state skips a NULL value
      const Contrib k = contrib_of(s, isf, iv[r], fv[r]);
      if (k.skip) continue;
      if (ent[r] >> 63) apply_glb(s, q.table + (ent[r] & ~(1ull << 63)) * W, k);
      else apply_lds(s, lds + q.lds_offset + rep_base + ent[r] * W, k);
    }
  }
}

// JOIN: the fused join probe (one query); else up to MAX_QUERIES queries over
// one table. Two instantiations, so each keeps only its own registers.
template <bool JOIN>
__global__ void __launch_bounds__(512) scan_agg_kernel(DScan sc_arg, const DQuery *__restrict__ qs) {
  // The descriptor is read in place in the kernel-argument segment (it is the
  // first argument, at offset 0): its column array is indexed with
  // per-query slots, which on the by-value copy would move the whole struct
  // to per-lane scratch memory; in place they are scalar loads.
  (void)sc_arg;
  const DScan &sc = *(const DScan *)__builtin_amdgcn_kernarg_segment_ptr();
  extern __shared__ uint64_t lds_raw[];
  // LDS: [0, MAX_QUERIES) selected-row counters, then the query tables.
  lds_u64 *lds_cnt = (lds_u64 *)lds_raw;
  lds_u64 *lds = lds_cnt + MAX_QUERIES;
  const int nq = sc.nqueries;
  const uint64_t lw = MAX_QUERIES + lds_words_total(qs, nq);
  for (uint64_t i = threadIdx.x; i < lw; i += blockDim.x) lds_cnt[i] = 0;
  __syncthreads();

  const uint64_t tile_rows = (uint64_t)blockDim.x * RPT;
  const uint64_t ntiles = (sc.nrows + tile_rows - 1) / tile_rows;
  uint64_t nmatch = 0;  // joined rows (fused probe), added to the counter once

  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    if (page_skipped(sc, tile * tile_rows)) continue;  // uniform across the workgroup
    const uint64_t row0 = tile * tile_rows + (uint64_t)threadIdx.x * RPT;
    uint32_t valid = 0;
    if (row0 < sc.nrows) {
      uint64_t left = sc.nrows - row0;
      valid = left >= RPT ? 0xFFu : ((1u << left) - 1);
    }
    if constexpr (!JOIN) {
      for (int qi = 0; qi < nq; ++qi) {
        const DQuery &q = qs[qi];
        uint32_t m = eval_query8(q, sc.cols, row0, valid);
        if (m) l_atomic_add(lds_cnt + qi, (uint64_t)__builtin_popcount(m));
        if (m && q.gmode != GM_NONE) agg_rows(sc, q, sc.cols, row0, m, lds + q.lds_offset);
      }
    } else {
      // Fused probe: filter the probe rows, then per match aggregate with the
      // build row's columns visible as side-1 columns (single query).
      const DQuery &q = qs[0];
      uint32_t m = eval_query8(q, sc.cols, row0, valid);
      if (!m) continue;
      const DCol &kc = sc.cols[sc.join_col];
      int64_t kv[RPT];
      if (sc.join_key_float) {
        double f[RPT]; load8_f(kc, row0, f);
        for (int r = 0; r < RPT; ++r) kv[r] = (int64_t)canon_f64_key_bits(f64_bits(f[r]));
      } else {
        load8_i(kc, row0, kv);
      }
      m &= ~load8_nulls(kc, row0);  // NULL never matches in joins
      if (sc.join_pay) {
        // unique build keys, build columns packed per key slot: one 16-byte gather per probe row
        gv4u p[RPT];
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          p[r] = gv4u{0u, 0u, 0u, 0u};
          if ((m >> r) & 1) {
            const uint64_t idx = (uint64_t)(kv[r] - sc.join_min);
            if (idx < sc.join_cap) p[r] = gld<gv4u>(sc.join_pay, idx * 16);
          }
        }
#pragma unroll
        for (int r = 0; r < RPT; ++r) nmatch += p[r].w & 1u;
        join_apply8p(sc, q, row0, p, lds);
      } else if (sc.join_unique) {
        // unique build keys: the index holds the build row; all 8 lookups in flight at once
        uint32_t b1[RPT];
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          b1[r] = 0;
          if ((m >> r) & 1) {
            const uint64_t idx = (uint64_t)(kv[r] - sc.join_min);
            if (idx < sc.join_cap) b1[r] = sc.join_unique[idx];
          }
        }
#pragma unroll
        for (int r = 0; r < RPT; ++r) nmatch += b1[r] ? 1u : 0u;
        join_apply8(sc, q, row0, b1, lds);
      } else {
        for (int r = 0; r < RPT; ++r) {
          if (!((m >> r) & 1)) continue;
          uint32_t start, count;
          if (!join_lookup(sc, kv[r], &start, &count)) continue;
          for (uint32_t j = 0; j < count; ++j) {
            ++nmatch;
            join_apply(sc, q, row0 + r, sc.join_rows[start + j], lds);
          }
        }
      }
    }
  }
  if (nmatch) l_atomic_add(lds_cnt, nmatch);
  __syncthreads();
  // selected-row counters: one global atomic per workgroup per query
  if (threadIdx.x < (unsigned)nq && lds_cnt[threadIdx.x])
    g_atomic_add(qs[threadIdx.x].selected, lds_cnt[threadIdx.x]);
  for (int qi = 0; qi < nq; ++qi) {
    const DQuery &q = qs[qi];
    if (q.lmode != LM_NONE) flush_lds(q, lds + q.lds_offset);
  }
}

// Zero-fill (graph-safe replacement for a memset node).
__global__ void fill_u64_kernel(uint64_t *p, uint64_t n, uint64_t v) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// Byte-exact zero fill of [p, p + n) at any alignment (graph-safe replacement
// for a memset node): 8-byte stores over the aligned middle, byte stores for
// the < 8 bytes before and after it.
__global__ void zero_bytes_kernel(uint8_t *p, uint64_t n) {
  const uint64_t mis = (uint64_t)((8 - ((uintptr_t)p & 7)) & 7);
  const uint64_t head = mis < n ? mis : n;
  const uint64_t words = (n - head) / 8;
  const uint64_t tail0 = head + words * 8;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t *w = reinterpret_cast<uint64_t *>(p + head);
  for (uint64_t k = i; k < words; k += stride) w[k] = 0;
  if (i < head) p[i] = 0;
  if (i < n - tail0) p[tail0 + i] = 0;
}

}  // namespace qm

// ---- host launchers ----
extern "C" hipError_t qmk_scan_agg(const qm::DScan *sc, const qm::DQuery *d_queries, int grid,
                                   int block, size_t lds_bytes, hipStream_t st) {
  const void *fn = sc->join_mode ? (const void *)qm::scan_agg_kernel<true> : (const void *)qm::scan_agg_kernel<false>;
  if (lds_bytes > 65536) {
    hipError_t e = qmk_func_lds(fn, lds_bytes);
    if (e != hipSuccess) return e;
  }
  qm::DScan a = *sc;
  const qm::DQuery *q = d_queries;
  void *args[] = {&a, &q};
  return hipLaunchKernel(fn, dim3(grid), dim3(block), args, lds_bytes, st);
}

extern "C" hipError_t qmk_fill_u64(uint64_t *p, uint64_t n, uint64_t v, hipStream_t st) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(qm::fill_u64_kernel, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

extern "C" hipError_t qmk_zero_bytes(void *p, uint64_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n / 8 + 255) / 256;
  if (blocks < 1) blocks = 1;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(qm::zero_bytes_kernel, dim3((unsigned)blocks), dim3(256), 0, st, static_cast<uint8_t *>(p), n);
  return hipGetLastError();
}

extern "C" hipError_t qmk_func_lds(const void *fn, size_t bytes) {
  static std::mutex mu;
  static std::map<const void *, size_t> set;
  std::lock_guard<std::mutex> g(mu);
  auto it = set.find(fn);
  if (it != set.end() && it->second >= bytes) return hipSuccess;
  const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e == hipSuccess) set[fn] = bytes;
  return e;
}

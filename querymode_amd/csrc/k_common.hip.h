// k_common.hip.h — device helpers shared by the gfx950 kernels: vectorised
// 8-row column loads, null bits, predicate evaluation and LDS / global
// 64-bit atomics.
#pragma once
#include <hip/hip_runtime.h>
#include "qm_device.h"

namespace qm {

// The same pointer as a global one, for loads, stores and atomics on result
// columns, group tables and counters (device global memory): a flat access
// would also count against every LDS wait of the kernel. Only for memory
// that is never LDS.
template <class T> __device__ __forceinline__ __attribute__((address_space(1))) T *gptr(T *p) {
  return (__attribute__((address_space(1))) T *)p;
}
// device-scope relaxed atomics on global memory
__device__ __forceinline__ uint64_t g_atomic_add(uint64_t *p, uint64_t v) {
  return __hip_atomic_fetch_add(gptr(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void g_atomic_max(uint64_t *p, uint64_t v) {
  __hip_atomic_fetch_max(gptr(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void g_atomic_or(uint64_t *p, uint64_t v) {
  __hip_atomic_fetch_or(gptr(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void g_atomic_or(uint32_t *p, uint32_t v) {
  __hip_atomic_fetch_or(gptr(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t g_atomic_cas(uint64_t *p, uint64_t expected, uint64_t desired) {
  __hip_atomic_compare_exchange_strong(gptr(p), &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return expected;  // old value
}
__device__ __forceinline__ uint64_t g_load(const uint64_t *p) {
  return __hip_atomic_load(gptr(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// 128-bit two's-complement add as two 64-bit atomics: the carry out of the
// low word is computed from the value that atomic saw, so the total is exact
// whatever the interleaving.
__device__ __forceinline__ void g_add128(uint64_t *p, uint64_t lo, uint64_t hi) {
  if (lo) {
    uint64_t old = g_atomic_add(p, lo);
    hi += (old + lo < old) ? 1 : 0;
  }
  if (hi) g_atomic_add(p + 1, hi);
}

// LDS variants: address-space-3 pointers so the compiler emits ds_* atomics.
typedef __attribute__((address_space(3))) uint64_t lds_u64;
__device__ __forceinline__ uint64_t l_atomic_add(lds_u64 *p, uint64_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void l_atomic_max(lds_u64 *p, uint64_t v) {
  __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void l_atomic_or(lds_u64 *p, uint64_t v) {
  __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint64_t l_atomic_cas(lds_u64 *p, uint64_t expected, uint64_t desired) {
  __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
  return expected;
}
__device__ __forceinline__ void l_add128(lds_u64 *p, uint64_t lo, uint64_t hi) {
  if (lo) {
    uint64_t old = l_atomic_add(p, lo);
    hi += (old + lo < old) ? 1 : 0;
  }
  if (hi) l_atomic_add(p + 1, hi);
}

// ---- 8-row loads. row0 is a multiple of 8 and columns are padded, so these
// are aligned 64/128-bit loads that never leave the allocation. ----
// Column memory (tables, results, join payloads, IN lists) is device global
// memory. The pointers reach the kernels inside descriptor structs, where the
// compiler cannot tell their address space; gld() loads through a global
// pointer, so these are global loads (a flat load would also count against
// every LDS wait of the kernel).
typedef uint32_t gv2u __attribute__((ext_vector_type(2)));
typedef uint32_t gv4u __attribute__((ext_vector_type(4)));
typedef int32_t gv4i __attribute__((ext_vector_type(4)));
typedef int64_t gv2l __attribute__((ext_vector_type(2)));
typedef float gv4f __attribute__((ext_vector_type(4)));
typedef double gv2d __attribute__((ext_vector_type(2)));
template <class V> __device__ __forceinline__ V gld(const void *p, uint64_t byte_off) {
  using G8 = const __attribute__((address_space(1))) uint8_t;
  using GV = const __attribute__((address_space(1))) V;
  return *(GV *)((G8 *)p + byte_off);
}

__device__ __forceinline__ void load8_i(const DCol &c, uint64_t row0, int64_t v[RPT]) {
  switch (c.dtype) {
    case DT_I8: {
      const gv2u w = gld<gv2u>(c.data, row0);
#pragma unroll
      for (int r = 0; r < 4; ++r) v[r] = (int8_t)((w.x >> (8 * r)) & 0xFF);
#pragma unroll
      for (int r = 0; r < 4; ++r) v[4 + r] = (int8_t)((w.y >> (8 * r)) & 0xFF);
      break;
    }
    case DT_BOOL: {
      const gv2u w = gld<gv2u>(c.data, row0);
#pragma unroll
      for (int r = 0; r < 4; ++r) v[r] = (w.x >> (8 * r)) & 0xFF;
#pragma unroll
      for (int r = 0; r < 4; ++r) v[4 + r] = (w.y >> (8 * r)) & 0xFF;
      break;
    }
    case DT_I16: {
      const gv4u w = gld<gv4u>(c.data, row0 * 2);
      uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        v[2 * r] = (int16_t)(ws[r] & 0xFFFF);
        v[2 * r + 1] = (int16_t)(ws[r] >> 16);
      }
      break;
    }
    case DT_I32: {
      const gv4i a = gld<gv4i>(c.data, row0 * 4), b = gld<gv4i>(c.data, row0 * 4 + 16);
      v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
      v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
      break;
    }
    default: {  // DT_I64
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const gv2l x = gld<gv2l>(c.data, row0 * 8 + 16 * r);
        v[2 * r] = x.x;
        v[2 * r + 1] = x.y;
      }
      break;
    }
  }
}

__device__ __forceinline__ void load8_f(const DCol &c, uint64_t row0, double v[RPT]) {
  if (c.dtype == DT_F32) {
    const gv4f a = gld<gv4f>(c.data, row0 * 4), b = gld<gv4f>(c.data, row0 * 4 + 16);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const gv2d x = gld<gv2d>(c.data, row0 * 8 + 16 * r);
      v[2 * r] = x.x;
      v[2 * r + 1] = x.y;
    }
  }
}

__device__ __forceinline__ bool dt_is_float(int32_t dt) { return dt == DT_F64 || dt == DT_F32; }

// 8 null bits (bit r = row0 + r is null). row0 % 8 == 0 -> one byte.
__device__ __forceinline__ uint32_t load8_nulls(const DCol &c, uint64_t row0) {
  return c.nulls ? (uint32_t)gld<uint8_t>(c.nulls, row0 >> 3) : 0u;
}

// single-row loads (gathers: join build side, compaction)
__device__ __forceinline__ int64_t load1_i(const DCol &c, uint64_t row) {
  switch (c.dtype) {
    case DT_I8: return gld<int8_t>(c.data, row);
    case DT_BOOL: return gld<uint8_t>(c.data, row);
    case DT_I16: return gld<int16_t>(c.data, row * 2);
    case DT_I32: return gld<int32_t>(c.data, row * 4);
    default: return gld<int64_t>(c.data, row * 8);
  }
}
__device__ __forceinline__ double load1_f(const DCol &c, uint64_t row) {
  if (c.dtype == DT_F32) return gld<float>(c.data, row * 4);
  return gld<double>(c.data, row * 8);
}
__device__ __forceinline__ bool load1_null(const DCol &c, uint64_t row) {
  return c.nulls ? ((gld<uint8_t>(c.nulls, row >> 3) >> (row & 7)) & 1) : false;
}

// ---- predicate evaluation on one value ----
__device__ __forceinline__ bool pred_int(const DPred &p, int64_t v) {
  bool in;
  if (p.kind == PK_INT_RANGE) {
    in = (v >= p.ilo) && (v <= p.ihi);
  } else {  // PK_INT_IN: binary search in the sorted list
    int lo = 0, hi = p.list_n;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (gld<int64_t>(p.list, (uint64_t)mid * 8) < v) lo = mid + 1; else hi = mid;
    }
    in = (lo < p.list_n) && (gld<int64_t>(p.list, (uint64_t)lo * 8) == v);
  }
  return in != (p.negate != 0);
}

__device__ __forceinline__ bool pred_f64(const DPred &p, double v) {
  if (v != v) {
    if (p.kind == PK_F64_IN) return (p.list_has_nan != 0) != (p.negate != 0);
    return p.nan_result != 0;
  }
  bool in;
  if (p.kind == PK_F64_RANGE) {
    bool a = p.lo_strict ? (v > p.flo) : (v >= p.flo);
    bool b = p.hi_strict ? (v < p.fhi) : (v <= p.fhi);
    in = a && b;
  } else {
    int lo = 0, hi = p.list_n;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (gld<double>(p.list, (uint64_t)mid * 8) < v) lo = mid + 1; else hi = mid;
    }
    in = (lo < p.list_n) && (gld<double>(p.list, (uint64_t)lo * 8) == v);
  }
  return in != (p.negate != 0);
}

// Evaluate one predicate over the thread's 8 rows -> 8-bit mask.
__device__ __forceinline__ uint32_t eval_pred8(const DPred &p, const DCol &c, uint64_t row0) {
  if (p.kind == PK_FALSE) return 0;
  uint32_t nm = load8_nulls(c, row0);
  if (p.kind == PK_IS_NULL) return nm;
  if (p.kind == PK_IS_NOT_NULL) return (~nm) & 0xFFu;
  uint32_t m = 0;
  if (p.kind == PK_F64_RANGE || p.kind == PK_F64_IN) {
    double v[RPT];
    load8_f(c, row0, v);
#pragma unroll
    for (int r = 0; r < RPT; ++r) m |= (pred_f64(p, v[r]) ? 1u : 0u) << r;
  } else {
    int64_t v[RPT];
    load8_i(c, row0, v);
#pragma unroll
    for (int r = 0; r < RPT; ++r) m |= (pred_int(p, v[r]) ? 1u : 0u) << r;
  }
  return m & ~nm;  // null never matches (src/decode.ts:489-491)
}

// One predicate on 8 values already in registers (int64 values or f64 bits).
__device__ __forceinline__ uint32_t eval_pred8_regs(const DPred &p, const int64_t v[RPT], uint32_t nm) {
  if (p.kind == PK_FALSE) return 0;
  if (p.kind == PK_IS_NULL) return nm;
  if (p.kind == PK_IS_NOT_NULL) return (~nm) & 0xFFu;
  uint32_t m = 0;
  if (p.kind == PK_F64_RANGE || p.kind == PK_F64_IN) {
#pragma unroll
    for (int r = 0; r < RPT; ++r) m |= (pred_f64(p, bits_f64((uint64_t)v[r])) ? 1u : 0u) << r;
  } else {
#pragma unroll
    for (int r = 0; r < RPT; ++r) m |= (pred_int(p, v[r]) ? 1u : 0u) << r;
  }
  return m & ~nm;
}

// AND-only queries over <= MAX_PSLOTS columns: every predicate column is
// loaded before any compare, so a tile costs one memory round trip.
__device__ __forceinline__ uint32_t eval_and8_fast(const DQuery &q, const DCol *cols, uint64_t row0,
                                                   uint32_t valid) {
  if (!valid) return 0;
  int64_t raw[MAX_PSLOTS][RPT];
  uint32_t nm[MAX_PSLOTS];
#pragma unroll
  for (int s = 0; s < MAX_PSLOTS; ++s) {
    nm[s] = 0;
    if (s < q.nslots) {
      const DCol &c = cols[q.slot_col[s]];
      nm[s] = load8_nulls(c, row0);
      if (dt_is_float(c.dtype)) {
        double f[RPT];
        load8_f(c, row0, f);
#pragma unroll
        for (int r = 0; r < RPT; ++r) raw[s][r] = (int64_t)f64_bits(f[r]);
      } else {
        load8_i(c, row0, raw[s]);
      }
    }
  }
  uint32_t m = valid;
#pragma unroll
  for (int s = 0; s < MAX_PSLOTS; ++s) {
    if (s >= q.nslots) break;
    for (int i = 0; i < q.npreds; ++i) {
      const DPred &p = q.preds[i];
      if (p.slot == s) m &= eval_pred8_regs(p, raw[s], nm[s]);
    }
  }
  return m;
}

// AND of group-0 predicates, AND with OR over groups 1..ngroups.
__device__ __forceinline__ uint32_t eval_query8(const DQuery &q, const DCol *cols, uint64_t row0,
                                                uint32_t valid) {
  if (q.and_only) return eval_and8_fast(q, cols, row0, valid);
  uint32_t m = valid;
  uint32_t orm = 0;
  int cur_group = -1;
  uint32_t gm = 0;
  for (int i = 0; i < q.npreds; ++i) {
    const DPred &p = q.preds[i];
    if (p.group == 0) {
      if (m) m &= eval_pred8(p, cols[p.col], row0);
    } else {
      if ((int)p.group != cur_group) {
        if (cur_group > 0) orm |= gm;
        cur_group = p.group;
        gm = m;
      }
      if (gm) gm &= eval_pred8(p, cols[p.col], row0);
    }
  }
  if (q.ngroups > 0) {
    if (cur_group > 0) orm |= gm;
    m &= orm;
  }
  return m;
}

// ---- zone maps: a tile starting at `row` lies in a page the query skips ----
// Interval of an order code x among sorted bounds (#{bounds <= x}) through a
// slot table: codes [lo + (k << shift), lo + ((k + 1) << shift)) lie in
// buckets ts[k] .. ts[k + 1] (ts[k] = the bounds below the slot's first code),
// so only the few bounds inside x's slot are searched (usually 0 to 2 steps)
// instead of all of them (the partitioned-runs row scatter).
__device__ __forceinline__ uint32_t bucket_slot(const uint64_t *sb, const uint16_t *ts, uint64_t lo, int shift,
                                                uint32_t slots, uint32_t nb, uint64_t x) {
  if (nb == 0 || x < lo) return 0;
  const uint64_t k = (x - lo) >> shift;
  if (k >= slots) return nb;
  uint32_t f = ts[k], h = ts[k + 1];
  while (f < h) {  // first bound above x in [f, h)
    const uint32_t m = (f + h) >> 1;
    if (sb[m] <= x) f = m + 1;
    else h = m;
  }
  return f;
}

__device__ __forceinline__ bool page_bit(const uint32_t *bits, uint64_t row) {
  if (!bits) return false;
  const uint64_t p = row / ZONE_ROWS;
  return (gld<uint32_t>(bits, (p >> 5) * 4) >> (p & 31)) & 1u;
}
__device__ __forceinline__ bool page_skipped(const DScan &sc, uint64_t row) { return page_bit(sc.page_skip, row); }

// ---- join probe: returns the build-row range of a probe key ----
__device__ __forceinline__ bool join_lookup(const DScan &sc, int64_t key, uint32_t *start,
                                            uint32_t *count) {
  if (sc.join_mode == 1) {
    uint64_t idx = (uint64_t)(key - sc.join_min);
    if (idx >= sc.join_cap) return false;
    uint64_t v = sc.join_table[idx];
    *start = (uint32_t)(v >> 32); *count = (uint32_t)v;
    return *count != 0;
  }
  uint64_t stored = (uint64_t)key + KEY_SALT;
  if (stored == 0) {  // reserved side entry after the table
    uint64_t v = sc.join_table[2 * sc.join_cap + 1];
    *start = (uint32_t)(v >> 32); *count = (uint32_t)v;
    return *count != 0;
  }
  const uint64_t mask = sc.join_cap - 1;
  uint64_t h = mix64(stored) & mask;
  for (uint64_t p = 0; p <= mask; ++p) {
    uint64_t slot = (h + p) & mask;
    uint64_t k = sc.join_table[2 * slot];
    if (k == stored) {
      uint64_t v = sc.join_table[2 * slot + 1];
      *start = (uint32_t)(v >> 32); *count = (uint32_t)v;
      return true;
    }
    if (k == 0) return false;
  }
  return false;
}

}  // namespace qm

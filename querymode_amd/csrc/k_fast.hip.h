// k_fast.hip.h — register-resident query descriptors for the hot kernels.
//
// The generic descriptors live in global memory and are walked with runtime
// loops, which costs scalar loads and dependent branches inside the tile
// loop. For the common shapes (AND of one range per column over <= 4 columns,
// <= 2 key columns, <= 4 value sources) the kernels copy the fields they need
// into these fixed-size structs once, before the tile loop; every loop over
// them is unrolled over the fixed bound, so the fields stay in (scalar)
// registers and all loads of a tile are issued before the first compare.
#pragma once
#include "k_agg.hip.h"

namespace qm {

static constexpr int MAX_FSRC = 4;

// Values known to be uniform across the wave, pinned to scalar registers.
__device__ __forceinline__ int32_t u32s(int32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ int64_t u64s(int64_t x) {
  uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)x);
  uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)((uint64_t)x >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double f64s(double x) { return bits_f64((uint64_t)u64s((int64_t)f64_bits(x))); }
// A column pointer pinned to scalar registers. The integer round trip would
// hide that it points to global memory (loads through it would become flat
// loads, which also count against LDS waits), so it goes back to a pointer
// through the global address space.
template <class T> __device__ __forceinline__ T *ptrs(T *p) {
  using G = __attribute__((address_space(1))) T;
  G *g = (G *)(uintptr_t)u64s((int64_t)(uintptr_t)p);
  return (T *)g;
}

struct FastFilter {
  int n;
  const void *data[MAX_PSLOTS];
  const uint8_t *nulls[MAX_PSLOTS];
  int32_t dtype[MAX_PSLOTS];
  uint8_t kind[MAX_PSLOTS];
  uint8_t negate[MAX_PSLOTS], nan_result[MAX_PSLOTS], lo_strict[MAX_PSLOTS], hi_strict[MAX_PSLOTS];
  int64_t ilo[MAX_PSLOTS], ihi[MAX_PSLOTS];
  double flo[MAX_PSLOTS], fhi[MAX_PSLOTS];
};

__device__ __forceinline__ FastFilter fast_filter_of(const DQuery &q, const DCol *cols) {
  FastFilter f;
  f.n = q.nslots;
#pragma unroll
  for (int s = 0; s < MAX_PSLOTS; ++s) {
    f.data[s] = nullptr; f.nulls[s] = nullptr; f.dtype[s] = DT_I64; f.kind[s] = PK_IS_NOT_NULL;
    f.negate[s] = 0; f.nan_result[s] = 0; f.lo_strict[s] = 0; f.hi_strict[s] = 0;
    f.ilo[s] = 0; f.ihi[s] = 0; f.flo[s] = 0; f.fhi[s] = 0;
    if (s < q.nslots) {
      const DCol &c = cols[q.slot_col[s]];
      f.data[s] = ptrs(c.data); f.nulls[s] = ptrs(c.nulls); f.dtype[s] = u32s(c.dtype);
    }
  }
  // exactly one predicate per slot (the planner merges ranges per column)
  for (int i = 0; i < q.npreds; ++i) {
    const DPred &p = q.preds[i];
#pragma unroll
    for (int s = 0; s < MAX_PSLOTS; ++s) {
      if (p.slot == s) {
        f.kind[s] = p.kind; f.negate[s] = p.negate; f.nan_result[s] = p.nan_result;
        f.lo_strict[s] = p.lo_strict; f.hi_strict[s] = p.hi_strict;
        f.ilo[s] = p.ilo; f.ihi[s] = p.ihi; f.flo[s] = p.flo; f.fhi[s] = p.fhi;
      }
    }
  }
#pragma unroll
  for (int s = 0; s < MAX_PSLOTS; ++s) {
    f.kind[s] = (uint8_t)u32s(f.kind[s]); f.negate[s] = (uint8_t)u32s(f.negate[s]);
    f.nan_result[s] = (uint8_t)u32s(f.nan_result[s]);
    f.lo_strict[s] = (uint8_t)u32s(f.lo_strict[s]); f.hi_strict[s] = (uint8_t)u32s(f.hi_strict[s]);
    f.ilo[s] = u64s(f.ilo[s]); f.ihi[s] = u64s(f.ihi[s]); f.flo[s] = f64s(f.flo[s]); f.fhi[s] = f64s(f.fhi[s]);
  }
  return f;
}

__device__ __forceinline__ void load8_any(const void *data, int32_t dtype, uint64_t row0, int64_t v[RPT]) {
  DCol c;
  c.data = data; c.nulls = nullptr; c.dtype = dtype; c.side = 0;
  if (dt_is_float(dtype)) {
    double f[RPT];
    load8_f(c, row0, f);
#pragma unroll
    for (int r = 0; r < RPT; ++r) v[r] = (int64_t)f64_bits(f[r]);
  } else {
    load8_i(c, row0, v);
  }
}

__device__ __forceinline__ uint32_t nulls8(const uint8_t *nulls, uint64_t row0) {
  return nulls ? (uint32_t)nulls[row0 >> 3] : 0u;
}

// loads of the NS predicate slots (issue before any compare)
template <int NS>
__device__ __forceinline__ void fast_load(const FastFilter &f, uint64_t row0, int64_t pv[NS][RPT], uint32_t pn[NS]) {
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    load8_any(f.data[s], f.dtype[s], row0, pv[s]);
    pn[s] = nulls8(f.nulls[s], row0);
  }
}

template <int NS>
__device__ __forceinline__ uint32_t fast_eval(const FastFilter &f, const int64_t pv[NS][RPT], const uint32_t pn[NS],
                                              uint32_t valid) {
  uint32_t m = valid;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const uint8_t k = f.kind[s];
    uint32_t r8 = 0;
    if (k == PK_FALSE) {
      r8 = 0;
    } else if (k == PK_IS_NULL) {
      r8 = pn[s];
    } else if (k == PK_IS_NOT_NULL) {
      r8 = ~pn[s] & 0xFFu;
    } else if (k == PK_F64_RANGE) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        double v = bits_f64((uint64_t)pv[s][r]);
        bool a = f.lo_strict[s] ? (v > f.flo[s]) : (v >= f.flo[s]);
        bool b = f.hi_strict[s] ? (v < f.fhi[s]) : (v <= f.fhi[s]);
        bool in = (a && b) != (f.negate[s] != 0);
        if (v != v) in = f.nan_result[s] != 0;
        r8 |= (in ? 1u : 0u) << r;
      }
      r8 &= ~pn[s];
    } else {  // PK_INT_RANGE
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        bool in = (pv[s][r] >= f.ilo[s]) && (pv[s][r] <= f.ihi[s]);
        r8 |= ((in != (f.negate[s] != 0)) ? 1u : 0u) << r;
      }
      r8 &= ~pn[s];
    }
    m &= r8;
  }
  return m;
}

struct FastKey {
  int n;
  uint8_t direct;
  const void *data[MAX_KEYS];
  const uint8_t *nulls[MAX_KEYS];
  int32_t dtype[MAX_KEYS];
  uint8_t is_float[MAX_KEYS];
};

__device__ __forceinline__ FastKey fast_key_of(const DQuery &q, const DCol *cols) {
  FastKey k;
  k.n = q.key.ncols;
  k.direct = q.gmode == GM_DIRECT;
#pragma unroll
  for (int i = 0; i < MAX_KEYS; ++i) {
    k.data[i] = nullptr; k.nulls[i] = nullptr; k.dtype[i] = DT_I64; k.is_float[i] = 0;
    if (i < q.key.ncols) {
      const DCol &c = cols[q.key.col[i]];
      k.data[i] = ptrs(c.data); k.nulls[i] = ptrs(c.nulls); k.dtype[i] = u32s(c.dtype);
      k.is_float[i] = (uint8_t)u32s(q.key.is_float[i]);
    }
  }
  return k;
}

template <int NK>
__device__ __forceinline__ void fast_key_load(const FastKey &k, uint64_t row0, int64_t kv[NK][RPT], uint32_t kn[NK]) {
#pragma unroll
  for (int i = 0; i < NK; ++i) {
    load8_any(k.data[i], k.dtype[i], row0, kv[i]);
    if (k.is_float[i]) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) kv[i][r] = (int64_t)canon_f64_key_bits((uint64_t)kv[i][r]);
    }
    kn[i] = nulls8(k.nulls[i], row0);
  }
}

// key code with the column count known at compile time (no dynamic indexing)
template <int NK>
__device__ __forceinline__ uint64_t key_code_fixed(const DKey &k, bool direct, const int64_t v[MAX_KEYS],
                                                   const uint32_t n[MAX_KEYS], int *side) {
  *side = -1;
  if (direct) {
    uint64_t idx = 0;
#pragma unroll
    for (int i = 0; i < NK; ++i) {
      uint64_t code = n[i] ? (k.dom[i] - 1) : (uint64_t)(v[i] - k.kmin[i]);
      idx += code * k.stride[i];
    }
    return idx;
  }
  if (NK == 1) {
    if (n[0]) { *side = 1; return 0; }
    uint64_t st = (uint64_t)v[0] + KEY_SALT;
    if (st == 0) { *side = 0; return 0; }
    return st;
  }
  uint64_t packed = 0;
  int shift = 0;
#pragma unroll
  for (int i = 0; i < NK; ++i) {
    uint64_t code = n[i] ? (k.dom[i] - 1) : (uint64_t)(v[i] - k.kmin[i]);
    packed |= code << shift;
    shift += k.bits[i];
  }
  return packed + KEY_SALT;
}

// value sources of a partitioned record; a source that is also a predicate
// slot reuses the slot's registers
struct FastSrc {
  int n;
  const void *a[MAX_FSRC];
  const void *b[MAX_FSRC];
  const uint8_t *na[MAX_FSRC];
  const uint8_t *nb[MAX_FSRC];
  int32_t dta[MAX_FSRC], dtb[MAX_FSRC];
  int8_t slot[MAX_FSRC];  // predicate slot holding column a (plain sources only), or -1
};

__device__ __forceinline__ FastSrc fast_src_of(const DPart &pp, const DQuery &q, const DCol *cols) {
  FastSrc s;
  s.n = pp.nsrc;
#pragma unroll
  for (int j = 0; j < MAX_FSRC; ++j) {
    s.a[j] = nullptr; s.b[j] = nullptr; s.na[j] = nullptr; s.nb[j] = nullptr;
    s.dta[j] = DT_I64; s.dtb[j] = DT_I64; s.slot[j] = -1;
    if (j < pp.nsrc) {
      const DCol &ca = cols[pp.src_a[j]];
      s.a[j] = ptrs(ca.data); s.na[j] = ptrs(ca.nulls); s.dta[j] = u32s(ca.dtype);
      if (pp.src_b[j] >= 0) {
        const DCol &cb = cols[pp.src_b[j]];
        s.b[j] = ptrs(cb.data); s.nb[j] = ptrs(cb.nulls); s.dtb[j] = u32s(cb.dtype);
      } else {
        for (int t = 0; t < q.nslots; ++t)
          if (q.slot_col[t] == pp.src_a[j]) s.slot[j] = (int8_t)t;
      }
      s.slot[j] = (int8_t)u32s(s.slot[j]);
    }
  }
  return s;
}

}  // namespace qm

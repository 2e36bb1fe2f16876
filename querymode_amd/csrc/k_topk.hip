// k_topk.hip — ORDER BY col [desc] LIMIT k (+ OFFSET) as a one-pass
// per-workgroup selection, then a single-workgroup merge of the candidates.
//
// Replaces TopKOperator (src/operators.ts:1716-1841). The reference keeps a
// binary heap of k+offset rows and replaces the root only on a strictly
// better value, so its tie order depends on the heap; here the order is the
// total order (value [desc], row id asc) with NULL and NaN last, as
// rowComparator puts them (src/types.ts:26-38).
//
// Each workgroup keeps a candidate buffer in LDS and a threshold: rows
// whose composite key is not below the threshold are rejected in registers;
// when the buffer could overflow it is bitonic-sorted in LDS, cut to k and
// the threshold tightened to the k-th key.
#include "k_common.hip.h"
#include "qm_kernels.h"

namespace qm {

static constexpr int TK_BLOCK = 256;
static constexpr uint32_t TK_TILE = TK_BLOCK * RPT;  // rows per tile
static constexpr uint32_t TK_MAX_CAP = 8192;

__device__ __forceinline__ bool key_less(const TopKey &x, const TopKey &y) {
  return x.a < y.a || (x.a == y.a && x.b < y.b);
}

__device__ __forceinline__ TopKey order_key_i(int64_t v, bool null, uint64_t row, int desc) {
  TopKey k;
  if (null) { k.a = ~0ull; k.b = (1ull << 48) | row; return k; }
  uint64_t c = code_i64(v);
  k.a = desc ? ~c : c; k.b = row;
  return k;
}
__device__ __forceinline__ TopKey order_key_f(double v, bool null, uint64_t row, int desc) {
  TopKey k;
  if (null || v != v) { k.a = ~0ull; k.b = (1ull << 48) | row; return k; }
  uint64_t c = code_f64_bits(f64_bits(v));
  k.a = desc ? ~c : c; k.b = row;
  return k;
}

// bitonic sort of buf[0..cap) ascending (cap a power of two)
__device__ void lds_bitonic(TopKey *buf, uint32_t cap) {
  for (uint32_t k = 2; k <= cap; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < cap; i += blockDim.x) {
        uint32_t ixj = i ^ j;
        if (ixj > i) {
          TopKey x = buf[i], y = buf[ixj];
          bool up = (i & k) == 0;
          bool sw = up ? key_less(y, x) : key_less(x, y);
          if (sw) { buf[i] = y; buf[ixj] = x; }
        }
      }
      __syncthreads();
    }
  }
}

// pad [count, cap) with +inf keys, sort, keep k; returns new count
__device__ uint32_t prune(TopKey *buf, uint32_t cap, uint32_t count, uint32_t k, TopKey *thr) {
  for (uint32_t i = count + threadIdx.x; i < cap; i += blockDim.x) { buf[i].a = ~0ull; buf[i].b = ~0ull; }
  __syncthreads();
  lds_bitonic(buf, cap);
  uint32_t n = count < k ? count : k;
  if (threadIdx.x == 0) {
    if (n == k) *thr = buf[k - 1];
    else { thr->a = ~0ull; thr->b = ~0ull; }
  }
  __syncthreads();
  return n;
}

// gthr: the smallest main order word of any workgroup's k-th candidate so far
// (~0: none yet). A row whose word is above it is worse than k rows of that
// workgroup, so no workgroup keeps it; each thread reads it one tile ahead.
__global__ void __launch_bounds__(TK_BLOCK) topk_scan_kernel(DScan sc, const DQuery *__restrict__ qs, int sort_col,
                                                             int desc, uint32_t k, uint32_t cap, TopKey *cand,
                                                             uint32_t *cand_counts, uint64_t *gthr) {
  extern __shared__ uint64_t lds_raw[];
  TopKey *buf = reinterpret_cast<TopKey *>(lds_raw);
  __shared__ uint32_t s_count;
  __shared__ TopKey s_thr;
  if (threadIdx.x == 0) { s_count = 0; s_thr.a = ~0ull; s_thr.b = ~0ull; }
  __syncthreads();
  const DQuery &q = qs[0];
  const DCol &c = sc.cols[sort_col];
  const bool isf = dt_is_float(c.dtype);
  const uint64_t ntiles = (sc.nrows + TK_TILE - 1) / TK_TILE;
  uint32_t count = 0;
  uint64_t ga_next = ~0ull;
  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint64_t ga = ga_next;
    ga_next = __hip_atomic_load(gthr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t row0 = tile * TK_TILE + (uint64_t)threadIdx.x * RPT;
    uint32_t valid = 0;
    if (row0 < sc.nrows) {
      uint64_t left = sc.nrows - row0;
      valid = left >= RPT ? 0xFFu : ((1u << left) - 1);
    }
    // a page the zone maps rule out holds no candidate: nothing of it is read
    const bool skip = page_skipped(sc, tile * TK_TILE);
    uint32_t m = skip ? 0u : ((q.npreds > 0) ? eval_query8(q, sc.cols, row0, valid) : valid);
    if (m) {
      TopKey thr = s_thr;
      if (ga != ~0ull) {
        TopKey gt;
        gt.a = ga + 1;
        gt.b = 0;
        if (key_less(gt, thr)) thr = gt;
      }
      uint32_t nm = load8_nulls(c, row0);
      TopKey kk[RPT];
      if (isf) {
        double v[RPT]; load8_f(c, row0, v);
#pragma unroll
        for (int r = 0; r < RPT; ++r) kk[r] = order_key_f(v[r], (nm >> r) & 1, row0 + r, desc);
      } else {
        int64_t v[RPT]; load8_i(c, row0, v);
#pragma unroll
        for (int r = 0; r < RPT; ++r) kk[r] = order_key_i(v[r], (nm >> r) & 1, row0 + r, desc);
      }
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        if (((m >> r) & 1) && key_less(kk[r], thr)) {
          uint32_t pos = atomicAdd(&s_count, 1u);
          buf[pos] = kk[r];
        }
      }
    }
    __syncthreads();
    count = s_count;
    if (count + TK_TILE > cap) {
      count = prune(buf, cap, count, k, &s_thr);
      if (threadIdx.x == 0) {
        s_count = count;
        if (s_thr.a != ~0ull) __hip_atomic_fetch_min(gthr, s_thr.a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      __syncthreads();
    }
  }
  __syncthreads();
  count = prune(buf, cap, s_count, k, &s_thr);
  TopKey *out = cand + (uint64_t)blockIdx.x * k;
  for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) out[i] = buf[i];
  if (threadIdx.x == 0) cand_counts[blockIdx.x] = count;
}

// workgroup b merges lists [b*lpb, (b+1)*lpb) of <= k candidates each into
// out + b*k / out_count[b]
__global__ void __launch_bounds__(TK_BLOCK) topk_merge_kernel(const TopKey *cand, const uint32_t *counts,
                                                              uint32_t nlists, uint32_t lpb, uint32_t k,
                                                              uint32_t cap, TopKey *out, uint32_t *out_count) {
  extern __shared__ uint64_t lds_raw[];
  TopKey *buf = reinterpret_cast<TopKey *>(lds_raw);
  __shared__ uint32_t s_count;
  __shared__ TopKey s_thr;
  if (threadIdx.x == 0) { s_count = 0; s_thr.a = ~0ull; s_thr.b = ~0ull; }
  __syncthreads();
  const uint32_t l0 = blockIdx.x * lpb;
  const uint32_t l1 = (l0 + lpb < nlists) ? l0 + lpb : nlists;
  for (uint32_t l = l0; l < l1; ++l) {
    const uint32_t n = counts[l];
    const TopKey *src = cand + (uint64_t)l * k;
    for (uint32_t i0 = 0; i0 < n; i0 += TK_BLOCK) {
      uint32_t i = i0 + threadIdx.x;
      TopKey thr = s_thr;
      if (i < n) {
        TopKey x = src[i];
        if (key_less(x, thr)) { uint32_t pos = atomicAdd(&s_count, 1u); buf[pos] = x; }
      }
      __syncthreads();
      uint32_t count = s_count;
      if (count + TK_BLOCK > cap) {
        count = prune(buf, cap, count, k, &s_thr);
        if (threadIdx.x == 0) s_count = count;
        __syncthreads();
      }
    }
  }
  __syncthreads();
  uint32_t count = prune(buf, cap, s_count, k, &s_thr);
  for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) out[(uint64_t)blockIdx.x * k + i] = buf[i];
  if (threadIdx.x == 0) out_count[blockIdx.x] = count;
}

__global__ void topk_rows_kernel(const TopKey *keys, const uint32_t *count, uint32_t offset, uint64_t *rows,
                                 uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    uint32_t c = *count;
    rows[i] = (offset + i < c) ? (keys[offset + i].b & ((1ull << 48) - 1)) : ~0ull;
  }
}

}  // namespace qm

using namespace qm;

extern "C" uint32_t qmk_topk_cap(uint32_t k) {
  uint32_t need = k + TK_TILE;
  uint32_t cap = 1;
  while (cap < need) cap <<= 1;
  return cap <= TK_MAX_CAP ? cap : 0;
}

extern "C" hipError_t qmk_topk_scan(const DScan *sc, const DQuery *d_q, int sort_col, int desc, uint32_t k,
                                    int grid, TopKey *cand, uint32_t *cand_counts, uint64_t *gthr, hipStream_t st) {
  if (!gthr) return hipErrorInvalidValue;
  uint32_t cap = qmk_topk_cap(k);
  if (!cap || k == 0) return hipErrorInvalidValue;
  if ((size_t)cap * sizeof(TopKey) > 65536) {
    hipError_t e = qmk_func_lds((const void *)topk_scan_kernel, (size_t)cap * sizeof(TopKey));
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(topk_scan_kernel, dim3(grid), dim3(TK_BLOCK), (size_t)cap * sizeof(TopKey), st, *sc, d_q,
                     sort_col, desc, k, cap, cand, cand_counts, gthr);
  return hipGetLastError();
}

extern "C" hipError_t qmk_topk_merge(const TopKey *cand, const uint32_t *cand_counts, uint32_t nlists,
                                     uint32_t lists_per_block, uint32_t k, TopKey *out, uint32_t *out_count,
                                     hipStream_t st) {
  uint32_t need = k + TK_BLOCK, cap = 1;
  while (cap < need) cap <<= 1;
  if (cap > TK_MAX_CAP || lists_per_block == 0) return hipErrorInvalidValue;
  uint32_t grid = (nlists + lists_per_block - 1) / lists_per_block;
  if (grid == 0) grid = 1;
  if ((size_t)cap * sizeof(TopKey) > 65536) {
    hipError_t e = qmk_func_lds((const void *)topk_merge_kernel, (size_t)cap * sizeof(TopKey));
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(topk_merge_kernel, dim3(grid), dim3(TK_BLOCK), (size_t)cap * sizeof(TopKey), st, cand,
                     cand_counts, nlists, lists_per_block, k, cap, out, out_count);
  return hipGetLastError();
}

extern "C" hipError_t qmk_topk_rows(const TopKey *keys, const uint32_t *count, uint32_t offset, uint64_t *rows,
                                    uint32_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(topk_rows_kernel, dim3((n + 255) / 256), dim3(256), 0, st, keys, count, offset, rows, n);
  return hipGetLastError();
}

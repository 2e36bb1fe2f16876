// k_sort.hip — full ORDER BY without LIMIT (InMemorySortOperator,
// src/operators.ts:1847-1889): a stable device radix sort of order codes
// (rocPRIM), NULL/NaN last, ties in input order — the order V8's stable
// Array.prototype.sort gives with rowComparator (src/types.ts:26-38).
#include <cstring>
#include "k_common.hip.h"
#include "qm_kernels.h"

namespace qm {

__global__ void sort_keys_kernel(DCol c, uint64_t n, int desc, uint64_t *keys, uint64_t *vals, uint8_t *cls) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    bool null = load1_null(c, i);
    uint64_t code;
    if (dt_is_float(c.dtype)) {
      double v = load1_f(c, i);
      if (v != v) null = true;
      code = code_f64_bits(f64_bits(v));
    } else {
      code = code_i64(load1_i(c, i));
    }
    keys[i] = null ? 0 : (desc ? ~code : code);
    cls[i] = null ? 1 : 0;
    vals[i] = i;
  }
}

__global__ void gather_cls_kernel(const uint8_t *cls, const uint64_t *idx, uint64_t n, uint8_t *out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = cls[idx[i]];
}

// ---- the same sort in key-range passes (when its temporaries exceed the budget) ----
__device__ __forceinline__ bool sort_key_of(const DCol &c, uint64_t i, int desc, uint64_t *key) {
  bool null = load1_null(c, i);
  uint64_t code;
  if (dt_is_float(c.dtype)) {
    const double v = load1_f(c, i);
    if (v != v) null = true;
    code = code_f64_bits(f64_bits(v));
  } else {
    code = code_i64(load1_i(c, i));
  }
  *key = desc ? ~code : code;
  return !null;
}

// rows per 2^16 bins of the top key bits (non-NULL, non-NaN rows)
__global__ void sort_hist_kernel(DCol c, uint64_t n, int desc, uint32_t *hist) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint64_t k;
    if (sort_key_of(c, i, desc, &k)) atomicAdd(&hist[k >> 48], 1u);
  }
}

// keep[i]: the row's key lies in [lo, hi] (NULL / NaN rows: `nulls` selects them instead)
__global__ void sort_range_kernel(DCol c, uint64_t n, int desc, uint64_t lo, uint64_t hi, int nulls, uint8_t *keep) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint64_t k;
    const bool ok = sort_key_of(c, i, desc, &k);
    keep[i] = nulls ? !ok : (ok && k >= lo && k <= hi);
  }
}

__global__ void sort_gather_keys_kernel(DCol c, const uint64_t *ids, uint64_t m, int desc, uint64_t *keys) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
    uint64_t k;
    (void)sort_key_of(c, ids[i], desc, &k);
    keys[i] = k;
  }
}

}  // namespace qm

using namespace qm;

static unsigned sgrid(uint64_t n) {
  const uint64_t g = (n + 255) / 256;
  return (unsigned)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

extern "C" hipError_t qmk_sort_hist(const DCol *c, uint64_t n, int desc, uint32_t *hist, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(sort_hist_kernel, dim3(sgrid(n)), dim3(256), 0, st, *c, n, desc, hist);
  return hipGetLastError();
}

extern "C" hipError_t qmk_sort_range(const DCol *c, uint64_t n, int desc, uint64_t lo, uint64_t hi, int nulls,
                                     uint8_t *keep, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(sort_range_kernel, dim3(sgrid(n)), dim3(256), 0, st, *c, n, desc, lo, hi, nulls, keep);
  return hipGetLastError();
}

extern "C" hipError_t qmk_sort_gather_keys(const DCol *c, const uint64_t *ids, uint64_t m, int desc, uint64_t *keys,
                                           hipStream_t st) {
  if (m == 0) return hipSuccess;
  hipLaunchKernelGGL(sort_gather_keys_kernel, dim3(sgrid(m)), dim3(256), 0, st, *c, ids, m, desc, keys);
  return hipGetLastError();
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// scratch of the two radix sorts (k_radix.hip): 64-bit order codes, then 1-bit classes
static size_t radix_bytes(uint64_t n) {
  const size_t a = qmk_radix_tmp_bytes(n, 8), b = qmk_radix_tmp_bytes(n, 1);
  return a > b ? a : b;
}

extern "C" size_t qmk_sort_tmp_bytes(uint64_t n) {
  return 4 * align256(n * 8) + 2 * align256(n) + align256(radix_bytes(n));
}

extern "C" hipError_t qmk_sort_keys(const DCol *c, uint64_t n, int desc, uint64_t *rows_out, void *tmp,
                                    size_t tmp_bytes, hipStream_t st) {
  if (n == 0) return hipSuccess;
  if (tmp_bytes < qmk_sort_tmp_bytes(n)) return hipErrorInvalidValue;
  char *p = static_cast<char *>(tmp);
  uint64_t *k0 = (uint64_t *)p; p += align256(n * 8);
  uint64_t *k1 = (uint64_t *)p; p += align256(n * 8);
  uint64_t *v0 = (uint64_t *)p; p += align256(n * 8);
  uint64_t *v1 = (uint64_t *)p; p += align256(n * 8);
  uint8_t *c0 = (uint8_t *)p; p += align256(n);
  uint8_t *c1 = (uint8_t *)p; p += align256(n);
  const size_t rb = radix_bytes(n);
  void *rtmp = p;
  unsigned grid = (unsigned)((n + 255) / 256 > 65536 ? 65536 : (n + 255) / 256);
  hipLaunchKernelGGL(sort_keys_kernel, dim3(grid), dim3(256), 0, st, *c, n, desc, k0, v0, c0);
  hipError_t e = qmk_radix_sort_u64(k0, k1, v0, v1, n, 0, 64, rtmp, rb, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(gather_cls_kernel, dim3(grid), dim3(256), 0, st, c0, v1, n, c1);
  e = qmk_radix_sort_u8(c1, c0, v1, rows_out, n, 0, 1, rtmp, rb, st);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

// k_post.hip — SQL post-stages over a device result (src/sql/executor.ts:30-170,
// src/operators.ts:1177-1227): the HAVING predicate program, the order codes
// of a multi-column ORDER BY, and the first-row marks of DISTINCT. Results of
// these stages are selections (ascending row ids) and permutations that the
// host applies with the gather kernels.
#include <cmath>
#include "k_common.hip.h"
#include "qm_kernels.h"

namespace qm {

enum : uint8_t { TV_FALSE = 0, TV_TRUE = 1, TV_NULL = 2 };

// One value of a result column: integers exactly (the reference's bigint or
// safe-integer number), floats as doubles.
struct HVal {
  bool null, is_int;
  int64_t i;
  double f;
};

__device__ __forceinline__ HVal hload(const DCol &c, uint64_t r) {
  HVal v;
  v.null = load1_null(c, r);
  v.is_int = !dt_is_float(c.dtype);
  v.i = 0;
  v.f = 0;
  if (!v.null) {
    if (v.is_int) v.i = load1_i(c, r);
    else v.f = load1_f(c, r);
  }
  return v;
}

__device__ __forceinline__ HVal hlit(uint8_t kind, double f, int64_t i) {
  HVal v;
  v.null = kind == HL_NULL;
  v.is_int = kind == HL_I64;
  v.i = i;
  v.f = f;
  return v;
}

constexpr double H_TWO63 = 9223372036854775808.0;

// compare() of the SQL evaluator (src/sql/evaluator.ts:169-200) for two
// non-null numeric values; the sign of the result is what the comparison
// operators test (NaN sorts last; a NaN result, Infinity - Infinity, makes
// every ordering test false).
__device__ __forceinline__ double h_cmp_ff(double a, double b) {
  if (a != a && b != b) return 0;
  if (a != a) return 1;
  if (b != b) return -1;
  return a - b;
}
// bigint a against number b: b truncated, then its fraction decides a tie
__device__ __forceinline__ double h_cmp_if(int64_t a, double b) {
  if (b != b) return -1;
  if (b == INFINITY) return -1;
  if (b == -INFINITY) return 1;
  const double t = trunc(b);
  if (t >= H_TWO63) return -1;
  if (t < -H_TWO63) return 1;
  const int64_t bb = (int64_t)t;
  if (a < bb) return -1;
  if (a > bb) return 1;
  const double fr = b - t;
  return fr > 0 ? -1 : (fr < 0 ? 1 : 0);
}
// number a against bigint b
__device__ __forceinline__ double h_cmp_fi(double a, int64_t b) {
  if (a != a) return 1;
  if (a == INFINITY) return 1;
  if (a == -INFINITY) return -1;
  const double t = trunc(a);
  if (t >= H_TWO63) return 1;
  if (t < -H_TWO63) return -1;
  const int64_t ab = (int64_t)t;
  if (ab < b) return -1;
  if (ab > b) return 1;
  const double fr = a - t;
  return fr > 0 ? 1 : (fr < 0 ? -1 : 0);
}
__device__ __forceinline__ double h_cmp(const HVal &a, const HVal &b) {
  if (a.is_int && b.is_int) return a.i < b.i ? -1 : (a.i > b.i ? 1 : 0);
  if (a.is_int) return h_cmp_if(a.i, b.f);
  if (b.is_int) return h_cmp_fi(a.f, b.i);
  return h_cmp_ff(a.f, b.f);
}

__device__ __forceinline__ bool h_safe_int(double x) {
  return x == trunc(x) && fabs(x) <= 9007199254740991.0;
}
// looseEqual() (src/sql/evaluator.ts:157-167) for two non-null numbers
__device__ __forceinline__ bool h_eq(const HVal &a, const HVal &b) {
  if (a.is_int && b.is_int) return a.i == b.i;
  if (a.is_int) return h_safe_int(b.f) ? a.i == (int64_t)b.f : (double)a.i == b.f;
  if (b.is_int) return h_safe_int(a.f) ? (int64_t)a.f == b.i : a.f == (double)b.i;
  return a.f == b.f;
}

__device__ __forceinline__ uint8_t h_test(int cmp, const HVal &a, const HVal &b) {
  if (cmp == 0) return h_eq(a, b) ? TV_TRUE : TV_FALSE;
  if (cmp == 1) return h_eq(a, b) ? TV_FALSE : TV_TRUE;
  const double d = h_cmp(a, b);
  bool r;
  switch (cmp) {
    case 2: r = d < 0; break;
    case 3: r = d <= 0; break;
    case 4: r = d > 0; break;
    default: r = d >= 0; break;
  }
  return r ? TV_TRUE : TV_FALSE;
}

// One thread per result row: the HAVING program on a stack of three-valued
// results; the row is kept when the program's value is true.
__global__ void having_kernel(HProg prog, uint64_t n, uint8_t *keep) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
    uint8_t st[HP_MAX_DEPTH];
    int sp = 0;
    for (int k = 0; k < prog.nops; ++k) {
      const HPOp &o = prog.ops[k];
      uint8_t v = TV_NULL;
      switch (o.kind) {
        case HP_CMP: {
          const HVal a = hload(prog.cols[o.col], r);
          if (!a.null) v = h_test(o.cmp, a, hlit(o.lk, o.f, o.i));
          st[sp++] = v;
          break;
        }
        case HP_CMP_COL: {
          const HVal a = hload(prog.cols[o.col], r), b = hload(prog.cols[o.col2], r);
          if (!a.null && !b.null) v = h_test(o.cmp, a, b);
          st[sp++] = v;
          break;
        }
        case HP_BETWEEN: {
          const HVal a = hload(prog.cols[o.col], r);
          if (!a.null) {
            const bool in = h_cmp(a, hlit(o.lk, o.f, o.i)) >= 0 && h_cmp(a, hlit(o.lk2, o.f2, o.i2)) <= 0;
            v = (in != (o.neg != 0)) ? TV_TRUE : TV_FALSE;
          }
          st[sp++] = v;
          break;
        }
        case HP_IN: {
          const HVal a = hload(prog.cols[o.col], r);
          if (!a.null) {
            bool found = false, has_null = false;
            for (uint32_t j = 0; j < o.list_n && !found; ++j) {
              const uint32_t x = o.list_off + j;
              const uint8_t kind = prog.list_kind[x];
              if (kind == HL_NULL) { has_null = true; continue; }
              found = h_eq(a, hlit(kind, prog.list_f[x], prog.list_i[x]));
            }
            if (found) v = o.neg ? TV_FALSE : TV_TRUE;
            else if (!has_null) v = o.neg ? TV_TRUE : TV_FALSE;
          }
          st[sp++] = v;
          break;
        }
        case HP_ISNULL: {
          const bool isn = load1_null(prog.cols[o.col], r);
          st[sp++] = (isn != (o.neg != 0)) ? TV_TRUE : TV_FALSE;
          break;
        }
        case HP_AND: {
          const uint8_t b = st[--sp], a = st[--sp];
          st[sp++] = (a == TV_FALSE || b == TV_FALSE) ? TV_FALSE : ((a == TV_NULL || b == TV_NULL) ? TV_NULL : TV_TRUE);
          break;
        }
        case HP_OR: {
          const uint8_t b = st[--sp], a = st[--sp];
          st[sp++] = (a == TV_TRUE || b == TV_TRUE) ? TV_TRUE : ((a == TV_NULL || b == TV_NULL) ? TV_NULL : TV_FALSE);
          break;
        }
        default: {  // HP_NOT
          const uint8_t a = st[sp - 1];
          st[sp - 1] = a == TV_NULL ? TV_NULL : (a == TV_TRUE ? TV_FALSE : TV_TRUE);
          break;
        }
      }
    }
    keep[r] = (sp == 1 && st[0] == TV_TRUE) ? 1 : 0;
  }
}

// ---- selection of the kept rows, in ascending order ----
constexpr int MK_BLOCK = 256, MK_PER = 8, MK_ROWS = MK_BLOCK * MK_PER;

__device__ __forceinline__ uint32_t mk_count(const uint8_t *keep, uint64_t n, uint64_t base) {
  uint32_t c = 0;
#pragma unroll
  for (int k = 0; k < MK_PER; ++k) {
    const uint64_t r = base + k;
    c += (r < n && keep[r]) ? 1u : 0u;
  }
  return c;
}

__global__ void mask_count_kernel(const uint8_t *keep, uint64_t n, uint64_t *counts) {
  __shared__ uint32_t part[MK_BLOCK / 64];
  const uint64_t base = (uint64_t)blockIdx.x * MK_ROWS + (uint64_t)threadIdx.x * MK_PER;
  uint32_t c = mk_count(keep, n, base);
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int w = 0; w < MK_BLOCK / 64; ++w) t += part[w];
    counts[blockIdx.x] = t;
  }
}

__global__ void mask_write_kernel(const uint8_t *keep, uint64_t n, const uint64_t *offs, uint64_t *ids) {
  __shared__ uint32_t pre[MK_BLOCK];
  const uint64_t base = (uint64_t)blockIdx.x * MK_ROWS + (uint64_t)threadIdx.x * MK_PER;
  const uint32_t c = mk_count(keep, n, base);
  pre[threadIdx.x] = c;
  __syncthreads();
  for (int off = 1; off < MK_BLOCK; off <<= 1) {  // inclusive scan over the block
    const uint32_t add = threadIdx.x >= (unsigned)off ? pre[threadIdx.x - off] : 0u;
    __syncthreads();
    pre[threadIdx.x] += add;
    __syncthreads();
  }
  uint64_t w = offs[blockIdx.x] + pre[threadIdx.x] - c;
#pragma unroll
  for (int k = 0; k < MK_PER; ++k) {
    const uint64_t r = base + k;
    if (r < n && keep[r]) ids[w++] = r;
  }
}

// ---- order codes (the stable multi-column sort, src/sql/executor.ts:116-140) ----
// Canonical order code of a non-null value: integers by value; floats with
// -0 as +0 and every NaN above +Infinity (the evaluator's NaN-last compare;
// DISTINCT's String() equality makes all NaN one value and -0 equal to 0).
__device__ __forceinline__ uint64_t canon_code(const DCol &c, uint64_t r) {
  if (dt_is_float(c.dtype)) {
    double v = load1_f(c, r);
    if (v != v) return ~0ull;
    if (v == 0.0) v = 0.0;
    const uint64_t b = __double_as_longlong(v);
    return (b >> 63) ? ~b : (b | (1ull << 63));
  }
  return (uint64_t)load1_i(c, r) ^ (1ull << 63);
}

__global__ void iota_kernel(uint64_t *p, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = i;
}

// keys[i] for the row perm[i]: pass 0 the value code (descending: inverted),
// pass 1 the NULL flag (NULLs last in both directions).
__global__ void order_code_kernel(DCol c, const uint64_t *perm, uint64_t n, int desc, int null_pass,
                                  uint64_t *keys) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t r = perm[i];
    const bool isn = load1_null(c, r);
    if (null_pass) {
      keys[i] = isn ? 1 : 0;
    } else {
      const uint64_t code = isn ? 0 : canon_code(c, r);
      keys[i] = desc ? ~code : code;
    }
  }
}

// keep[i] = 1 when perm[i] starts a new DISTINCT key (the first row, or one
// whose key differs from the previous row's in any column; NULLs are one value)
__global__ void distinct_mark_kernel(const DCol *cols, int ncols, const uint64_t *perm, uint64_t n, uint8_t *keep) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint8_t k = 1;
    if (i > 0) {
      const uint64_t a = perm[i - 1], b = perm[i];
      k = 0;
      for (int j = 0; j < ncols && !k; ++j) {
        const DCol &c = cols[j];
        const bool na = load1_null(c, a), nb = load1_null(c, b);
        if (na != nb) k = 1;
        else if (!na && canon_code(c, a) != canon_code(c, b)) k = 1;
      }
    }
    keep[i] = k;
  }
}

__global__ void gather_u64_kernel(const uint64_t *src, const uint64_t *idx, uint64_t n, uint64_t *dst) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[idx[i]];
}

// ---- QMCB encoding of a row range (src/columnar.ts:523-613) ----
// dst[j] = the QMCB bits of rows [row0 + 8j, row0 + 8j + 8) within the range:
// a NULL bitmap shifted to start at row0, or a one-byte-per-row BOOL column
// packed; bits past the range are 0. `any` gets 1 when a bit is set.
__global__ void qmcb_bits_kernel(const uint8_t *src, int is_bool, uint64_t row0, uint64_t nrows, uint8_t *dst,
                                 uint32_t *any) {
  const uint64_t nb = (nrows + 7) / 8;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint32_t seen = 0;
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nb; j += stride) {
    uint32_t byte = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t k = j * 8 + b;
      if (k >= nrows) break;
      const uint64_t r = row0 + k;
      const uint32_t bit = is_bool ? (gld<uint8_t>(src, r) != 0) : ((gld<uint8_t>(src, r >> 3) >> (r & 7)) & 1u);
      byte |= bit << b;
    }
    dst[j] = (uint8_t)byte;
    seen |= byte;
  }
  if (any && seen) atomicOr(any, 1u);
}

}  // namespace qm

using namespace qm;

static unsigned post_grid(uint64_t n) {
  const uint64_t g = (n + 255) / 256;
  return (unsigned)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

extern "C" hipError_t qmk_having(const HProg *prog, uint64_t n, uint8_t *keep, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(having_kernel, dim3(post_grid(n)), dim3(256), 0, st, *prog, n, keep);
  return hipGetLastError();
}

extern "C" uint64_t qmk_mask_blocks(uint64_t n) { return (n + MK_ROWS - 1) / MK_ROWS; }

extern "C" hipError_t qmk_mask_count(const uint8_t *keep, uint64_t n, uint64_t *counts, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(mask_count_kernel, dim3((unsigned)qmk_mask_blocks(n)), dim3(MK_BLOCK), 0, st, keep, n, counts);
  return hipGetLastError();
}

extern "C" hipError_t qmk_mask_write(const uint8_t *keep, uint64_t n, const uint64_t *offs, uint64_t *ids,
                                     hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(mask_write_kernel, dim3((unsigned)qmk_mask_blocks(n)), dim3(MK_BLOCK), 0, st, keep, n, offs,
                     ids);
  return hipGetLastError();
}

extern "C" hipError_t qmk_iota(uint64_t *p, uint64_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(iota_kernel, dim3(post_grid(n)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

extern "C" hipError_t qmk_order_codes(const DCol *c, const uint64_t *perm, uint64_t n, int desc, int null_pass,
                                      uint64_t *keys, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(order_code_kernel, dim3(post_grid(n)), dim3(256), 0, st, *c, perm, n, desc, null_pass, keys);
  return hipGetLastError();
}

extern "C" hipError_t qmk_distinct_marks(const DCol *d_cols, int ncols, const uint64_t *perm, uint64_t n,
                                         uint8_t *keep, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(distinct_mark_kernel, dim3(post_grid(n)), dim3(256), 0, st, d_cols, ncols, perm, n, keep);
  return hipGetLastError();
}

extern "C" hipError_t qmk_gather_u64(const uint64_t *src, const uint64_t *idx, uint64_t n, uint64_t *dst,
                                     hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(gather_u64_kernel, dim3(post_grid(n)), dim3(256), 0, st, src, idx, n, dst);
  return hipGetLastError();
}

extern "C" size_t qmk_radix_pairs_tmp_bytes(uint64_t n) { return qmk_radix_tmp_bytes(n, 8); }

// Stable radix sort of (key, value) pairs on bits [0, end_bit) of the keys (k_radix.hip).
extern "C" hipError_t qmk_radix_pairs(const uint64_t *kin, uint64_t *kout, const uint64_t *vin, uint64_t *vout,
                                      uint64_t n, int end_bit, void *tmp, size_t tmp_bytes, hipStream_t st) {
  return qmk_radix_sort_u64(kin, kout, vin, vout, n, 0, end_bit, tmp, tmp_bytes, st);
}

extern "C" hipError_t qmk_qmcb_bits(const uint8_t *src, int is_bool, uint64_t row0, uint64_t nrows, uint8_t *dst,
                                    uint32_t *any, hipStream_t st) {
  if (nrows == 0) return hipSuccess;
  hipLaunchKernelGGL(qmcb_bits_kernel, dim3(post_grid((nrows + 7) / 8)), dim3(256), 0, st, src, is_bool, row0, nrows,
                     dst, any);
  return hipGetLastError();
}

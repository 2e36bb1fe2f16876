// k_stats.hip — per-column statistics computed once at upload: min/max
// (the zone-map values the reference keeps per page, src/decode.ts:9-100),
// the largest finite |x| (fixed-point scale of exact FP64 sums), NaN and
// NULL counts, and a 4096-register HyperLogLog sketch of the distinct
// values (hash-table sizing for GROUP BY and joins).
#include "k_common.hip.h"
#include "qm_kernels.h"

namespace qm {

// stats words: [0] max(~code) (MIN), [1] max(code) (MAX), [2] max finite |x| bits,
// [3] NaN count, [4] NULL count, [5] non-null count, [6] descents of an integer
// column without NULLs (rows i with x[i] > x[i+1]; 0 = sorted ascending), [7]
// max(~bits) of the nonzero finite |x| (the smallest one is ~stats[7]; 0 = none)
__global__ void col_stats_kernel(DCol c, uint64_t n, uint64_t *stats, uint32_t *hll) {
  __shared__ uint32_t sh_hll[4096];
  for (int i = threadIdx.x; i < 4096; i += blockDim.x) sh_hll[i] = 0;
  __syncthreads();
  uint64_t mn = 0, mx = 0, mabs = 0, mnz = 0, nan = 0, nul = 0, nn = 0, desc = 0;
  const bool isf = dt_is_float(c.dtype);
  const bool order_check = !isf && c.nulls == nullptr;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    if (load1_null(c, i)) { ++nul; continue; }
    ++nn;
    uint64_t keybits;
    if (isf) {
      double v = load1_f(c, i);
      uint64_t b = f64_bits(v);
      keybits = canon_f64_key_bits(b);
      if (v != v) { ++nan; }
      else {
        uint64_t code = code_f64_bits(b);
        if (~code > mn) mn = ~code;
        if (code > mx) mx = code;
        if (!f64_is_inf_bits(b)) {
          uint64_t a = b & 0x7FFFFFFFFFFFFFFFull;
          if (a > mabs) mabs = a;
          if (a && ~a > mnz) mnz = ~a;
        }
      }
    } else {
      int64_t v = load1_i(c, i);
      keybits = (uint64_t)v;
      if (order_check && i + 1 < n && load1_i(c, i + 1) < v) ++desc;
      uint64_t code = code_i64(v);
      if (~code > mn) mn = ~code;
      if (code > mx) mx = code;
    }
    uint64_t h = mix64(keybits ^ 0x5bd1e9955bd1e995ull);
    uint32_t reg = (uint32_t)(h >> 52);
    uint64_t w = (h << 12) | (1ull << 11);
    uint32_t rank = (uint32_t)__clzll((long long)w) + 1;
    atomicMax(&sh_hll[reg], rank);
  }
  for (int off = 32; off > 0; off >>= 1) {
    uint64_t t;
    t = __shfl_xor(mn, off); if (t > mn) mn = t;
    t = __shfl_xor(mx, off); if (t > mx) mx = t;
    t = __shfl_xor(mabs, off); if (t > mabs) mabs = t;
    t = __shfl_xor(mnz, off); if (t > mnz) mnz = t;
    nan += __shfl_xor(nan, off);
    nul += __shfl_xor(nul, off);
    nn += __shfl_xor(nn, off);
    desc += __shfl_xor(desc, off);
  }
  if ((threadIdx.x & 63) == 0) {
    if (mn) g_atomic_max(&stats[0], mn);
    if (mx) g_atomic_max(&stats[1], mx);
    if (mabs) g_atomic_max(&stats[2], mabs);
    if (mnz) g_atomic_max(&stats[7], mnz);
    if (nan) g_atomic_add(&stats[3], nan);
    if (nul) g_atomic_add(&stats[4], nul);
    if (nn) g_atomic_add(&stats[5], nn);
    if (desc) g_atomic_add(&stats[6], desc);
  }
  __syncthreads();
  for (int r = threadIdx.x; r < 4096; r += blockDim.x)
    if (sh_hll[r]) atomicMax(&hll[r], sh_hll[r]);
}

// key histogram: rows per bin of 2^g consecutive key values (bin = (v - kmin) >> g);
// up to KH_LDS_BINS bins are counted in LDS per workgroup first
static constexpr uint32_t KH_LDS_BINS = 16384;
template <bool LDS>
__global__ void __launch_bounds__(256) key_hist_kernel(DCol c, uint64_t n, int64_t kmin, int g, uint32_t nbins,
                                                       uint32_t *hist) {
  __shared__ uint32_t s_h[LDS ? KH_LDS_BINS : 1];
  if constexpr (LDS) {
    for (uint32_t i = threadIdx.x; i < nbins; i += 256) s_h[i] = 0;
    __syncthreads();
  }
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x; r < n; r += stride) {
    const uint64_t b = (uint64_t)(load1_i(c, r) - kmin) >> g;
    if constexpr (LDS) atomicAdd(&s_h[b], 1u);
    else atomicAdd(&hist[b], 1u);
  }
  if constexpr (LDS) {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nbins; i += 256)
      if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
  }
}

// streaming read: every 16-byte word once (eight loads in flight per
// thread), summed so the loads are not dropped
__global__ void __launch_bounds__(256) stream_read_kernel(const uint4 *p, uint64_t n16, uint64_t *sink) {
  uint32_t acc = 0;
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  for (; i + 7 * stride < n16; i += 8 * stride) {
    gv4u v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = gld<gv4u>(p, (i + k * stride) * 16);
#pragma unroll
    for (int k = 0; k < 8; ++k) acc ^= v[k].x ^ v[k].y ^ v[k].z ^ v[k].w;
  }
  for (; i < n16; i += stride) {
    const gv4u a = gld<gv4u>(p, i * 16);
    acc ^= a.x ^ a.y ^ a.z ^ a.w;
  }
  if (acc == 0x9E3779B9u) sink[0] = acc;  // practically never: keeps the loads
}

}  // namespace qm

using namespace qm;

extern "C" hipError_t qmk_col_stats(const DCol *c, uint64_t n, uint64_t *stats, uint32_t *hll, hipStream_t st) {
  if (n == 0) return hipSuccess;
  uint64_t g = (n + 255) / 256;
  if (g > 1024) g = 1024;
  hipLaunchKernelGGL(col_stats_kernel, dim3((unsigned)g), dim3(256), 0, st, *c, n, stats, hll);
  return hipGetLastError();
}

namespace qm {

// Zone maps: per page of ZONE_ROWS rows, the min and max order code of its
// non-NULL values (a page of NULLs only: min ~0, max 0), so a range filter
// can skip pages it cannot match (the reference's canSkipPage on page
// min/max, src/decode.ts:9-100). One workgroup per page.
__global__ void __launch_bounds__(256) zone_kernel(DCol c, uint64_t n, uint64_t *zmin, uint64_t *zmax) {
  __shared__ uint64_t s_mn[4], s_mx[4];
  const uint64_t p = blockIdx.x;
  const uint64_t r0 = p * ZONE_ROWS, r1 = (r0 + ZONE_ROWS < n) ? r0 + ZONE_ROWS : n;
  const bool isf = dt_is_float(c.dtype);
  uint64_t mn = ~0ull, mx = 0;
  for (uint64_t r = r0 + threadIdx.x; r < r1; r += 256) {
    if (load1_null(c, r)) continue;
    const uint64_t code = isf ? code_f64_bits(f64_bits(load1_f(c, r))) : code_i64(load1_i(c, r));
    mn = code < mn ? code : mn;
    mx = code > mx ? code : mx;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t a = __shfl_xor(mn, o, 64), b = __shfl_xor(mx, o, 64);
    mn = a < mn ? a : mn;
    mx = b > mx ? b : mx;
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { s_mn[w] = mn; s_mx[w] = mx; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; ++k) {
      mn = s_mn[k] < mn ? s_mn[k] : mn;
      mx = s_mx[k] > mx ? s_mx[k] : mx;
    }
    zmin[p] = mn;
    zmax[p] = mx;
  }
}

}  // namespace qm

using namespace qm;

extern "C" hipError_t qmk_zone_maps(const DCol *c, uint64_t n, uint64_t *zmin, uint64_t *zmax, hipStream_t st) {
  if (n == 0) return hipSuccess;
  const uint64_t pages = (n + ZONE_ROWS - 1) / ZONE_ROWS;
  hipLaunchKernelGGL(zone_kernel, dim3((unsigned)pages), dim3(256), 0, st, *c, n, zmin, zmax);
  return hipGetLastError();
}

extern "C" hipError_t qmk_key_hist(const DCol *c, uint64_t n, int64_t kmin, int g, uint32_t nbins, uint32_t *hist,
                                   hipStream_t st) {
  if (n == 0) return hipSuccess;
  uint64_t grid = (n + 255) / 256;
  if (nbins <= KH_LDS_BINS) {
    if (grid > 2048) grid = 2048;
    hipLaunchKernelGGL(key_hist_kernel<true>, dim3((unsigned)grid), dim3(256), 0, st, *c, n, kmin, g, nbins, hist);
  } else {
    if (grid > 65536) grid = 65536;
    hipLaunchKernelGGL(key_hist_kernel<false>, dim3((unsigned)grid), dim3(256), 0, st, *c, n, kmin, g, nbins, hist);
  }
  return hipGetLastError();
}

extern "C" hipError_t qmk_stream_read(const void *p, uint64_t bytes, uint32_t grid, uint64_t *sink, hipStream_t st) {
  hipLaunchKernelGGL(stream_read_kernel, dim3(grid), dim3(256), 0, st, static_cast<const uint4 *>(p), bytes / 16, sink);
  return hipGetLastError();
}

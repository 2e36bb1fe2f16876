// k_radix.hip — stable LSD radix sort of (key, 64-bit value) pairs for gfx950.
//
// Used where the query needs a total order rather than the k best: ORDER BY
// without LIMIT (src/operators.ts:1751-1802 InMemorySort, here by the order
// codes of k_sort.hip), the multi-column ORDER BY and DISTINCT stages
// (k_post.hip) and the key-range passes of an ORDER BY over the memory
// budget. Stability is what those callers rely on: equal keys keep their
// input order, so a sort by one key after another is a sort by both, and a
// tie keeps row order as the reference's comparator does.
//
// One pass per digit of up to 8 bits, least significant first, each in three
// launches over tiles of RX_TILE keys:
//   1. radix_hist_kernel: the tile's digit counts (LDS), stored digit-major;
//   2. an exclusive scan of those counts (qmk_exscan_u64): where each
//      (digit, tile) run starts in the output;
//   3. radix_scatter_kernel: the tile again, in rounds of one key per thread
//      in input order; a key's rank among the equal digits of its wave comes
//      from one ballot per digit bit, the counts of the earlier waves and
//      rounds from a per-digit prefix in LDS; key and value go to run start +
//      rank.
// Passes alternate between the output and one temporary pair of buffers so
// the last pass lands in the output. Nothing is allocated: the caller passes
// qmk_radix_tmp_bytes(n, key bytes) of scratch.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "k_common.hip.h"
#include "qm_kernels.h"

namespace qm {
namespace {

constexpr int RX_BLOCK = 256;                      // threads per tile (4 waves)
constexpr int RX_ROUNDS = 16;                      // keys per thread per tile
constexpr uint32_t RX_TILE = RX_BLOCK * RX_ROUNDS;  // keys per tile
constexpr int RX_WAVES = RX_BLOCK / 64;

template <class K>
__device__ __forceinline__ uint32_t rx_digit(K key, int shift, uint32_t dmask) {
  return (uint32_t)((uint64_t)key >> shift) & dmask;
}

// 1. digit counts of tile b: counts[d * ntiles + b]
template <class K>
__global__ void __launch_bounds__(RX_BLOCK) radix_hist_kernel(const K *__restrict__ keys, uint64_t n, int shift,
                                                             int bits, uint64_t *__restrict__ counts,
                                                             uint32_t ntiles) {
  __shared__ uint32_t s_cnt[256];
  const int t = threadIdx.x;
  const uint32_t ndig = 1u << bits, dmask = ndig - 1u;
  s_cnt[t] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * RX_TILE;
#pragma unroll 4
  for (int r = 0; r < RX_ROUNDS; ++r) {
    const uint64_t i = base + (uint64_t)r * RX_BLOCK + (uint64_t)t;
    if (i < n) atomicAdd(&s_cnt[rx_digit(keys[i], shift, dmask)], 1u);
  }
  __syncthreads();
  if ((uint32_t)t < ndig) *gptr(&counts[(uint64_t)t * ntiles + blockIdx.x]) = s_cnt[t];
}

// 3. stable scatter of tile b by its digit runs' starts (offsets, digit-major)
template <class K>
__global__ void __launch_bounds__(RX_BLOCK) radix_scatter_kernel(const K *__restrict__ kin, K *__restrict__ kout,
                                                                const uint64_t *__restrict__ vin,
                                                                uint64_t *__restrict__ vout, uint64_t n, int shift,
                                                                int bits, const uint64_t *__restrict__ offsets,
                                                                uint32_t ntiles) {
  __shared__ uint64_t s_off[256];                 // output start of each digit's run of this tile
  __shared__ uint32_t s_wc[2][RX_WAVES][256];     // per round: keys of each (wave, digit), then their first rank
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const uint32_t ndig = 1u << bits, dmask = ndig - 1u;
  if ((uint32_t)t < ndig) s_off[t] = offsets[(uint64_t)t * ntiles + blockIdx.x];
#pragma unroll
  for (int w = 0; w < RX_WAVES; ++w) {
    s_wc[0][w][t] = 0;
    s_wc[1][w][t] = 0;
  }
  uint32_t run = 0;  // thread t: keys of digit t in the earlier rounds
  const uint64_t lt = (1ull << lane) - 1ull;
  const uint64_t base = (uint64_t)blockIdx.x * RX_TILE;
  __syncthreads();
  for (int r = 0; r < RX_ROUNDS; ++r) {
    const int buf = r & 1;
    const uint64_t i = base + (uint64_t)r * RX_BLOCK + (uint64_t)t;
    const bool ok = i < n;
    K k = 0;
    uint64_t v = 0;
    uint32_t d = 0;
    if (ok) {
      k = kin[i];
      v = vin[i];
      d = rx_digit(k, shift, dmask);
    }
    // lanes of this wave holding the same digit: one ballot per digit bit
    uint64_t same = __ballot(ok);
    for (int b = 0; b < bits; ++b) {
      const uint64_t on = __ballot(ok && ((d >> b) & 1u));
      same &= ((d >> b) & 1u) ? on : ~on;
    }
    const uint32_t below = (uint32_t)__builtin_popcountll(same & lt);
    if (ok && below == 0) s_wc[buf][wave][d] = (uint32_t)__builtin_popcountll(same);
    __syncthreads();
    // thread t (digit t): first rank of each wave's keys of digit t in this
    // round; clears the other buffer for the next round
    if ((uint32_t)t < ndig) {
      uint32_t acc = run;
#pragma unroll
      for (int w = 0; w < RX_WAVES; ++w) {
        const uint32_t c = s_wc[buf][w][t];
        s_wc[buf][w][t] = acc;
        acc += c;
        s_wc[buf ^ 1][w][t] = 0;
      }
      run = acc;
    }
    __syncthreads();
    if (ok) {
      const uint64_t pos = s_off[d] + s_wc[buf][wave][d] + below;
      *gptr(&kout[pos]) = k;
      *gptr(&vout[pos]) = v;
    }
  }
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

uint32_t rx_tiles(uint64_t n) { return (uint32_t)((n + RX_TILE - 1) / RX_TILE); }

size_t rx_tmp_bytes(uint64_t n, size_t key_bytes) {
  const uint64_t ncnt = (uint64_t)256 * rx_tiles(n);
  return align256(n * key_bytes) + align256(n * 8) + align256(ncnt * 8) + align256(qmk_exscan_tmp_words(ncnt) * 8) +
         256;
}

// Stable sort of n (key, value) pairs on key bits [begin_bit, end_bit).
template <class K>
hipError_t rx_sort(const K *kin, K *kout, const uint64_t *vin, uint64_t *vout, uint64_t n, int begin_bit,
                   int end_bit, void *tmp, size_t tmp_bytes, hipStream_t st) {
  if (n == 0) return hipSuccess;
  if (!kin || !kout || !vin || !vout || kin == kout || vin == vout || begin_bit < 0 ||
      end_bit > (int)(8 * sizeof(K)) || tmp_bytes < rx_tmp_bytes(n, sizeof(K)))
    return hipErrorInvalidValue;
  if (end_bit <= begin_bit) {  // no key bits: the input order is the order
    if (hipMemcpyAsync(kout, kin, n * sizeof(K), hipMemcpyDeviceToDevice, st) != hipSuccess) return hipGetLastError();
    return hipMemcpyAsync(vout, vin, n * 8, hipMemcpyDeviceToDevice, st);
  }
  const uint32_t ntiles = rx_tiles(n);
  const uint64_t ncnt = (uint64_t)256 * ntiles;
  char *p = static_cast<char *>(tmp);
  K *kt = reinterpret_cast<K *>(p);
  p += align256(n * sizeof(K));
  uint64_t *vt = reinterpret_cast<uint64_t *>(p);
  p += align256(n * 8);
  uint64_t *counts = reinterpret_cast<uint64_t *>(p);
  p += align256(ncnt * 8);
  uint64_t *stmp = reinterpret_cast<uint64_t *>(p);
  p += align256(qmk_exscan_tmp_words(ncnt) * 8);
  uint64_t *total = reinterpret_cast<uint64_t *>(p);
  const int npass = (end_bit - begin_bit + 7) / 8;
  const K *ks = kin;
  const uint64_t *vs = vin;
  for (int pass = 0; pass < npass; ++pass) {
    const int shift = begin_bit + 8 * pass;
    const int bits = end_bit - shift < 8 ? end_bit - shift : 8;
    // the last pass writes the output; the ones before alternate so that it does
    const bool to_out = ((npass - 1 - pass) & 1) == 0;
    K *kd = to_out ? kout : kt;
    uint64_t *vd = to_out ? vout : vt;
    const uint32_t ndig = 1u << bits;
    const uint64_t nc = (uint64_t)ndig * ntiles;
    hipLaunchKernelGGL(radix_hist_kernel<K>, dim3(ntiles), dim3(RX_BLOCK), 0, st, ks, n, shift, bits, counts, ntiles);
    hipError_t e = qmk_exscan_u64(counts, nc, counts, total, stmp, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(radix_scatter_kernel<K>, dim3(ntiles), dim3(RX_BLOCK), 0, st, ks, kd, vs, vd, n, shift, bits,
                       counts, ntiles);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    ks = kd;
    vs = vd;
  }
  return hipSuccess;
}

}  // namespace
}  // namespace qm

using namespace qm;

extern "C" size_t qmk_radix_tmp_bytes(uint64_t n, int key_bytes) {
  return rx_tmp_bytes(n, key_bytes == 1 ? 1 : 8);
}

extern "C" hipError_t qmk_radix_sort_u64(const uint64_t *kin, uint64_t *kout, const uint64_t *vin, uint64_t *vout,
                                         uint64_t n, int begin_bit, int end_bit, void *tmp, size_t tmp_bytes,
                                         hipStream_t st) {
  return rx_sort<uint64_t>(kin, kout, vin, vout, n, begin_bit, end_bit, tmp, tmp_bytes, st);
}

extern "C" hipError_t qmk_radix_sort_u8(const uint8_t *kin, uint8_t *kout, const uint64_t *vin, uint64_t *vout,
                                        uint64_t n, int begin_bit, int end_bit, void *tmp, size_t tmp_bytes,
                                        hipStream_t st) {
  return rx_sort<uint8_t>(kin, kout, vin, vout, n, begin_bit, end_bit, tmp, tmp_bytes, st);
}

// k_join.hip — hash-join build of the right (build) side and the
// materialised inner-join probe.
//
// Replaces HashJoinOperator's in-memory path (src/operators.ts:2349-2436
// build, 2673-2713 probe). The reference maps String(key) -> Row[] in
// insertion order and skips NULL keys; here the build produces a bucketed
// index: per distinct key a (start, count) range into an array of build row
// ids sorted ascending (= insertion order), either directly indexed by
// key - min for dense key domains or in an open-addressing table.
#include "k_common.hip.h"
#include "qm_kernels.h"

namespace qm {

__device__ __forceinline__ int64_t join_key_of(const DCol &c, uint64_t row, int key_float) {
  return key_float ? (int64_t)canon_f64_key_bits(f64_bits(load1_f(c, row))) : load1_i(c, row);
}

// table entry word holding (start << 32 | count) for a key; inserts in HASH mode
__device__ __forceinline__ uint64_t *join_slot(uint64_t *table, int mode, int64_t key, int64_t kmin, uint64_t cap,
                                               uint32_t *status) {
  if (mode == 1) return table + (uint64_t)(key - kmin);
  uint64_t stored = (uint64_t)key + KEY_SALT;
  if (stored == 0) return table + 2 * cap + 1;
  const uint64_t mask = cap - 1;
  uint64_t h = mix64(stored) & mask;
  for (uint64_t p = 0; p <= mask; ++p) {
    uint64_t slot = (h + p) & mask;
    uint64_t *e = table + 2 * slot;
    uint64_t k = g_load(e);
    if (k == stored) return e + 1;
    if (k == 0) {
      uint64_t old = g_atomic_cas(e, 0, stored);
      if (old == 0 || old == stored) return e + 1;
    }
  }
  status[0] = 1;
  return nullptr;
}

__global__ void join_count_build_kernel(DCol c, uint64_t n, const uint64_t *sel, int mode, int key_float,
                                        int64_t kmin, uint64_t cap, uint64_t *table, uint32_t *status) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; j < n; j += stride) {
    uint64_t i = sel ? sel[j] : j;
    if (load1_null(c, i)) continue;  // NULL never matches (src/operators.ts:2424-2428)
    uint64_t *w = join_slot(table, mode, join_key_of(c, i, key_float), kmin, cap, status);
    if (w) g_atomic_add(w, 1);
  }
}

// counts (low 32 bits of each entry word) -> u64 array
__global__ void join_extract_counts_kernel(const uint64_t *table, int mode, uint64_t nent, uint64_t *counts) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < nent; i += stride) {
    uint64_t w = (mode == 1) ? table[i] : table[2 * i + 1];
    counts[i] = w & 0xFFFFFFFFull;
  }
}

__global__ void join_store_starts_kernel(uint64_t *table, int mode, uint64_t nent, const uint64_t *starts) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < nent; i += stride) {
    uint64_t *w = (mode == 1) ? &table[i] : &table[2 * i + 1];
    *w = (starts[i] << 32) | (*w & 0xFFFFFFFFull);
  }
}

__global__ void join_scatter_kernel(DCol c, uint64_t n, const uint64_t *sel, int mode, int key_float, int64_t kmin,
                                    uint64_t cap, uint64_t *table, uint32_t *rows, uint32_t *cursor,
                                    uint32_t *status) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; j < n; j += stride) {
    uint64_t i = sel ? sel[j] : j;
    if (load1_null(c, i)) continue;
    uint64_t *w = join_slot(table, mode, join_key_of(c, i, key_float), kmin, cap, status);
    if (!w) continue;
    uint64_t ent = (mode == 1) ? (uint64_t)(w - table) : (uint64_t)((w - table) - 1) / 2;
    uint32_t pos = atomicAdd(&cursor[ent], 1u);
    rows[(uint32_t)(*w >> 32) + pos] = (uint32_t)i;
  }
}

// Each key's row-id range is sorted ascending (the reference's insertion
// order). A range of at most JOIN_SMALL_SEG rows is insertion-sorted by one
// thread; a longer one is appended to `big` and sorted by a workgroup below.
constexpr uint32_t JOIN_SMALL_SEG = 32;

__global__ void join_sort_segments_kernel(const uint64_t *table, int mode, uint64_t nent, uint32_t *rows,
                                          uint32_t *big, uint32_t *nbig) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < nent; i += stride) {
    uint64_t w = (mode == 1) ? table[i] : table[2 * i + 1];
    uint32_t cnt = (uint32_t)w, st = (uint32_t)(w >> 32);
    if (cnt < 2) continue;
    if (cnt > JOIN_SMALL_SEG) {
      big[atomicAdd(nbig, 1u)] = (uint32_t)i;  // at most nrows / (JOIN_SMALL_SEG + 1) such ranges
      continue;
    }
    uint32_t *s = rows + st;
    for (uint32_t a = 1; a < cnt; ++a) {
      uint32_t x = s[a];
      int b = (int)a - 1;
      while (b >= 0 && s[b] > x) { s[b + 1] = s[b]; --b; }
      s[b + 1] = x;
    }
  }
}

// A long range is sorted by marking its row ids (all distinct) in an LDS
// bitmap one window of JOIN_BM_BITS values at a time and writing the set bits
// out in order: O(cnt * windows + range / 32) per range, with windows =
// (max - min) / JOIN_BM_BITS + 1. Output goes to `out` at the range's offset,
// then back into `rows`.
constexpr uint32_t JOIN_BIG_BLOCK = 512;
constexpr uint32_t JOIN_BM_WORDS = 8192;  // 32 KB of LDS
constexpr uint64_t JOIN_BM_BITS = (uint64_t)JOIN_BM_WORDS * 32;
constexpr uint32_t JOIN_BM_WPT = JOIN_BM_WORDS / JOIN_BIG_BLOCK;

__device__ __forceinline__ uint32_t block_exscan_u32(uint32_t v, uint32_t *total, uint32_t *sh /*>=9*/) {
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(x, off, 64);
    if (lane >= (uint32_t)off) x += y;
  }
  if (lane == 63) sh[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (uint32_t w = 0; w < JOIN_BIG_BLOCK / 64; ++w) { uint32_t t = sh[w]; sh[w] = acc; acc += t; }
    sh[JOIN_BIG_BLOCK / 64] = acc;
  }
  __syncthreads();
  const uint32_t res = sh[wid] + x - v;
  *total = sh[JOIN_BIG_BLOCK / 64];
  __syncthreads();
  return res;
}

__global__ void __launch_bounds__(JOIN_BIG_BLOCK) join_sort_big_kernel(const uint64_t *table, int mode,
                                                                       uint32_t *rows, const uint32_t *big,
                                                                       const uint32_t *nbig, uint32_t *out) {
  __shared__ uint32_t bm[JOIN_BM_WORDS];
  __shared__ uint32_t sh[JOIN_BIG_BLOCK / 64 + 1];
  __shared__ uint32_t smin[JOIN_BIG_BLOCK / 64], smax[JOIN_BIG_BLOCK / 64];
  const uint32_t tid = threadIdx.x, nb = *nbig;
  for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
    const uint64_t ent = big[b];
    const uint64_t w = (mode == 1) ? table[ent] : table[2 * ent + 1];
    const uint32_t cnt = (uint32_t)w;
    uint32_t *s = rows + (uint32_t)(w >> 32);
    uint32_t *o = out + (uint32_t)(w >> 32);
    uint32_t lo = 0xFFFFFFFFu, hi = 0;
    for (uint32_t e = tid; e < cnt; e += JOIN_BIG_BLOCK) {
      const uint32_t v = s[e];
      lo = min(lo, v);
      hi = max(hi, v);
    }
    for (int off = 32; off > 0; off >>= 1) {
      lo = min(lo, (uint32_t)__shfl_xor(lo, off));
      hi = max(hi, (uint32_t)__shfl_xor(hi, off));
    }
    if ((tid & 63) == 0) { smin[tid >> 6] = lo; smax[tid >> 6] = hi; }
    __syncthreads();
    lo = smin[0];
    hi = smax[0];
    for (uint32_t k = 1; k < JOIN_BIG_BLOCK / 64; ++k) { lo = min(lo, smin[k]); hi = max(hi, smax[k]); }
    uint32_t written = 0;
    for (uint64_t base = lo; base <= hi; base += JOIN_BM_BITS) {
#pragma unroll
      for (uint32_t k = 0; k < JOIN_BM_WPT; ++k) bm[tid * JOIN_BM_WPT + k] = 0;
      __syncthreads();
      for (uint32_t e = tid; e < cnt; e += JOIN_BIG_BLOCK) {
        const uint64_t v = s[e];
        if (v >= base && v - base < JOIN_BM_BITS) {
          const uint32_t d = (uint32_t)(v - base);
          atomicOr(&bm[d >> 5], 1u << (d & 31));
        }
      }
      __syncthreads();
      uint32_t words[JOIN_BM_WPT], c = 0;
#pragma unroll
      for (uint32_t k = 0; k < JOIN_BM_WPT; ++k) {
        words[k] = bm[tid * JOIN_BM_WPT + k];
        c += __popc(words[k]);
      }
      uint32_t total;
      uint32_t pos = written + block_exscan_u32(c, &total, sh);
#pragma unroll
      for (uint32_t k = 0; k < JOIN_BM_WPT; ++k) {
        uint32_t m = words[k];
        const uint64_t v0 = base + (uint64_t)(tid * JOIN_BM_WPT + k) * 32;
        while (m) {
          const uint32_t bit = __ffs(m) - 1;
          m &= m - 1;
          o[pos++] = (uint32_t)(v0 + bit);
        }
      }
      written += total;
    }
    __syncthreads();  // the range's sorted ids are in `out`
    for (uint32_t e = tid; e < cnt; e += JOIN_BIG_BLOCK) s[e] = o[e];
    __syncthreads();
  }
}

// ---- materialised probe: pairs in probe order, build rows ascending ----
__device__ __forceinline__ uint32_t probe_matches8(const DScan &sc, const DQuery &q, uint64_t row0, uint32_t valid,
                                                   uint32_t st[RPT], uint32_t ct[RPT]) {
  uint32_t m = eval_query8(q, sc.cols, row0, valid);
  const DCol &kc = sc.cols[sc.join_col];
  uint32_t total = 0;
#pragma unroll
  for (int r = 0; r < RPT; ++r) { st[r] = 0; ct[r] = 0; }
  if (!m) return 0;
  m &= ~load8_nulls(kc, row0);
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    if (!((m >> r) & 1)) continue;
    int64_t key = join_key_of(kc, row0 + r, sc.join_key_float);
    uint32_t s, c;
    if (join_lookup(sc, key, &s, &c)) { st[r] = s; ct[r] = c; total += c; }
  }
  return total;
}

__global__ void __launch_bounds__(256) join_pairs_count_kernel(DScan sc, const DQuery *__restrict__ qs,
                                                               uint64_t tpb, uint64_t *block_counts) {
  const DQuery &q = qs[0];
  const uint64_t tile_rows = (uint64_t)blockDim.x * RPT;
  uint64_t t0 = (uint64_t)blockIdx.x * tpb;
  uint64_t cnt = 0;
  for (uint64_t t = t0; t < t0 + tpb; ++t) {
    uint64_t row0 = t * tile_rows + (uint64_t)threadIdx.x * RPT;
    if (row0 >= sc.nrows) break;
    uint64_t left = sc.nrows - row0;
    uint32_t valid = left >= RPT ? 0xFFu : ((1u << left) - 1);
    uint32_t st[RPT], ct[RPT];
    cnt += probe_matches8(sc, q, row0, valid, st, ct);
  }
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
  __shared__ uint64_t sh[4];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) block_counts[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

__device__ __forceinline__ uint64_t block_exscan_u64(uint64_t v, uint64_t *total, uint64_t *sh /*>=17*/) {
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint64_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint64_t y = __shfl_up(x, off, 64);
    if (lane >= (uint32_t)off) x += y;
  }
  if (lane == 63) sh[wid] = x;
  __syncthreads();
  const uint32_t nw = (blockDim.x + 63) >> 6;
  if (threadIdx.x == 0) {
    uint64_t acc = 0;
    for (uint32_t w = 0; w < nw; ++w) { uint64_t t = sh[w]; sh[w] = acc; acc += t; }
    sh[16] = acc;
  }
  __syncthreads();
  uint64_t res = sh[wid] + x - v;
  *total = sh[16];
  __syncthreads();
  return res;
}

__global__ void __launch_bounds__(256) join_pairs_write_kernel(DScan sc, const DQuery *__restrict__ qs, uint64_t tpb,
                                                               const uint64_t *block_offsets, uint64_t out_cap,
                                                               uint64_t *out_left, uint64_t *out_right) {
  __shared__ uint64_t sh[17];
  const DQuery &q = qs[0];
  const uint64_t tile_rows = (uint64_t)blockDim.x * RPT;
  uint64_t t0 = (uint64_t)blockIdx.x * tpb;
  uint64_t out_base = block_offsets[blockIdx.x];
  for (uint64_t t = t0; t < t0 + tpb; ++t) {
    if (t * tile_rows >= sc.nrows) break;
    if (out_base >= out_cap) break;
    uint64_t row0 = t * tile_rows + (uint64_t)threadIdx.x * RPT;
    uint32_t st[RPT], ct[RPT];
    uint64_t mine = 0;
    if (row0 < sc.nrows) {
      uint64_t left = sc.nrows - row0;
      uint32_t valid = left >= RPT ? 0xFFu : ((1u << left) - 1);
      mine = probe_matches8(sc, q, row0, valid, st, ct);
    } else {
      for (int r = 0; r < RPT; ++r) { st[r] = 0; ct[r] = 0; }
    }
    uint64_t total;
    uint64_t pos = out_base + block_exscan_u64(mine, &total, sh);
    for (int r = 0; r < RPT; ++r) {
      for (uint32_t j = 0; j < ct[r]; ++j) {
        if (pos < out_cap) { out_left[pos] = row0 + r; out_right[pos] = sc.join_rows[st[r] + j]; }
        ++pos;
      }
    }
    out_base += total;
  }
}

}  // namespace qm

using namespace qm;

static unsigned gridn(uint64_t n) {
  uint64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

extern "C" uint64_t qmk_join_build_tmp_words(uint64_t nent, uint64_t nrows) {
  const uint64_t scan = 2 * nent + qmk_exscan_tmp_words(nent);
  const uint64_t sort = (nrows + nrows / (JOIN_SMALL_SEG + 1) + 2) / 2 + 1;
  return std::max(scan, sort);
}

// tmp must hold qmk_join_build_tmp_words(nent, nrows) u64 words, cursor nent + 1 u32 (zero)
// DIRECT index with unique keys: per slot the build row + 1 (0: none);
// status[1] <- 1 when a key has several build rows (the index is then unused)
__global__ void join_unique_kernel(const uint64_t *table, uint64_t nent, const uint32_t *rows, uint32_t *uniq,
                                   uint32_t *status) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < nent; i += stride) {
    const uint64_t v = table[i];
    const uint32_t cnt = (uint32_t)v, start = (uint32_t)(v >> 32);
    if (cnt > 1) atomicOr(&status[1], 1u);
    uniq[i] = cnt == 1 ? rows[start] + 1u : 0u;
  }
}

// The join payload of every DIRECT slot: the build row's packed column
// values and flags (layout in DScan::join_pay).
__global__ void join_payload_kernel(const uint32_t *uniq, uint64_t cap, DJoinPay jp, gv4u *out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < cap; i += stride) {
    const uint32_t u = uniq[i];
    uint64_t b0 = 0, b1 = 0;
    if (u) {
      const uint64_t row = u - 1u;
      uint32_t flags = 1;
      for (int j = 0; j < jp.n; ++j) {
        const DCol &c = jp.cols[j];
        uint64_t v;
        switch (c.dtype) {
          case DT_I8: case DT_BOOL: v = gld<uint8_t>(c.data, row); break;
          case DT_I16: v = gld<uint16_t>(c.data, row * 2); break;
          case DT_I32: case DT_F32: v = gld<uint32_t>(c.data, row * 4); break;
          default: v = gld<uint64_t>(c.data, row * 8); break;
        }
        const int off = jp.off[j];
        if (off < 8) b0 |= v << (8 * off);
        else b1 |= v << (8 * (off - 8));
        if (jp.nbit[j] && load1_null(c, row)) flags |= 1u << jp.nbit[j];
      }
      b1 |= (uint64_t)flags << 32;
    }
    gv4u w;
    w.x = (uint32_t)b0; w.y = (uint32_t)(b0 >> 32); w.z = (uint32_t)b1; w.w = (uint32_t)(b1 >> 32);
    out[i] = w;
  }
}

extern "C" hipError_t qmk_join_payload(const uint32_t *uniq, uint64_t cap, const DJoinPay *jp, uint32_t *out,
                                       hipStream_t st) {
  hipLaunchKernelGGL(join_payload_kernel, dim3(gridn(cap)), dim3(256), 0, st, uniq, cap, *jp,
                     reinterpret_cast<gv4u *>(out));
  return hipGetLastError();
}

// DIRECT index straight from the build rows when every key is unique: each
// row claims its slot (build row + 1); a slot claimed twice sets status[1]
// and the caller falls back to the counted build.
__global__ void join_claim_kernel(DCol c, uint64_t n, const uint64_t *sel, int64_t kmin, uint64_t cap,
                                  uint32_t *uniq, uint32_t *status) {
  uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; j < n; j += stride) {
    const uint64_t i = sel ? sel[j] : j;
    if (load1_null(c, i)) continue;  // NULL never matches (src/operators.ts:2424-2428)
    const uint64_t idx = (uint64_t)(load1_i(c, i) - kmin);
    if (idx >= cap) continue;  // outside the key statistics' domain: cannot happen for this table
    if (atomicCAS(&uniq[idx], 0u, (uint32_t)(i + 1)) != 0u) atomicOr(&status[1], 1u);
  }
}

extern "C" hipError_t qmk_join_claim(const DCol *key, uint64_t nrows, const uint64_t *sel, int64_t kmin, uint64_t cap,
                                     uint32_t *uniq, uint32_t *status, hipStream_t st) {
  hipLaunchKernelGGL(join_claim_kernel, dim3(gridn(nrows)), dim3(256), 0, st, *key, nrows, sel, kmin, cap, uniq,
                     status);
  return hipGetLastError();
}

extern "C" hipError_t qmk_join_unique(const uint64_t *table, uint64_t nent, const uint32_t *rows, uint32_t *uniq,
                                      uint32_t *status, hipStream_t st) {
  hipLaunchKernelGGL(join_unique_kernel, dim3(gridn(nent)), dim3(256), 0, st, table, nent, rows, uniq, status);
  return hipGetLastError();
}

extern "C" hipError_t qmk_join_build(const DCol *key, uint64_t nrows, const uint64_t *sel, int mode, int key_float,
                                     int64_t kmin, uint64_t cap, uint64_t *table, uint32_t *rows, uint32_t *cursor,
                                     uint64_t *tmp, uint32_t *status, hipStream_t st) {
  uint64_t nent = (mode == 1) ? cap : cap + 1;
  hipLaunchKernelGGL(join_count_build_kernel, dim3(gridn(nrows)), dim3(256), 0, st, *key, nrows, sel, mode,
                     key_float, kmin, cap, table, status);
  uint64_t *counts = tmp, *starts = tmp + nent, *stmp = tmp + 2 * nent;
  hipLaunchKernelGGL(join_extract_counts_kernel, dim3(gridn(nent)), dim3(256), 0, st, table, mode, nent, counts);
  hipError_t e = qmk_exscan_u64(counts, nent, starts, nullptr, stmp, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(join_store_starts_kernel, dim3(gridn(nent)), dim3(256), 0, st, table, mode, nent, starts);
  hipLaunchKernelGGL(join_scatter_kernel, dim3(gridn(nrows)), dim3(256), 0, st, *key, nrows, sel, mode, key_float,
                     kmin, cap, table, rows, cursor, status);
  // the scatter is done with counts/starts: tmp now holds the sorted copy of
  // long ranges (nrows u32) and their list; the list's counter is cursor[nent],
  // zero like the rest of cursor
  uint32_t *sorted = reinterpret_cast<uint32_t *>(tmp);
  uint32_t *big = sorted + nrows;
  uint32_t *nbig = cursor + nent;
  hipLaunchKernelGGL(join_sort_segments_kernel, dim3(gridn(nent)), dim3(256), 0, st, table, mode, nent, rows, big,
                     nbig);
  const uint64_t maxbig = nrows / (JOIN_SMALL_SEG + 1);
  if (maxbig > 0)
    hipLaunchKernelGGL(join_sort_big_kernel, dim3((unsigned)std::min<uint64_t>(maxbig, 512)), dim3(JOIN_BIG_BLOCK),
                       0, st, table, mode, rows, big, nbig, sorted);
  return hipGetLastError();
}

extern "C" hipError_t qmk_join_count(const DScan *sc, const DQuery *d_q, uint64_t nblocks, uint64_t tpb,
                                     uint64_t *block_counts, hipStream_t st) {
  hipLaunchKernelGGL(join_pairs_count_kernel, dim3((unsigned)nblocks), dim3(256), 0, st, *sc, d_q, tpb,
                     block_counts);
  return hipGetLastError();
}

extern "C" hipError_t qmk_join_write(const DScan *sc, const DQuery *d_q, uint64_t nblocks, uint64_t tpb,
                                     const uint64_t *block_offsets, uint64_t out_cap, uint64_t *out_left,
                                     uint64_t *out_right, hipStream_t st) {
  hipLaunchKernelGGL(join_pairs_write_kernel, dim3((unsigned)nblocks), dim3(256), 0, st, *sc, d_q, tpb,
                     block_offsets, out_cap, out_left, out_right);
  return hipGetLastError();
}

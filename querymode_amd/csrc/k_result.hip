// k_result.hip — result-side kernels: group-table compaction + finalisation
// (finalizePartialAgg / resolveValue, src/partial-agg.ts:105-134, 306-352),
// filter-only stream compaction to ascending row ids (the selection vector
// ScanOperator returns, src/operators.ts:282-404), column gathers, and a
// device-wide exclusive scan.
#include "k_common.hip.h"
#include "qm_kernels.h"

namespace qm {

// ---------------------------------------------------------------------------
// block-level exclusive scan of one u32 per thread (blockDim <= 1024)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t block_exscan_u32(uint32_t v, uint32_t *total, uint32_t *sh /*>=16*/) {
  const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(x, off, 64);
    if (lane >= (uint32_t)off) x += y;
  }
  if (lane == 63) sh[wid] = x;
  __syncthreads();
  const uint32_t nw = (blockDim.x + 63) >> 6;
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (uint32_t w = 0; w < nw; ++w) { uint32_t t = sh[w]; sh[w] = acc; acc += t; }
    sh[16] = acc;
  }
  __syncthreads();
  uint32_t res = sh[wid] + x - v;
  *total = sh[16];
  __syncthreads();
  return res;
}

// ---------------------------------------------------------------------------
// device-wide exclusive scan of u64 values (3 phases)
// ---------------------------------------------------------------------------
static constexpr int SCAN_BLOCK = 256, SCAN_PER_THREAD = 16;
static constexpr uint64_t SCAN_CHUNK = (uint64_t)SCAN_BLOCK * SCAN_PER_THREAD;

__global__ void scan_reduce_kernel(const uint64_t *in, uint64_t n, uint64_t *sums) {
  __shared__ uint64_t sh[SCAN_BLOCK / 64];
  uint64_t base = (uint64_t)blockIdx.x * SCAN_CHUNK;
  uint64_t acc = 0;
  for (int i = 0; i < SCAN_PER_THREAD; ++i) {
    uint64_t idx = base + (uint64_t)i * SCAN_BLOCK + threadIdx.x;
    if (idx < n) acc += in[idx];
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int w = 0; w < SCAN_BLOCK / 64; ++w) t += sh[w];
    sums[blockIdx.x] = t;
  }
}

// single block of 1024 threads: in-place exclusive scan of `n` block sums,
// total -> *total
__global__ void __launch_bounds__(1024) scan_sums_kernel(uint64_t *sums, uint64_t n, uint64_t *total) {
  __shared__ uint64_t sh[1024];
  const uint64_t per = (n + blockDim.x - 1) / blockDim.x;
  const uint64_t b = (uint64_t)threadIdx.x * per;
  const uint64_t e = (b + per < n) ? b + per : n;
  uint64_t run = 0;
  for (uint64_t i = b; i < e; ++i) run += sums[i];
  sh[threadIdx.x] = run;
  __syncthreads();
  // Hillis-Steele inclusive scan of the thread totals
  for (uint32_t off = 1; off < blockDim.x; off <<= 1) {
    uint64_t v = (threadIdx.x >= off) ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += v;
    __syncthreads();
  }
  uint64_t acc = sh[threadIdx.x] - run;
  for (uint64_t i = b; i < e; ++i) { uint64_t t = sums[i]; sums[i] = acc; acc += t; }
  if (threadIdx.x == blockDim.x - 1 && total) *total = sh[threadIdx.x];
}

__global__ void scan_apply_kernel(const uint64_t *in, uint64_t n, const uint64_t *sums, uint64_t *out) {
  __shared__ uint64_t sh[SCAN_BLOCK];
  uint64_t base = (uint64_t)blockIdx.x * SCAN_CHUNK;
  // each thread scans a contiguous run of SCAN_PER_THREAD elements
  uint64_t mine[SCAN_PER_THREAD];
  uint64_t run = 0;
  for (int i = 0; i < SCAN_PER_THREAD; ++i) {
    uint64_t idx = base + (uint64_t)threadIdx.x * SCAN_PER_THREAD + i;
    mine[i] = idx < n ? in[idx] : 0;
    run += mine[i];
  }
  sh[threadIdx.x] = run;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t acc = sums[blockIdx.x];
    for (int t = 0; t < SCAN_BLOCK; ++t) { uint64_t v = sh[t]; sh[t] = acc; acc += v; }
  }
  __syncthreads();
  uint64_t acc = sh[threadIdx.x];
  for (int i = 0; i < SCAN_PER_THREAD; ++i) {
    uint64_t idx = base + (uint64_t)threadIdx.x * SCAN_PER_THREAD + i;
    if (idx < n) out[idx] = acc;
    acc += mine[i];
  }
}

// ---------------------------------------------------------------------------
// group-table compaction + finalisation
// ---------------------------------------------------------------------------
static constexpr int CMP_BLOCK = 256;
static constexpr uint64_t CMP_SLOTS_PER_BLOCK = 4096;

__device__ __forceinline__ bool slot_occupied(const DFinal &f, uint64_t slot) {
  if (f.always_one_row) return slot == 0;
  const uint64_t *e = f.table + slot * f.entry_words;
  return e[f.count_word] != 0;
}

__global__ void compact_count_kernel(DFinal f, uint64_t *block_counts) {
  uint64_t base = (uint64_t)blockIdx.x * CMP_SLOTS_PER_BLOCK;
  uint64_t cnt = 0;
  for (uint64_t i = threadIdx.x; i < CMP_SLOTS_PER_BLOCK; i += blockDim.x) {
    uint64_t slot = base + i;
    if (slot < f.nslots && slot_occupied(f, slot)) ++cnt;
  }
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
  __shared__ uint64_t sh[CMP_BLOCK / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int w = 0; w < CMP_BLOCK / 64; ++w) t += sh[w];
    block_counts[blockIdx.x] = t;
  }
}

__device__ void decode_key(const DFinal &f, uint64_t slot, int64_t v[MAX_KEYS], bool isnull[MAX_KEYS]) {
  const DKey &k = f.key;
  for (int i = 0; i < MAX_KEYS; ++i) { v[i] = 0; isnull[i] = false; }
  if (f.gmode == GM_DIRECT) {
    for (int i = 0; i < k.ncols; ++i) {
      uint64_t code = (slot / k.stride[i]) % k.dom[i];
      if (k.nullable[i] && code == k.dom[i] - 1) isnull[i] = true;
      else v[i] = (int64_t)code + k.kmin[i];
    }
    return;
  }
  // GM_HASH
  if (slot == f.nslots - 1) { for (int i = 0; i < k.ncols; ++i) isnull[i] = true; return; }
  uint64_t raw;
  if (slot == f.nslots - 2) raw = KEY_UNREPRESENTABLE;
  else raw = f.table[slot * f.entry_words] - KEY_SALT;
  if (k.ncols == 1) { v[0] = (int64_t)raw; return; }
  int shift = 0;
  for (int i = 0; i < k.ncols; ++i) {
    uint64_t m = (k.bits[i] >= 64) ? ~0ull : ((1ull << k.bits[i]) - 1);
    uint64_t code = (raw >> shift) & m;
    shift += k.bits[i];
    if (k.nullable[i] && code == k.dom[i] - 1) isnull[i] = true;
    else v[i] = (int64_t)code + k.kmin[i];
  }
}

__device__ __forceinline__ void store_typed_i(void *out, int32_t dt, uint64_t i, int64_t v) {
  switch (dt) {
    case DT_I8: static_cast<int8_t *>(out)[i] = (int8_t)v; break;
    case DT_BOOL: static_cast<uint8_t *>(out)[i] = (uint8_t)v; break;
    case DT_I16: static_cast<int16_t *>(out)[i] = (int16_t)v; break;
    case DT_I32: static_cast<int32_t *>(out)[i] = (int32_t)v; break;
    default: static_cast<int64_t *>(out)[i] = v; break;
  }
}
__device__ __forceinline__ void store_typed_f(void *out, int32_t dt, uint64_t i, double v) {
  if (dt == DT_F32) static_cast<float *>(out)[i] = (float)v;
  else static_cast<double *>(out)[i] = v;
}

__device__ void finalize_slot(const DFinal &f, uint64_t slot, uint64_t pos) {
  const uint64_t *e = f.table + slot * f.entry_words;
  int64_t kv[MAX_KEYS]; bool kn[MAX_KEYS];
  bool have_key = false;
  for (int oi = 0; oi < f.nouts; ++oi) {
    const DOut &o = f.outs[oi];
    uint8_t isnull = 0;
    uint64_t cnt = (o.cnt_word >= 0) ? e[o.cnt_word] : 0;
    switch (o.kind) {
      case OK_KEY: {
        if (!have_key) { decode_key(f, slot, kv, kn); have_key = true; }
        if (kn[o.key_idx]) { isnull = 1; store_typed_i(o.data, o.out_dtype == DT_F64 ? DT_I64 : o.out_dtype, pos, 0); }
        else if (o.out_dtype == DT_F64) static_cast<double *>(o.data)[pos] = bits_f64((uint64_t)kv[o.key_idx]);
        else store_typed_i(o.data, o.out_dtype, pos, kv[o.key_idx]);
        break;
      }
      case OK_COUNT:
        static_cast<int64_t *>(o.data)[pos] = (int64_t)cnt;
        break;
      case OK_SUM_I64:
        if (cnt == 0) { isnull = 1; static_cast<int64_t *>(o.data)[pos] = 0; }
        else static_cast<int64_t *>(o.data)[pos] = (int64_t)e[o.word];
        break;
      case OK_SUM_I128: {
        if (cnt == 0) { isnull = 1; static_cast<int64_t *>(o.data)[pos] = 0; break; }
        uint64_t lo = e[o.word], hi = e[o.word + 1];
        bool fits = (hi == 0 && (int64_t)lo >= 0) || (hi == ~0ull && (int64_t)lo < 0);
        if (!fits && f.status) f.status[1] = 1;
        static_cast<int64_t *>(o.data)[pos] = (int64_t)lo;
        break;
      }
      case OK_SUM_F64: {
        if (cnt == 0) { isnull = 1; static_cast<double *>(o.data)[pos] = 0; break; }
        u128 q; q.lo = e[o.word]; q.hi = e[o.word + 1];
        static_cast<double *>(o.data)[pos] = finish_f64_sum(q, e[o.word + 2], o.scale);
        break;
      }
      case OK_AVG_F64: {
        if (cnt == 0) { isnull = 1; static_cast<double *>(o.data)[pos] = 0; break; }
        u128 q; q.lo = e[o.word]; q.hi = e[o.word + 1];
        static_cast<double *>(o.data)[pos] = finish_f64_sum(q, e[o.word + 2], o.scale) / (double)cnt;
        break;
      }
      case OK_SUM_WIDE: {
        if (cnt == 0) { isnull = 1; static_cast<double *>(o.data)[pos] = 0; break; }
        static_cast<double *>(o.data)[pos] = wide_to_f64(e + o.word + 1, o.nw, o.scale, e[o.word]);
        break;
      }
      case OK_AVG_WIDE: {
        if (cnt == 0) { isnull = 1; static_cast<double *>(o.data)[pos] = 0; break; }
        static_cast<double *>(o.data)[pos] = wide_to_f64(e + o.word + 1, o.nw, o.scale, e[o.word]) / (double)cnt;
        break;
      }
      case OK_AVG_I64: {
        if (cnt == 0) { isnull = 1; static_cast<double *>(o.data)[pos] = 0; break; }
        static_cast<double *>(o.data)[pos] = (double)(int64_t)e[o.word] / (double)cnt;
        break;
      }
      case OK_AVG_BIG: {
        if (cnt == 0) { isnull = 1; static_cast<double *>(o.data)[pos] = 0; break; }
        u128 q; q.lo = e[o.word]; q.hi = e[o.word + 1];
        static_cast<double *>(o.data)[pos] = bigint_avg(q, cnt);
        break;
      }
      default: {  // OK_MIN / OK_MAX
        uint64_t st = e[o.word];
        if (cnt == 0) {
          isnull = 1;
          if (o.src_is_float) store_typed_f(o.data, o.out_dtype, pos, 0.0);
          else store_typed_i(o.data, o.out_dtype, pos, 0);
          break;
        }
        if (o.src_is_float) {
          double v;
          if (st == 0) v = (o.kind == OK_MIN) ? bits_f64(0x7FF0000000000000ull) : bits_f64(0xFFF0000000000000ull);
          else v = bits_f64(decode_f64_code(o.kind == OK_MIN ? ~st : st));
          store_typed_f(o.data, o.out_dtype, pos, v);
        } else {
          store_typed_i(o.data, o.out_dtype, pos, decode_i64(o.kind == OK_MIN ? ~st : st));
        }
        break;
      }
    }
    if (o.nulls && isnull) atomicOr(&o.nulls[pos >> 5], 1u << (pos & 31));
  }
}

__global__ void compact_write_kernel(DFinal f, const uint64_t *block_offsets) {
  __shared__ uint32_t sh[17];
  uint64_t base = (uint64_t)blockIdx.x * CMP_SLOTS_PER_BLOCK;
  uint64_t out_base = block_offsets[blockIdx.x];
  for (uint64_t i0 = 0; i0 < CMP_SLOTS_PER_BLOCK; i0 += blockDim.x) {
    uint64_t slot = base + i0 + threadIdx.x;
    uint32_t occ = (slot < f.nslots && slot_occupied(f, slot)) ? 1u : 0u;
    uint32_t total;
    uint32_t off = block_exscan_u32(occ, &total, sh);
    if (occ) finalize_slot(f, slot, out_base + off);
    out_base += total;
  }
}

// Many small tables compacted in one launch: workgroup i serves fs[i] (at
// most CMP_SMALL_SLOTS slots), clears its output NULL bitmaps, then writes its
// occupied slots in slot order; counts[i] = rows written.
static constexpr uint64_t CMP_SMALL_SLOTS = 4096;
__global__ void __launch_bounds__(CMP_BLOCK) compact_small_kernel(const DFinal *fs, uint64_t *counts) {
  __shared__ uint32_t sh[17];
  const DFinal &f = fs[blockIdx.x];
  const uint64_t nslots = f.nslots < CMP_SMALL_SLOTS ? f.nslots : CMP_SMALL_SLOTS;
  const uint64_t words = (nslots + 31) / 32;
  for (int oi = 0; oi < f.nouts; ++oi) {
    uint32_t *nb = f.outs[oi].nulls;
    if (nb)
      for (uint64_t w = threadIdx.x; w < words; w += blockDim.x) nb[w] = 0;
  }
  __syncthreads();
  uint64_t out = 0;
  for (uint64_t i0 = 0; i0 < nslots; i0 += blockDim.x) {
    const uint64_t slot = i0 + threadIdx.x;
    const uint32_t occ = (slot < nslots && slot_occupied(f, slot)) ? 1u : 0u;
    uint32_t total;
    const uint32_t off = block_exscan_u32(occ, &total, sh);
    if (occ) finalize_slot(f, slot, out + off);
    out += total;
  }
  if (threadIdx.x == 0) counts[blockIdx.x] = out;
}

// ---------------------------------------------------------------------------
// filter-only selection: ascending row ids of rows passing the predicates.
// Each block owns a contiguous range of tiles so the output order is the
// table order (the order of the reference's index lists).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) select_count_kernel(DScan sc, const DQuery *__restrict__ qs,
                                                           uint64_t tiles_per_block, uint64_t *block_counts) {
  const DQuery &q = qs[0];
  const uint64_t tile_rows = (uint64_t)blockDim.x * RPT;
  uint64_t t0 = (uint64_t)blockIdx.x * tiles_per_block;
  uint64_t cnt = 0;
  for (uint64_t t = t0; t < t0 + tiles_per_block; ++t) {
    uint64_t row0 = t * tile_rows + (uint64_t)threadIdx.x * RPT;
    if (row0 >= sc.nrows) break;
    if (page_skipped(sc, t * tile_rows)) continue;
    uint64_t left = sc.nrows - row0;
    uint32_t valid = left >= RPT ? 0xFFu : ((1u << left) - 1);
    cnt += __builtin_popcount(eval_query8(q, sc.cols, row0, valid));
  }
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
  __shared__ uint64_t sh[4];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) block_counts[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ void __launch_bounds__(256) select_write_kernel(DScan sc, const DQuery *__restrict__ qs,
                                                           uint64_t tiles_per_block, const uint64_t *block_offsets,
                                                           uint64_t out_cap, uint64_t *out_rows) {
  __shared__ uint32_t sh[17];
  const DQuery &q = qs[0];
  const uint64_t tile_rows = (uint64_t)blockDim.x * RPT;
  uint64_t t0 = (uint64_t)blockIdx.x * tiles_per_block;
  uint64_t out_base = block_offsets[blockIdx.x];
  for (uint64_t t = t0; t < t0 + tiles_per_block; ++t) {
    if (t * tile_rows >= sc.nrows) break;   // uniform across the block
    if (out_base >= out_cap) break;          // limit reached
    if (page_skipped(sc, t * tile_rows)) continue;  // uniform: no row of a skipped page is written
    uint64_t row0 = t * tile_rows + (uint64_t)threadIdx.x * RPT;
    uint32_t m = 0;
    if (row0 < sc.nrows) {
      uint64_t left = sc.nrows - row0;
      uint32_t valid = left >= RPT ? 0xFFu : ((1u << left) - 1);
      m = eval_query8(q, sc.cols, row0, valid);
    }
    uint32_t total;
    uint32_t off = block_exscan_u32(__builtin_popcount(m), &total, sh);
    uint64_t pos = out_base + off;
    for (int r = 0; r < RPT; ++r) {
      if ((m >> r) & 1) {
        if (pos < out_cap) out_rows[pos] = row0 + r;
        ++pos;
      }
    }
    out_base += total;
  }
}

// ---------------------------------------------------------------------------
// gather through a u32 index (the join build side): out[i] = col[idx[rows[i]]]
__global__ void gather2_kernel(DCol c, const uint64_t *rows, const uint32_t *idx, uint64_t n, int32_t out_dt,
                               void *out, uint32_t *nulls) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    uint64_t r = idx ? idx[rows[i]] : rows[i];
    if (dt_is_float(c.dtype)) store_typed_f(out, out_dt, i, load1_f(c, r));
    else store_typed_i(out, out_dt, i, load1_i(c, r));
    if (nulls && load1_null(c, r)) atomicOr(&nulls[i >> 5], 1u << (i & 31));
  }
}

}  // namespace qm

using namespace qm;

static unsigned grid_for(uint64_t n, unsigned block, unsigned cap = 65536) {
  uint64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

extern "C" hipError_t qmk_exscan_u64(const uint64_t *in, uint64_t n, uint64_t *out, uint64_t *total,
                                     uint64_t *tmp_sums, hipStream_t st) {
  if (n == 0) return hipSuccess;
  uint64_t nb = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
  hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, st, in, n, tmp_sums);
  hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(1024), 0, st, tmp_sums, nb, total);
  hipLaunchKernelGGL(scan_apply_kernel, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, st, in, n, tmp_sums, out);
  return hipGetLastError();
}
extern "C" uint64_t qmk_exscan_tmp_words(uint64_t n) { return (n + SCAN_CHUNK - 1) / SCAN_CHUNK + 1; }

extern "C" uint64_t qmk_compact_blocks(uint64_t nslots) {
  return (nslots + CMP_SLOTS_PER_BLOCK - 1) / CMP_SLOTS_PER_BLOCK;
}
extern "C" hipError_t qmk_compact_count(const DFinal *f, uint64_t *block_counts, hipStream_t st) {
  uint64_t nb = qmk_compact_blocks(f->nslots);
  hipLaunchKernelGGL(compact_count_kernel, dim3((unsigned)nb), dim3(CMP_BLOCK), 0, st, *f, block_counts);
  return hipGetLastError();
}
extern "C" uint64_t qmk_compact_small_slots() { return CMP_SMALL_SLOTS; }
extern "C" hipError_t qmk_compact_small(const DFinal *d_f, uint32_t n, uint64_t *counts, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(compact_small_kernel, dim3(n), dim3(CMP_BLOCK), 0, st, d_f, counts);
  return hipGetLastError();
}
extern "C" hipError_t qmk_compact_write(const DFinal *f, const uint64_t *block_offsets, hipStream_t st) {
  uint64_t nb = qmk_compact_blocks(f->nslots);
  hipLaunchKernelGGL(compact_write_kernel, dim3((unsigned)nb), dim3(CMP_BLOCK), 0, st, *f, block_offsets);
  return hipGetLastError();
}

extern "C" hipError_t qmk_select_count(const DScan *sc, const DQuery *d_q, uint64_t nblocks, uint64_t tpb,
                                       uint64_t *block_counts, hipStream_t st) {
  hipLaunchKernelGGL(select_count_kernel, dim3((unsigned)nblocks), dim3(256), 0, st, *sc, d_q, tpb, block_counts);
  return hipGetLastError();
}
extern "C" hipError_t qmk_select_write(const DScan *sc, const DQuery *d_q, uint64_t nblocks, uint64_t tpb,
                                       const uint64_t *block_offsets, uint64_t out_cap, uint64_t *out_rows,
                                       hipStream_t st) {
  hipLaunchKernelGGL(select_write_kernel, dim3((unsigned)nblocks), dim3(256), 0, st, *sc, d_q, tpb,
                     block_offsets, out_cap, out_rows);
  return hipGetLastError();
}

extern "C" hipError_t qmk_gather(const DCol *c, const uint64_t *rows, const uint32_t *idx, uint64_t n,
                                 int32_t out_dt, void *out, uint32_t *nulls, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(gather2_kernel, dim3(grid_for(n, 256)), dim3(256), 0, st, *c, rows, idx, n, out_dt, out,
                     nulls);
  return hipGetLastError();
}

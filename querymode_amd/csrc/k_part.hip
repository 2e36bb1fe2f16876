// k_part.hip — partitioned hash aggregation for high-cardinality GROUP BY.
//
// Same semantics as the fused scan (AggregateOperator + partial-agg states,
// src/operators.ts:1646-1710, src/partial-agg.ts:141-352); different data
// movement. Global atomics run at a fixed device-wide rate, so per-row
// atomics on a table of 10^5..10^7 groups bound the fused kernel far below
// HBM bandwidth. Here the selected rows are moved, not aggregated in place:
//   pass 1  filter + compact selected rows into records in a region private
//           to the workgroup (wave-level reservation through an LDS counter,
//           no global atomics), count records per key partition in LDS;
//   pass 2  scatter records into partition order (LDS cursors);
//   pass 3  one workgroup per partition segment aggregates in an LDS table and
//           stores its groups into the global table (plain stores when the
//           workgroup owns the whole partition).
// Global atomics drop from several per selected row to about one per group.
#include "k_fast.hip.h"
#include "qm_kernels.h"

namespace qm {

static constexpr int PART_BLOCK = 256;
static constexpr int STAGE_WORDS = 512;  // per-wave LDS record stage (4 KB): 128 records of 4 words
static constexpr int UNROLL = 4;
static constexpr int MAX_RW = 8;  // record words: 4 or 8 (32/64-byte records, copied as 16-byte vectors)

// record copy / load in 16-byte pieces (records are 32-byte aligned)
__device__ __forceinline__ void rec_load(const uint64_t *src, int R, uint64_t rw[MAX_RW]) {
  const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
#pragma unroll
  for (int q = 0; q < MAX_RW / 2; ++q) {
    if (2 * q < R) {
      uint4 v = s4[q];
      rw[2 * q] = ((uint64_t)v.y << 32) | v.x;
      rw[2 * q + 1] = ((uint64_t)v.w << 32) | v.z;
    }
  }
}
__device__ __forceinline__ void rec_copy(uint64_t *dst, const uint64_t *src, int R) {
  const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
  uint4 *d4 = reinterpret_cast<uint4 *>(dst);
  uint4 v[MAX_RW / 2];
#pragma unroll
  for (int q = 0; q < MAX_RW / 2; ++q) if (2 * q < R) v[q] = s4[q];
#pragma unroll
  for (int q = 0; q < MAX_RW / 2; ++q) if (2 * q < R) d4[q] = v[q];
}

// Hash partition of a key code: the high 32 bits of the mixed code scaled to
// [0, nparts) (any partition count; the LDS tables of pass 3 index by the low
// bits, which stay independent).
__device__ __forceinline__ uint32_t hash_part(uint64_t code, uint32_t nparts) {
  return (uint32_t)(((mix64(code) >> 32) * (uint64_t)nparts) >> 32);
}

__device__ __forceinline__ uint32_t part_of(const DQuery &q, const DPart &pp, uint64_t code) {
  if (q.gmode == GM_DIRECT) return (uint32_t)(code >> pp.sbits);
  return pp.pbits ? hash_part(code, pp.nparts) : 0u;
}

// heavy-hitter table lookup (HOT_SLOTS slots, linear probing)
__device__ __forceinline__ int hot_slot(const uint64_t *hot, uint64_t hk) {
  uint32_t h = (uint32_t)mix64(hk ^ 0xA5A5A5A5A5A5A5A5ull) & (HOT_SLOTS - 1);
#pragma unroll
  for (int p = 0; p < HOT_PROBES; ++p) {
    uint32_t s = (h + p) & (HOT_SLOTS - 1);
    uint64_t k = hot[s];
    if (k == hk) return (int)s;
    if (k == 0) return -1;
  }
  return -1;
}

__device__ __forceinline__ void load_hot(const DPart &pp, uint64_t *hot_lds) {
  for (int i = threadIdx.x; i < HOT_SLOTS; i += blockDim.x) hot_lds[i] = pp.hot_n ? pp.hot_keys[i] : 0;
}

// ---- pass 1: filter + compact + partition histogram ----
__global__ void __launch_bounds__(PART_BLOCK) part_scan_kernel(DScan sc, const DQuery *__restrict__ qs, DPart pp,
                                                               uint64_t tiles_per_block) {
  extern __shared__ uint32_t hist_lds[];  // nparts + 1 counters (the last: heavy hitters)
  __shared__ uint32_t s_fill;
  __shared__ uint64_t hot_lds[HOT_SLOTS];
  const DQuery &q = qs[0];
  for (int i = threadIdx.x; i <= pp.nparts; i += blockDim.x) hist_lds[i] = 0;
  if (threadIdx.x == 0) s_fill = 0;
  load_hot(pp, hot_lds);
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t tile_rows = (uint64_t)blockDim.x * RPT;
  const uint64_t t0 = (uint64_t)blockIdx.x * tiles_per_block;
  const uint64_t ntiles = (sc.nrows + tile_rows - 1) / tile_rows;
  const uint64_t t1 = (t0 + tiles_per_block < ntiles) ? t0 + tiles_per_block : ntiles;
  const int R = pp.rec_words;
  uint64_t *region = pp.recs + (uint64_t)blockIdx.x * pp.blk_cap * R;
  for (uint64_t t = t0; t < t1; ++t) {
    const uint64_t row0 = t * tile_rows + (uint64_t)threadIdx.x * RPT;
    uint32_t valid = 0;
    if (row0 < sc.nrows) {
      uint64_t left = sc.nrows - row0;
      valid = left >= RPT ? 0xFFu : ((1u << left) - 1);
    }
    const uint32_t m = eval_query8(q, sc.cols, row0, valid);
    // wave-level exclusive scan of the per-lane counts, one LDS reservation per wave
    const uint32_t cnt = __builtin_popcount(m);
    uint32_t x = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      uint32_t y = __shfl_up(x, off, 64);
      if (lane >= (uint32_t)off) x += y;
    }
    const uint32_t wave_total = __shfl(x, 63, 64);
    if (wave_total == 0) continue;  // uniform per wave
    uint32_t base = 0;
    if (lane == 63) base = atomicAdd(&s_fill, wave_total);
    base = __shfl(base, 63, 64);
    if (!m) continue;
    uint64_t pos = (uint64_t)base + x - cnt;
    // key codes
    int64_t kv[MAX_KEYS][RPT];
    uint32_t kn[MAX_KEYS] = {0, 0};
#pragma unroll
    for (int i = 0; i < MAX_KEYS; ++i) {
      if (i >= q.key.ncols) break;
      const DCol &c = sc.cols[q.key.col[i]];
      kn[i] = load8_nulls(c, row0);
      if (q.key.is_float[i]) {
        double f[RPT];
        load8_f(c, row0, f);
#pragma unroll
        for (int r = 0; r < RPT; ++r) kv[i][r] = (int64_t)canon_f64_key_bits(f64_bits(f[r]));
      } else {
        load8_i(c, row0, kv[i]);
      }
    }
    uint64_t rec_idx[RPT], mask_w[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      rec_idx[r] = ~0ull;
      mask_w[r] = 0;
      if (!((m >> r) & 1)) continue;
      int64_t v[MAX_KEYS] = {0, 0};
      uint32_t n[MAX_KEYS] = {0, 0};
#pragma unroll
      for (int i = 0; i < MAX_KEYS; ++i)
        if (i < q.key.ncols) { v[i] = kv[i][r]; n[i] = (kn[i] >> r) & 1; }
      uint64_t code;
      if (q.gmode == GM_DIRECT) {
        code = key_index_direct(q.key, v, n);
      } else {
        int side;
        code = key_code_hash(q.key, v, n, &side);
        if (side >= 0) mask_w[r] = (1ull << 63) | ((uint64_t)side << 62);
      }
      rec_idx[r] = pos++;
      region[rec_idx[r] * R] = code;
      uint32_t p = (mask_w[r] >> 63) ? 0u : part_of(q, pp, code);
      if (pp.hot_n && !(mask_w[r] >> 63) && hot_slot(hot_lds, q.gmode == GM_DIRECT ? code + 1 : code) >= 0)
        p = pp.nparts;
      atomicAdd(&hist_lds[p], 1u);
    }
    for (int j = 0; j < pp.nsrc; ++j) {
      const DCol &a = sc.cols[pp.src_a[j]];
      uint32_t nm = load8_nulls(a, row0);
      uint64_t w[RPT];
      if (pp.src_b[j] >= 0) {
        const DCol &b = sc.cols[pp.src_b[j]];
        double xa[RPT], yb[RPT];
        if (dt_is_float(a.dtype)) load8_f(a, row0, xa);
        else { int64_t t8[RPT]; load8_i(a, row0, t8); for (int r = 0; r < RPT; ++r) xa[r] = (double)t8[r]; }
        if (dt_is_float(b.dtype)) load8_f(b, row0, yb);
        else { int64_t t8[RPT]; load8_i(b, row0, t8); for (int r = 0; r < RPT; ++r) yb[r] = (double)t8[r]; }
#pragma unroll
        for (int r = 0; r < RPT; ++r) w[r] = f64_bits(xa[r] * yb[r]);
        nm |= load8_nulls(b, row0);
      } else if (dt_is_float(a.dtype)) {
        double xa[RPT];
        load8_f(a, row0, xa);
#pragma unroll
        for (int r = 0; r < RPT; ++r) w[r] = f64_bits(xa[r]);
      } else {
        int64_t xa[RPT];
        load8_i(a, row0, xa);
#pragma unroll
        for (int r = 0; r < RPT; ++r) w[r] = (uint64_t)xa[r];
      }
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        if (rec_idx[r] == ~0ull) continue;
        region[rec_idx[r] * R + 1 + j] = w[r];
        if ((nm >> r) & 1) mask_w[r] |= (1ull << j);
      }
    }
    if (pp.has_mask) {
#pragma unroll
      for (int r = 0; r < RPT; ++r)
        if (rec_idx[r] != ~0ull) region[rec_idx[r] * R + R - 1] = mask_w[r];
    }
  }
  __syncthreads();
  for (int p = threadIdx.x; p <= pp.nparts; p += blockDim.x)
    pp.hist[(uint64_t)p * gridDim.x + blockIdx.x] = hist_lds[p];
  if (threadIdx.x == 0) pp.blk_count[blockIdx.x] = s_fill;
}

// ---- pass 1, register-descriptor form: AND of one range per column, <= 4
// sources. All predicate columns (and, with `prefetch`, the key columns) of a
// tile are loaded before the first compare. ----
template <int NS, int NK>
__global__ void __launch_bounds__(PART_BLOCK) __attribute__((amdgpu_waves_per_eu(3))) part_scan_fast_kernel(DScan sc, const DQuery *__restrict__ qs,
                                                                    DPart pp, uint64_t tiles_per_block,
                                                                    int prefetch) {
  extern __shared__ uint32_t hist_lds[];
  __shared__ uint32_t s_fill;
  __shared__ uint64_t hot_lds[HOT_SLOTS];
  __shared__ __align__(16) uint64_t stage[PART_BLOCK / 64][STAGE_WORDS];
  const DQuery &q = qs[0];
  for (int i = threadIdx.x; i <= pp.nparts; i += blockDim.x) hist_lds[i] = 0;
  if (threadIdx.x == 0) s_fill = 0;
  load_hot(pp, hot_lds);
  const int hot_n = pp.hot_n;
  uint64_t *stage_w = stage[threadIdx.x >> 6];
  const FastFilter F = fast_filter_of(q, sc.cols);
  const FastKey K = fast_key_of(q, sc.cols);
  const FastSrc S = fast_src_of(pp, q, sc.cols);
  const DKey key = q.key;
  const int R = pp.rec_words;
  const int has_mask = pp.has_mask;
  const bool direct = q.gmode == GM_DIRECT;
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t tile_rows = (uint64_t)blockDim.x * RPT;
  const uint64_t t0 = (uint64_t)blockIdx.x * tiles_per_block;
  const uint64_t ntiles = (sc.nrows + tile_rows - 1) / tile_rows;
  const uint64_t t1 = (t0 + tiles_per_block < ntiles) ? t0 + tiles_per_block : ntiles;
  uint64_t *region = pp.recs + (uint64_t)blockIdx.x * pp.blk_cap * R;
  for (uint64_t t = t0; t < t1; ++t) {
    const uint64_t row0 = t * tile_rows + (uint64_t)threadIdx.x * RPT;
    uint32_t valid = 0;
    if (row0 < sc.nrows) {
      uint64_t left = sc.nrows - row0;
      valid = left >= RPT ? 0xFFu : ((1u << left) - 1);
    }
    int64_t pv[NS][RPT];
    uint32_t pn[NS];
    int64_t kv[NK][RPT];
    uint32_t kn[NK];
    if (valid) {
      fast_load<NS>(F, row0, pv, pn);
      if (prefetch) fast_key_load<NK>(K, row0, kv, kn);
    }
    const uint32_t m = valid ? fast_eval<NS>(F, pv, pn, valid) : 0u;
    const uint32_t cnt = __builtin_popcount(m);
    uint32_t x = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      uint32_t y = __shfl_up(x, off, 64);
      if (lane >= (uint32_t)off) x += y;
    }
    const uint32_t wave_total = __shfl(x, 63, 64);
    if (wave_total == 0) continue;  // uniform per wave
    uint32_t base = 0;
    if (lane == 63) base = atomicAdd(&s_fill, wave_total);
    base = __shfl(base, 63, 64);
    // The wave's records go to region[base, base + wave_total). When they fit
    // in the wave's LDS stage they are assembled there and copied out as
    // coalesced 16-byte stores; otherwise each lane stores its own words.
    const bool staged = wave_total * (uint32_t)R <= (uint32_t)STAGE_WORDS;  // uniform per wave
    uint64_t *dst = staged ? stage_w : region;
    const uint64_t pos0 = staged ? (uint64_t)(x - cnt) : (uint64_t)base + x - cnt;
    if (m) {
    if (!prefetch) fast_key_load<NK>(K, row0, kv, kn);
    uint32_t side_m = 0, nullk_m = 0;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      if (!((m >> r) & 1)) continue;
      int64_t v[MAX_KEYS] = {0, 0};
      uint32_t n[MAX_KEYS] = {0, 0};
#pragma unroll
      for (int i = 0; i < NK; ++i) { v[i] = kv[i][r]; n[i] = (kn[i] >> r) & 1; }
      int sd;
      const uint64_t code = key_code_fixed<NK>(key, direct, v, n, &sd);
      uint32_t p;
      if (direct) {
        p = (uint32_t)(code >> pp.sbits);
      } else {
        if (sd >= 0) { side_m |= 1u << r; if (sd) nullk_m |= 1u << r; }
        p = (sd >= 0 || !pp.pbits) ? 0u : hash_part(code, pp.nparts);
      }
      if (hot_n && sd < 0 && hot_slot(hot_lds, direct ? code + 1 : code) >= 0) p = pp.nparts;
      const uint64_t ri = pos0 + __builtin_popcount(m & ((1u << r) - 1));
      dst[ri * R] = code;
      atomicAdd(&hist_lds[p], 1u);
    }
    uint32_t null_src[MAX_FSRC] = {0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < MAX_FSRC; ++j) {
      if (j >= S.n) break;
      if (S.b[j]) {
        int64_t xa[RPT], yb[RPT];
        load8_any(S.a[j], S.dta[j], row0, xa);
        load8_any(S.b[j], S.dtb[j], row0, yb);
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          if (!((m >> r) & 1)) continue;
          double fa = dt_is_float(S.dta[j]) ? bits_f64((uint64_t)xa[r]) : (double)xa[r];
          double fb = dt_is_float(S.dtb[j]) ? bits_f64((uint64_t)yb[r]) : (double)yb[r];
          dst[(pos0 + __builtin_popcount(m & ((1u << r) - 1))) * R + 1 + j] = f64_bits(fa * fb);
        }
        null_src[j] = nulls8(S.na[j], row0) | nulls8(S.nb[j], row0);
      } else if (S.slot[j] >= 0) {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          if (S.slot[j] != s) continue;
#pragma unroll
          for (int r = 0; r < RPT; ++r) {
            if (!((m >> r) & 1)) continue;
            dst[(pos0 + __builtin_popcount(m & ((1u << r) - 1))) * R + 1 + j] = (uint64_t)pv[s][r];
          }
        }
        null_src[j] = nulls8(S.na[j], row0);
      } else {
        int64_t xa[RPT];
        load8_any(S.a[j], S.dta[j], row0, xa);
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          if (!((m >> r) & 1)) continue;
          dst[(pos0 + __builtin_popcount(m & ((1u << r) - 1))) * R + 1 + j] = (uint64_t)xa[r];
        }
        null_src[j] = nulls8(S.na[j], row0);
      }
    }
    if (has_mask) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        if (!((m >> r) & 1)) continue;
        uint64_t mw = 0;
#pragma unroll
        for (int j = 0; j < MAX_FSRC; ++j) mw |= (uint64_t)((null_src[j] >> r) & 1) << j;
        if ((side_m >> r) & 1) mw |= (1ull << 63) | ((uint64_t)((nullk_m >> r) & 1) << 62);
        dst[(pos0 + __builtin_popcount(m & ((1u << r) - 1))) * R + R - 1] = mw;
      }
    }
    }  // m
    if (staged) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const uint32_t n16 = wave_total * (uint32_t)R / 2;
      const uint4 *src4 = reinterpret_cast<const uint4 *>(stage_w);
      uint4 *dst4 = reinterpret_cast<uint4 *>(region + (uint64_t)base * R);
      for (uint32_t u = lane; u < n16; u += 64) dst4[u] = src4[u];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
  __syncthreads();
  for (int p = threadIdx.x; p <= pp.nparts; p += blockDim.x)
    pp.hist[(uint64_t)p * gridDim.x + blockIdx.x] = hist_lds[p];
  if (threadIdx.x == 0) pp.blk_count[blockIdx.x] = s_fill;
}

// apply one record's state contributions to an LDS or global entry
__device__ __forceinline__ void rec_apply(const DQuery &q, const DPart &pp, const uint64_t *rw, int R,
                                          lds_u64 *lent, uint64_t *gent) {
  for (int si = 0; si < q.nstates; ++si) {
    const DState &s = q.states[si];
    const int8_t src = pp.state_src[si];
    Contrib k;
    if (src < 0) {
      k.a = 1; k.b = k.c = 0; k.skip = false;
    } else {
      if (pp.has_mask && ((rw[R - 1] >> src) & 1)) continue;
      uint64_t w = rw[1 + src];
      bool isf = pp.src_float[src] != 0;
      k = contrib_of(s, isf, isf ? 0 : (int64_t)w, isf ? bits_f64(w) : 0.0);
      if (k.skip) continue;
    }
    if (lent) apply_lds(s, lent, k);
    else apply_glb(s, gent, k);
  }
}

// ---- pass 2: scatter block regions into partition order. Workgroup b2
// handles pass-1 regions [b2*rpb, (b2+1)*rpb): the partition ranges of those
// regions are adjacent in the partition-major offsets, so each workgroup
// writes one contiguous run per partition. Heavy-hitter records are
// aggregated in LDS here instead of being written. ----
__global__ void __launch_bounds__(1024) part_scatter_kernel(const DQuery *__restrict__ qs, DPart pp, uint32_t g1,
                                                            uint32_t rpb) {
  extern __shared__ uint64_t cur[];  // nparts + 1 cursors, then HOT_SLOTS hot entries
  __shared__ uint64_t hot_lds[HOT_SLOTS];
  const DQuery &q = qs[0];
  const int W = q.entry_words;
  const uint32_t b0 = blockIdx.x * rpb;
  const uint32_t b1 = (b0 + rpb < g1) ? b0 + rpb : g1;
  const int nparts = pp.nparts;
  lds_u64 *hent = (lds_u64 *)(cur + nparts + 1);
  for (int p = threadIdx.x; p <= nparts; p += blockDim.x) cur[p] = pp.part_off[(uint64_t)p * g1 + b0];
  load_hot(pp, hot_lds);
  const int hot_n = pp.hot_n;
  if (hot_n)
    for (int i = threadIdx.x; i < HOT_SLOTS * W; i += blockDim.x) hent[i] = 0;
  __syncthreads();
  const int R = pp.rec_words;
  const int pbits = pp.pbits, sbits = pp.sbits;
  const bool direct = q.gmode == GM_DIRECT;
  const int has_mask = pp.has_mask;
  for (uint32_t rb = b0; rb < b1; ++rb) {
    const uint64_t n = pp.blk_count[rb];
    const uint64_t *region = pp.recs + (uint64_t)rb * pp.blk_cap * R;
    for (uint64_t i0 = 0; i0 < n; i0 += (uint64_t)UNROLL * blockDim.x) {
      uint64_t idx[UNROLL], key[UNROLL], dst[UNROLL];
      bool side[UNROLL];
      int hs[UNROLL];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        idx[u] = i0 + (uint64_t)u * blockDim.x + threadIdx.x;
        key[u] = 0;
        side[u] = false;
        hs[u] = -1;
        if (idx[u] < n) {
          key[u] = region[idx[u] * R];
          if (has_mask) side[u] = (region[idx[u] * R + R - 1] >> 63) != 0;
        }
      }
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        if (idx[u] >= n) continue;
        if (hot_n && !side[u]) hs[u] = hot_slot(hot_lds, direct ? key[u] + 1 : key[u]);
        if (hs[u] >= 0) {
          uint64_t rw[MAX_RW];
          rec_load(region + idx[u] * R, R, rw);
          rec_apply(q, pp, rw, R, hent + (uint64_t)hs[u] * W, nullptr);
          continue;
        }
        uint32_t p = side[u] ? 0u
                             : (direct ? (uint32_t)(key[u] >> sbits)
                                       : (pbits ? hash_part(key[u], (uint32_t)nparts) : 0u));
        dst[u] = (uint64_t)atomicAdd((unsigned long long *)&cur[p], 1ull);
      }
#pragma unroll
      for (int u = 0; u < UNROLL; ++u)
        if (idx[u] < n && hs[u] < 0) rec_copy(pp.precs + dst[u] * R, region + idx[u] * R, R);
    }
  }
  if (!hot_n) return;
  __syncthreads();
  // merge the heavy hitters into the global table
  for (int sl = threadIdx.x; sl < HOT_SLOTS; sl += blockDim.x) {
    const uint64_t hk = hot_lds[sl];
    lds_u64 *e = hent + (uint64_t)sl * W;
    if (hk == 0 || e[q.count_word] == 0) continue;
    uint64_t acc[24];
    for (int w = 0; w < W; ++w) acc[w] = e[w];
    uint64_t *dst;
    if (direct) {
      dst = q.table + (hk - 1) * W;
    } else {
      acc[0] = hk;
      uint64_t g = glb_find(q, hk);
      if (g == ~0ull) continue;
      dst = q.table + g * W;
    }
    merge_entry_glb(q, dst, acc);
  }
}

__global__ void __launch_bounds__(1024) part_agg_kernel(const DQuery *__restrict__ qs, DPart pp, uint32_t nseg) {
  extern __shared__ uint64_t lds_raw[];
  lds_u64 *lds = (lds_u64 *)lds_raw;
  __shared__ uint64_t s_ins;
  const DQuery &q = qs[0];
  const int W = q.entry_words;
  const int R = pp.rec_words;
  const uint64_t LC = (uint64_t)pp.lds_entries;
  const bool direct = q.gmode == GM_DIRECT;
  const uint64_t smask = direct ? ((1ull << pp.sbits) - 1) : (LC - 1);
  const int has_mask = pp.has_mask;
  if (threadIdx.x == 0) s_ins = 0;
  uint32_t my_ins = 0;
  for (uint32_t sgi = blockIdx.x; sgi < nseg; sgi += gridDim.x) {
    const uint64_t *sg = pp.seg + 4 * (uint64_t)sgi;
    const uint64_t part = sg[0], start = sg[1], count = sg[2];
    const bool exclusive = sg[3] != 0;
    for (uint64_t i = threadIdx.x; i < LC * W; i += blockDim.x) lds[i] = 0;
    __syncthreads();
    const uint64_t *base = pp.precs + start * R;
    for (uint64_t i0 = 0; i0 < count; i0 += (uint64_t)UNROLL * blockDim.x) {
      uint64_t rw[UNROLL][MAX_RW];
      bool ok[UNROLL];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        uint64_t i = i0 + (uint64_t)u * blockDim.x + threadIdx.x;
        ok[u] = i < count;
        if (ok[u]) rec_load(base + i * R, R, rw[u]);
        else rw[u][0] = 0;
      }
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) {
        if (!ok[u]) continue;
        const uint64_t code = rw[u][0];
        uint64_t *gent = nullptr;
        lds_u64 *lent = nullptr;
        if (has_mask && (rw[u][R - 1] >> 63)) {
          gent = q.table + (q.gcap + ((rw[u][R - 1] >> 62) & 1)) * W;  // NULL / unrepresentable key
        } else if (direct) {
          lent = lds + (code & smask) * W;
        } else {
          uint32_t h = (uint32_t)mix64(code);
          for (int p = 0; p < LDS_HASH_PROBES; ++p) {
            uint64_t slot = (h + p) & smask;
            lds_u64 *e = lds + slot * W;
            uint64_t k = *e;
            if (k == 0) k = l_atomic_cas(e, 0, code);
            if (k == 0 || k == code) { lent = e; break; }
          }
          if (!lent) {
            uint64_t g = glb_find(q, code);
            if (g == ~0ull) continue;
            gent = q.table + g * W;
          }
        }
        rec_apply(q, pp, rw[u], R, lent, gent);
      }
    }
    __syncthreads();
    // write the segment's groups into the global table
    for (uint64_t slot = threadIdx.x; slot < LC; slot += blockDim.x) {
      lds_u64 *e = lds + slot * W;
      if (e[q.count_word] == 0) continue;
      uint64_t acc[24];
      for (int w = 0; w < W; ++w) acc[w] = e[w];
      uint64_t *dst;
      bool fresh;
      if (direct) {
        // DIRECT slice: global index = partition base + slot
        uint64_t idx = (part << pp.sbits) + slot;
        dst = q.table + idx * W;
        fresh = exclusive;
      } else {
        uint64_t g = glb_find_ins(q, acc[0], &fresh);
        if (g == ~0ull) continue;
        dst = q.table + g * W;
        if (fresh) ++my_ins;
        fresh = fresh && exclusive;
      }
      if (fresh) {
        for (int w = (direct ? 0 : 1); w < W; ++w) dst[w] = acc[w];
      } else {
        merge_entry_glb(q, dst, acc);
      }
    }
    __syncthreads();
  }
  if (!direct && my_ins) atomicAdd((unsigned long long *)&s_ins, (unsigned long long)my_ins);
  __syncthreads();
  if (!direct && threadIdx.x == 0 && s_ins) {
    uint64_t n = g_atomic_add(pp.inserted, s_ins) + s_ins;
    if (n > (q.gcap >> 1) + (q.gcap >> 2)) q.status[0] = 1;
  }
}

__global__ void part_starts_kernel(const uint64_t *part_off, uint32_t nparts, uint32_t grid, uint64_t total,
                                   uint64_t *starts) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nparts) starts[p] = part_off[(uint64_t)p * grid];
  if (p == nparts) starts[p] = total;
}

}  // namespace qm

using namespace qm;

// fast = NS * 4 + NK (predicate slots 1..4, key columns 1..2)
static const void *fast_fn(int fast) {
  switch (fast) {
    case 1 * 4 + 1: return (const void *)part_scan_fast_kernel<1, 1>;
    case 1 * 4 + 2: return (const void *)part_scan_fast_kernel<1, 2>;
    case 2 * 4 + 1: return (const void *)part_scan_fast_kernel<2, 1>;
    case 2 * 4 + 2: return (const void *)part_scan_fast_kernel<2, 2>;
    case 3 * 4 + 1: return (const void *)part_scan_fast_kernel<3, 1>;
    case 3 * 4 + 2: return (const void *)part_scan_fast_kernel<3, 2>;
    case 4 * 4 + 1: return (const void *)part_scan_fast_kernel<4, 1>;
    case 4 * 4 + 2: return (const void *)part_scan_fast_kernel<4, 2>;
  }
  return nullptr;
}

extern "C" hipError_t qmk_part_scan(const DScan *sc, const DQuery *d_q, const DPart *pp, uint32_t grid,
                                    uint64_t tpb, int fast, int prefetch, hipStream_t st) {
  size_t lds = (size_t)(pp->nparts + 1) * 4;
  const void *fn = fast ? fast_fn(fast) : (const void *)part_scan_kernel;
  if (!fn) return hipErrorInvalidValue;
  if (lds > 65536) {
    hipError_t e = qmk_func_lds(fn, (size_t)lds);
    if (e != hipSuccess) return e;
  }
  if (fast) {
    void *args[] = {(void *)sc, (void *)&d_q, (void *)pp, (void *)&tpb, (void *)&prefetch};
    return hipLaunchKernel(fn, dim3(grid), dim3(PART_BLOCK), args, lds, st);
  }
  hipLaunchKernelGGL(part_scan_kernel, dim3(grid), dim3(PART_BLOCK), lds, st, *sc, d_q, *pp, tpb);
  return hipGetLastError();
}

extern "C" hipError_t qmk_part_scatter(const DQuery *d_q, const DPart *pp, uint32_t g1, uint32_t rpb,
                                       int entry_words, hipStream_t st) {
  size_t lds = (size_t)(pp->nparts + 1) * 8 + (pp->hot_n ? (size_t)HOT_SLOTS * entry_words * 8 : 0);
  if (lds > 65536) {
    hipError_t e = qmk_func_lds((const void *)part_scatter_kernel, (size_t)lds);
    if (e != hipSuccess) return e;
  }
  uint32_t g2 = (g1 + rpb - 1) / rpb;
  hipLaunchKernelGGL(part_scatter_kernel, dim3(g2), dim3(1024), lds, st, d_q, *pp, g1, rpb);
  return hipGetLastError();
}

extern "C" hipError_t qmk_part_starts(const uint64_t *part_off, uint32_t nparts, uint32_t grid, uint64_t total,
                                      uint64_t *starts, hipStream_t st) {
  hipLaunchKernelGGL(part_starts_kernel, dim3((nparts + 1 + 255) / 256), dim3(256), 0, st, part_off, nparts, grid,
                     total, starts);
  return hipGetLastError();
}

extern "C" hipError_t qmk_part_agg(const DQuery *d_q, const DPart *pp, uint32_t nseg, uint32_t grid, int entry_words,
                                   hipStream_t st) {
  size_t lds = (size_t)pp->lds_entries * entry_words * 8;
  if (lds > 65536) {
    hipError_t e = qmk_func_lds((const void *)part_agg_kernel, (size_t)lds);
    if (e != hipSuccess) return e;
  }
  if (nseg == 0) return hipSuccess;
  if (grid > nseg) grid = nseg;
  hipLaunchKernelGGL(part_agg_kernel, dim3(grid), dim3(1024), lds, st, d_q, *pp, nseg);
  return hipGetLastError();
}

extern "C" int qmk_part_scan_blocks_per_cu(int fast, size_t lds) {
  int n = 0;
  const void *fn = fast ? fast_fn(fast) : (const void *)part_scan_kernel;
  if (!fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, PART_BLOCK, lds) != hipSuccess) return 2;
  return n > 0 ? n : 1;
}

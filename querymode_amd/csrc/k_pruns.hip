// k_pruns.hip — many aggregate queries grouped by one integer key stored in
// any row order: the batched high-cardinality group-by.
//
// Every query of a launch filters by (at most) one range on one common column
// and shares one aggregate list (COUNT(*), then additive states — COUNT(col),
// exact integer SUM — then MIN / MAX order codes). The launch's sorted
// distinct range bounds cut that column into buckets (bucket(x) =
// #{bounds <= x}, NULL: nbounds + 1); query q selects the bucket range
// [e_lo, e_hi). Per batch, the table is read once:
//
//   pass 1 (pr_scatter_kernel<0>): each row's bucket (binary search over the
//     bounds in LDS); rows in buckets no query covers are dropped; the others
//     become records — word 0 = the key code's bits below the level-1 digit
//     | bucket << 16, then the raw words of the state columns (and a NULL-bit
//     word) — counting-sorted in LDS by the level-1 digit (key code >> sh1)
//     and written as contiguous runs into level-1 regions (one atomic
//     reservation per digit and tile);
//   pass 2 (pr_scatter_kernel<1>): the same over each level-1 region, by the
//     next key bits, into final partitions of 2^s consecutive keys;
//   count (pr_count_kernel): per final partition, a (key, bucket) bitmap in
//     LDS; lanes = keys, one query per wave step: the groups of every
//     (query, partition) — their exclusive scan (query-major) places every
//     group before anything is written;
//   write (pr_runs_kernel): per final partition, a counting sort of its
//     records by (key, bucket) in LDS, inclusive prefix sums of the additive
//     states over that order; then lanes = keys, one query per wave step: the
//     query's rows of key k are the sorted span [start(k, e_lo),
//     start(k, e_hi)), so COUNT / SUM are differences of two entries and
//     MIN / MAX the max over the span. A wave writes the non-empty groups of
//     64 consecutive keys of one query as consecutive result rows.
// Results come out per query in ascending key order, independent of the
// table's row order and of scheduling.
//
// Aggregate semantics are those of src/partial-agg.ts:38-134 (COUNT, exact
// integer SUM, MIN / MAX by value, NaN never updates MIN / MAX); filters
// those of src/decode.ts:477-578 (NULL never matches) through the bucket
// ranges the host derives from each query's coerced range.
#include <algorithm>
#include "k_common.hip.h"
#include "qm_kernels.h"

namespace qm {

namespace {

static constexpr int PR_WAVES = PR_BLOCK / 64;
// write pass: 1024 threads (16 waves per CU to hide the LDS latency of the
// per-group chains), 4 records per thread
static constexpr int PR_RBLOCK = 1024;
static constexpr int PR_RWAVES = PR_RBLOCK / 64;
static constexpr int PR_RIPT = 4;  // records per thread in the write pass (cap <= PR_RIPT * PR_RBLOCK)
static constexpr int PR_GPT = 4;   // groups a thread holds while the write pass stores the outputs

__device__ __forceinline__ bool pr_null(const uint8_t *nb, uint64_t row) {
  return nb != nullptr && ((gld<uint8_t>(nb, row >> 3) >> (row & 7)) & 1);
}
__device__ __forceinline__ bool pr_wide(int32_t dt) { return dt == DT_I64 || dt == DT_F64; }

// value bits as stored in a record: 64-bit types raw, the others in one word
// (integers sign-extended to 32 bits, f32 bits, bool 0 / 1)
__device__ __forceinline__ uint64_t pr_bits(const void *p, int32_t dt, uint64_t row) {
  switch (dt) {
    case DT_I8: return (uint32_t)(int32_t)gld<int8_t>(p, row);
    case DT_BOOL: return gld<uint8_t>(p, row);
    case DT_I16: return (uint32_t)(int32_t)gld<int16_t>(p, row * 2);
    case DT_I32: case DT_F32: return gld<uint32_t>(p, row * 4);
    default: return gld<uint64_t>(p, row * 8);
  }
}
__device__ __forceinline__ int64_t pr_ival(uint64_t w, int32_t dt) {
  return dt == DT_I64 ? (int64_t)w : (int64_t)(int32_t)(uint32_t)w;
}
__device__ __forceinline__ double pr_fval(uint64_t w, int32_t dt) {
  return dt == DT_F32 ? (double)__uint_as_float((uint32_t)w) : bits_f64(w);
}
__device__ __forceinline__ uint64_t pr_code(uint64_t w, int32_t dt) {
  return dt_is_float(dt) ? code_f64_bits(f64_bits(pr_fval(w, dt))) : code_i64(pr_ival(w, dt));
}

// bucket of a filter order code (bounds padded to 2^bsteps - 1 entries)
__device__ __forceinline__ uint32_t pr_bucket(const uint64_t *sb, int bsteps, uint32_t nb, uint64_t x) {
  uint32_t f = 0;
  for (int step = 1 << (bsteps - 1); step > 0; step >>= 1)
    if (sb[f + step - 1] <= x) f += step;
  return f < nb ? f : nb;
}

__device__ __forceinline__ uint64_t u64max(uint64_t a, uint64_t b) { return a > b ? a : b; }
// a value every lane holds the same copy of, as a scalar
__device__ __forceinline__ uint64_t rfl64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// block-wide exclusive scan of one u32 per thread (PR_BLOCK threads)
__device__ __forceinline__ uint32_t pr_block_exscan(uint32_t v, uint32_t *sw /*[PR_WAVES + 1]*/, uint32_t *total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) sw[wave] = x;
  __syncthreads();
  uint32_t base = 0, all = 0;
#pragma unroll
  for (int w = 0; w < PR_WAVES; ++w) {
    const uint32_t c = sw[w];
    base += w < wave ? c : 0u;
    all += c;
  }
  *total = all;
  __syncthreads();
  return base + x - v;
}
__device__ __forceinline__ uint64_t pr_block_exscan64(uint64_t v, uint64_t *sw /*[PR_WAVES]*/) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint64_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) sw[wave] = x;
  __syncthreads();
  uint64_t base = 0;
#pragma unroll
  for (int w = 0; w < PR_WAVES; ++w) base += w < wave ? sw[w] : 0ull;
  __syncthreads();
  return base + x - v;
}

template <int WAVES>
__device__ __forceinline__ uint32_t pr_block_exscan_n(uint32_t v, uint32_t *sw, uint32_t *total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) sw[wave] = x;
  __syncthreads();
  uint32_t base = 0, all = 0;
#pragma unroll
  for (int w = 0; w < WAVES; ++w) {
    const uint32_t c = sw[w];
    base += w < wave ? c : 0u;
    all += c;
  }
  *total = all;
  __syncthreads();
  return base + x - v;
}
template <int WAVES>
__device__ __forceinline__ uint64_t pr_block_exscan64_n(uint64_t v, uint64_t *sw) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint64_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) sw[wave] = x;
  __syncthreads();
  uint64_t base = 0;
#pragma unroll
  for (int w = 0; w < WAVES; ++w) base += w < wave ? sw[w] : 0ull;
  __syncthreads();
  return base + x - v;
}

}  // namespace

// ---------------------------------------------------------------------------
// passes 1 and 2: scatter by a key digit through an LDS-staged counting sort
// ---------------------------------------------------------------------------
template <int SRC, int RW>
// (at most 128 VGPRs: two workgroups per CU, which their LDS allows, need four waves per SIMD)
__global__ void __launch_bounds__(PR_BLOCK, 4) pr_scatter_kernel(PRScan ps) {
  constexpr int IPT = PR_STAGE_WORDS / RW / PR_BLOCK;
  constexpr uint32_t TILE = (uint32_t)IPT * PR_BLOCK;
  constexpr int NVX = (RW - 1) < PR_MAX_V ? (RW - 1) : PR_MAX_V;  // value columns a record of RW words can hold
  __shared__ __align__(16) uint32_t s_stage[TILE * RW];
  __shared__ uint16_t s_dig[TILE];
  __shared__ uint32_t s_hist[PR_MAX_P];
  __shared__ uint32_t s_lim[PR_MAX_P];
  __shared__ uint64_t s_gofs[PR_MAX_P];
  __shared__ uint64_t s_bounds[SRC == 0 ? R_MAX_BUCKETS : 1];
  __shared__ uint16_t s_bslot[SRC == 0 ? PR_BSLOTS + 1 : 1];
  __shared__ uint32_t s_chunk0[SRC == 1 ? PR_MAX_P + 1 : 1];  // first unit of each level-1 region
  __shared__ uint32_t s_brow[SRC == 0 ? R_MAX_BUCKETS : 1];
  __shared__ uint32_t s_sw[PR_WAVES + 1];
  const int t = threadIdx.x;
  const uint32_t ndig = SRC == 0 ? ps.np1 : (1u << (ps.sh1 - ps.s));
  const uint32_t dpt = (ndig + PR_BLOCK - 1) / PR_BLOCK;  // digits per thread in the scan (<= 2)
  const uint32_t nb = (uint32_t)ps.nbounds;
  if constexpr (SRC == 0) {
    const int nbp = (1 << ps.bsteps) - 1;
    for (int i = t; i < nbp; i += PR_BLOCK) s_bounds[i] = ps.bounds[i];
    if (ps.bslot)
      for (uint32_t i = t; i <= ps.bslots; i += PR_BLOCK) s_bslot[i] = ps.bslot[i];
    for (int i = t; i < R_MAX_BUCKETS; i += PR_BLOCK) s_brow[i] = 0;
  }
  if constexpr (SRC == 1)
    for (uint32_t i = t; i <= ps.np1; i += PR_BLOCK) s_chunk0[i] = ps.chunk0[i];
  const uint64_t nunits = SRC == 0 ? (ps.nrows + TILE - 1) / TILE : (uint64_t)ps.chunk0[ps.np1];
  const uint32_t kmask1 = (1u << ps.sh1) - 1u;
  for (uint64_t u = blockIdx.x; u < nunits; u += gridDim.x) {
    for (uint32_t d = t; d < ndig; d += PR_BLOCK) s_hist[d] = 0;
    __syncthreads();
    // the unit's items: rows [src0, src0 + srcn) or records of one level-1 region
    uint64_t src0 = 0, srcn = 0;
    uint32_t p1 = 0;
    if constexpr (SRC == 0) {
      src0 = u * TILE;
      srcn = ps.nrows - src0 < TILE ? ps.nrows - src0 : TILE;
    } else {
      uint32_t lo = 0, hi = ps.np1;  // last p1 with chunk0[p1] <= u (in LDS)
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_chunk0[mid] <= u) lo = mid; else hi = mid;
      }
      p1 = lo;
      const uint64_t ci = u - s_chunk0[p1];
      const uint64_t cap1 = ps.start1[p1 + 1] - ps.start1[p1];
      const uint64_t n1 = ps.cur1[p1] < cap1 ? ps.cur1[p1] : cap1;
      src0 = ps.start1[p1] + ci * TILE;
      srcn = ci * TILE < n1 ? (n1 - ci * TILE < TILE ? n1 - ci * TILE : TILE) : 0;
    }
    uint32_t kb[IPT], dg[IPT], rk[IPT], nbits[IPT];
    uint64_t v[IPT][NVX > 0 ? NVX : 1];
    uint32_t w[IPT][RW];
    if constexpr (SRC == 0) {
      // every load of the thread's rows first (the rows are independent), then
      // the buckets: no load waits for another row's bucket search
      uint64_t kraw[IPT], praw[IPT];
      uint32_t live = 0, pnul = 0;
#pragma unroll
      for (int i = 0; i < IPT; ++i) {
        const uint64_t it = (uint64_t)t + (uint64_t)i * PR_BLOCK;
        rk[i] = ~0u;
        dg[i] = 0;
        kb[i] = 0;
        nbits[i] = 0;
        kraw[i] = 0;
        praw[i] = 0;
#pragma unroll
        for (int j = 0; j < NVX; ++j) v[i][j] = 0;
        if (it >= srcn) continue;
        const uint64_t row = src0 + it;
        if (page_bit(ps.page_skip, row)) continue;  // no query can match a row of this page
        live |= 1u << i;
        DCol kc;
        kc.data = ps.key; kc.nulls = nullptr; kc.dtype = ps.key_dt; kc.side = 0;
        kraw[i] = (uint64_t)load1_i(kc, row);
        praw[i] = pr_bits(ps.pcol, ps.p_dt, row);
        if (pr_null(ps.pnull, row)) pnul |= 1u << i;
#pragma unroll
        for (int j = 0; j < NVX; ++j) {
          if (j < ps.nv) {
            v[i][j] = pr_bits(ps.vcol[j], ps.v_dt[j], row);
            if (pr_null(ps.vnull[j], row)) nbits[i] |= 1u << j;
          }
        }
      }
#pragma unroll
      for (int i = 0; i < IPT; ++i) {
        if (!((live >> i) & 1u)) continue;
        const uint64_t code = kraw[i] - (uint64_t)ps.kmin;
        uint32_t e;
        if ((pnul >> i) & 1u) e = nb + 1;
        else {
          const uint64_t x = pr_code(praw[i], ps.p_dt);
          e = ps.bslot ? bucket_slot(s_bounds, s_bslot, ps.bslot_lo, ps.bslot_shift, ps.bslots, nb, x)
                       : pr_bucket(s_bounds, ps.bsteps, nb, x);
        }
        if (!((ps.bucket_on[e >> 5] >> (e & 31)) & 1u)) continue;
        const uint64_t d = code >> ps.sh1;
        if (d >= ndig) { atomicOr(ps.status, 1u); continue; }
        dg[i] = (uint32_t)d;
        kb[i] = ((uint32_t)code & kmask1) | (e << 16);
        atomicAdd(&s_brow[e], 1u);
        rk[i] = atomicAdd(&s_hist[dg[i]], 1u);
      }
    } else {
      // every record of the thread is loaded before any is counted, so the
      // loads are in flight together
#pragma unroll
      for (int i = 0; i < IPT; ++i) {
        const uint64_t it = (uint64_t)t + (uint64_t)i * PR_BLOCK;
        if (it >= srcn) continue;
        const uint32_t *r = ps.rec1 + (src0 + it) * RW;
        if constexpr (RW % 4 == 0) {  // records of 4 or 8 words start 16-byte aligned
#pragma unroll
          for (int k = 0; k < RW; k += 4) {
            const uint4 x = *reinterpret_cast<const uint4 *>(r + k);
            w[i][k] = x.x; w[i][k + 1] = x.y; w[i][k + 2] = x.z; w[i][k + 3] = x.w;
          }
        } else {
#pragma unroll
          for (int k = 0; k < RW; ++k) w[i][k] = r[k];
        }
      }
#pragma unroll
      for (int i = 0; i < IPT; ++i) {
        const uint64_t it = (uint64_t)t + (uint64_t)i * PR_BLOCK;
        rk[i] = ~0u;
        dg[i] = 0;
        kb[i] = 0;
        nbits[i] = 0;
        if (it >= srcn) continue;
        dg[i] = (w[i][0] & 0xFFFFu) >> ps.s;
        rk[i] = atomicAdd(&s_hist[dg[i]], 1u);
      }
    }
    __syncthreads();
    // digit starts in the tile, and one reservation per digit in its region
    constexpr int DPT = PR_MAX_P / PR_BLOCK;
    uint32_t cnt[DPT], mine = 0;
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
      cnt[k] = 0;
      const uint32_t d = (uint32_t)t * dpt + k;
      if (k < (int)dpt && d < ndig) cnt[k] = s_hist[d];
      mine += cnt[k];
    }
    uint32_t total;
    uint32_t run = pr_block_exscan(mine, s_sw, &total);
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
      const uint32_t d = (uint32_t)t * dpt + k;
      if (k >= (int)dpt || d >= ndig) continue;
      s_hist[d] = run;
      uint32_t lim = run;
      if (cnt[k]) {
        const uint32_t dst = SRC == 0 ? d : ((p1 << (ps.sh1 - ps.s)) | d);
        const uint32_t nreg = SRC == 0 ? ps.np1 : ps.np;
        uint64_t allowed = 0, g = 0, st = 0;
        if (dst < nreg) {
          const uint64_t *start = SRC == 0 ? ps.start1 : ps.start2;
          uint32_t *cur = SRC == 0 ? ps.cur1 : ps.cur2;
          st = start[dst];
          const uint64_t capr = start[dst + 1] - st;
          g = atomicAdd(&cur[dst], cnt[k]);
          allowed = g >= capr ? 0 : (capr - g < cnt[k] ? capr - g : cnt[k]);
        }
        if (allowed < cnt[k]) atomicOr(ps.status, 1u);
        s_gofs[d] = st + g - run;
        lim = run + (uint32_t)allowed;
      }
      s_lim[d] = lim;
      run += cnt[k];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      if (rk[i] == ~0u) continue;
      const uint32_t slot = s_hist[dg[i]] + rk[i];
      s_dig[slot] = (uint16_t)dg[i];
      uint32_t *o = s_stage + slot * RW;
      if constexpr (SRC == 0) {
        o[0] = kb[i];
#pragma unroll
        for (int j = 0; j < NVX; ++j) {
          if (j >= ps.nv) break;
          o[ps.v_word[j]] = (uint32_t)v[i][j];
          if (pr_wide(ps.v_dt[j])) o[ps.v_word[j] + 1] = (uint32_t)(v[i][j] >> 32);
        }
        if (ps.null_word >= 0) o[ps.null_word] = nbits[i];
      } else {
#pragma unroll
        for (int k = 0; k < RW; ++k) o[k] = w[i][k];
      }
    }
    __syncthreads();
    uint32_t *dstrec = SRC == 0 ? ps.rec1 : ps.rec2;
    for (uint32_t j = t; j < total; j += PR_BLOCK) {
      const uint32_t d = s_dig[j];
      if (j >= s_lim[d]) continue;
      uint32_t *o = dstrec + (s_gofs[d] + j) * RW;
      const uint32_t *in = s_stage + j * RW;
      if constexpr (RW % 4 == 0) {
#pragma unroll
        for (int k = 0; k < RW; k += 4) *reinterpret_cast<uint4 *>(o + k) = *reinterpret_cast<const uint4 *>(in + k);
      } else if constexpr (RW % 2 == 0) {
#pragma unroll
        for (int k = 0; k < RW; k += 2) *reinterpret_cast<uint2 *>(o + k) = *reinterpret_cast<const uint2 *>(in + k);
      } else {
#pragma unroll
        for (int k = 0; k < RW; ++k) o[k] = in[k];
      }
    }
    __syncthreads();
  }
  if constexpr (SRC == 0) {
    for (int i = t; i < R_MAX_BUCKETS; i += PR_BLOCK)
      if (s_brow[i]) atomicAdd((unsigned long long *)&ps.bucket_rows[i], (unsigned long long)s_brow[i]);
  }
}

// ---------------------------------------------------------------------------
// count pass: groups of every (query, final partition)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(PR_BLOCK) pr_count_kernel(PRScan ps) {
  constexpr int BW = R_MAX_BUCKETS / 32 + 1;  // bitmap words per key (+1: no bank conflicts across keys)
  __shared__ uint32_t s_bm[64 * BW];
  __shared__ uint32_t s_qe[R_MAX_QUERIES];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const uint32_t P = blockIdx.x;
  if (t < ps.nq) s_qe[t] = ps.qe[t];
  const uint64_t base = ps.start2[P];
  const uint64_t capr = ps.start2[P + 1] - base;
  const uint64_t n = ps.cur2[P] < capr ? ps.cur2[P] : capr;
  for (int i = t; i < 64 * BW; i += PR_BLOCK) s_bm[i] = 0;
  __syncthreads();
  const uint32_t km = (1u << ps.s) - 1u;
  // a thread's record words are all loaded before any of them is applied, so
  // its loads are in flight together (a partition is at most a few rounds)
  constexpr int CPT = 8;
  for (uint64_t j0 = t; j0 < n; j0 += (uint64_t)CPT * PR_BLOCK) {
    uint32_t kb[CPT];
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
      const uint64_t j = j0 + (uint64_t)i * PR_BLOCK;
      kb[i] = j < n ? gld<uint32_t>(ps.rec2, (base + j) * (uint64_t)ps.rw * 4) : 0u;
    }
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
      if (j0 + (uint64_t)i * PR_BLOCK >= n) break;
      const uint32_t e = kb[i] >> 16;
      atomicOr(&s_bm[(kb[i] & km) * BW + (e >> 5)], 1u << (e & 31));
    }
  }
  __syncthreads();
  const bool kon = (uint32_t)lane <= km;
  const uint32_t *bm = s_bm + lane * BW;
  for (int q = wave; q < ps.nq; q += PR_WAVES) {
    const uint32_t qe = s_qe[q];
    const uint32_t elo = qe & 0xFFFFu, ehi = qe >> 16;
    bool any = false;
    if (elo < ehi) {
      for (uint32_t wd = elo >> 5; wd <= (ehi - 1) >> 5; ++wd) {
        const uint32_t lo = elo > wd * 32 ? elo - wd * 32 : 0u;
        const uint32_t hi = ehi - wd * 32 < 32 ? ehi - wd * 32 : 32u;
        const uint32_t m = (hi == 32 ? ~0u : ((1u << hi) - 1u)) & ~((1u << lo) - 1u);
        any = any || (bm[wd] & m) != 0;
      }
    }
    const uint64_t b = __ballot(kon && any);
    if (lane == 0) ps.gcount[(uint64_t)q * ps.np + P] = (uint64_t)__builtin_popcountll(b);
  }
}

// ---------------------------------------------------------------------------
// write pass
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t pr_start(const uint32_t *bins, uint32_t x) {
  const uint32_t w = bins[x >> 1];
  return (x & 1) ? (w >> 16) : (w & 0xFFFFu);
}

// log2 of the byte width an output is stored with
__device__ __forceinline__ int pr_out_lw(const ROut &ro) {
  const int32_t dt = ro.out_dtype;
  const int iw = dt == DT_I64 ? 3 : (dt == DT_I32 ? 2 : (dt == DT_I16 ? 1 : 0));
  if (ro.kind == OK_KEY) return iw;
  if (ro.kind == OK_MIN || ro.kind == OK_MAX) return ro.src_is_float ? (dt == DT_F32 ? 2 : 3) : iw;
  return 3;
}

// one record's words (rw uniform; NW = the array length, rw <= NW): one
// vector load when the record is 2, 4, 6 or 8 words and the buffer is
// 32-byte aligned, else a load per word
template <int NW>
__device__ __forceinline__ void pr_load_rec(const uint32_t *rec, uint64_t idx, uint32_t rw, bool aligned,
                                            uint32_t (&w)[NW]) {
  const uint32_t *r = rec + idx * (uint64_t)rw;
#pragma unroll
  for (int k = 0; k < NW; ++k) w[k] = 0;
  if constexpr (NW == 4) {
    if (aligned) {
      const uint4 x = *reinterpret_cast<const uint4 *>(r);
      w[0] = x.x; w[1] = x.y; w[2] = x.z; w[3] = x.w;
      return;
    }
  } else {
    if (aligned && rw == 4) {
      const uint4 x = *reinterpret_cast<const uint4 *>(r);
      w[0] = x.x; w[1] = x.y; w[2] = x.z; w[3] = x.w;
      return;
    } else if (aligned && rw == 8) {
      const uint4 x = reinterpret_cast<const uint4 *>(r)[0], y = reinterpret_cast<const uint4 *>(r)[1];
      w[0] = x.x; w[1] = x.y; w[2] = x.z; w[3] = x.w;
      w[4] = y.x; w[5] = y.y; w[6] = y.z; w[7] = y.w;
      return;
    } else if (aligned && rw == 2) {
      const uint2 x = *reinterpret_cast<const uint2 *>(r);
      w[0] = x.x; w[1] = x.y;
      return;
    } else if (aligned && rw == 6) {
      const uint2 x = reinterpret_cast<const uint2 *>(r)[0], y = reinterpret_cast<const uint2 *>(r)[1],
                  z = reinterpret_cast<const uint2 *>(r)[2];
      w[0] = x.x; w[1] = x.y; w[2] = y.x; w[3] = y.y; w[4] = z.x; w[5] = z.y;
      return;
    }
  }
#pragma unroll
  for (int k = 0; k < NW; ++k)
    if ((uint32_t)k < rw) w[k] = r[k];
}

// contribution of one record (raw words w) to non-COUNT(*) state 1 + s2
template <int NW>
__device__ __forceinline__ uint64_t pr_contrib(const PRScan &ps, int s2, const uint32_t (&w)[NW]) {
  const DState &st = ps.st[1 + s2];
  const int vj = ps.st_v[1 + s2];
  if (vj < 0) return st.kind == SK_COUNT ? 1ull : 0ull;  // COUNT(*)
  uint32_t nbits = 0, lo = 0, hi = 0;
  // words by runtime index, selected over the unrolled record (no dynamic register indexing)
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    nbits = k == ps.null_word ? w[k] : nbits;
    lo = k == ps.v_word[vj] ? w[k] : lo;
    hi = k == ps.v_word[vj] + 1 ? w[k] : hi;
  }
  if ((nbits >> vj) & 1u) return 0;
  const int32_t dt = ps.v_dt[vj];
  const uint64_t bits = pr_wide(dt) ? ((uint64_t)lo | ((uint64_t)hi << 32)) : (uint64_t)lo;
  if (st.kind == SK_COUNT) return 1;
  if (st.kind == SK_SUM_I64) return (uint64_t)pr_ival(bits, dt);
  // SK_MAX_CODE: MAX keeps code, MIN ~code; NaN never updates either
  if (dt_is_float(dt)) {
    const double fv = pr_fval(bits, dt);
    if (fv != fv) return 0;
    const uint64_t code = code_f64_bits(f64_bits(fv));
    return st.is_min ? ~code : code;
  }
  const uint64_t code = code_i64(pr_ival(bits, dt));
  return st.is_min ? ~code : code;
}

// An order-code state over an integer column of at most 32 bits keeps 32-bit
// codes in LDS: (uint32_t)v ^ 2^31 (MIN: its complement), the low word of the
// 64-bit code with the top bit flipped; 0 (no value) stays 0. A real code can
// be 0 too (MIN of INT32_MAX, MAX of INT32_MIN), but a group's max is only
// read when it counted a value, and then the real codes' max is >= every 0.
__device__ __forceinline__ bool pr_narrow(const PRScan &ps, int st) {
  const int vj = ps.st_v[st];
  if (vj < 0) return false;
  const int32_t dt = ps.v_dt[vj];
  return !dt_is_float(dt) && dt != DT_I64;
}
__device__ __forceinline__ uint32_t pr_code32(uint64_t c) { return c ? ((uint32_t)c ^ 0x80000000u) : 0u; }

// A workgroup takes a contiguous range of final partitions, so the rows it
// writes for one query are consecutive across its partitions (each result
// column is written as long runs, not as short pieces from many workgroups).
// RW: the record width in words when fixed at compile time (4), else 0 (any
// width up to PR_MAX_RW, read from the launch).
template <int NSUM, int NMAX, int RW>
__global__ void __launch_bounds__(PR_RBLOCK) pr_runs_kernel(PRScan ps) {
  constexpr int NW = RW > 0 ? RW : PR_MAX_RW;
  constexpr int NSA = NSUM > 0 ? NSUM : 1, NMA = NMAX > 0 ? NMAX : 1;
  extern __shared__ __align__(16) uint8_t s_dyn[];
  __shared__ uint32_t s_sw[PR_RWAVES + 1];
  __shared__ uint64_t s_sw64[PR_RWAVES];
  __shared__ uint32_t s_qe[R_MAX_QUERIES];
  __shared__ uint64_t s_qrow[R_MAX_QUERIES];  // first result row of (query, loaded partition)
  __shared__ uint32_t s_qcnt[R_MAX_QUERIES], s_qbase[R_MAX_QUERIES];  // groups of (query, partition), flat start
  __shared__ uint8_t s_gk[R_MAX_QUERIES * 64];  // keys (partition lanes) of each query's groups, ascending
  __shared__ uint8_t s_gq[R_MAX_QUERIES * 64];  // query of each flat group
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const uint32_t per = (ps.np + gridDim.x - 1) / gridDim.x;
  const uint32_t P0 = blockIdx.x * per;
  const uint32_t P1 = P0 + per < ps.np ? P0 + per : ps.np;
  if (P0 >= P1) return;
  const uint32_t NBK = (uint32_t)ps.nbounds + 2;
  const uint32_t nbin = (NBK << ps.s);
  const uint32_t nbw = nbin / 2 + 1;              // packed u16 entries 0 .. nbin
  const uint32_t cap = ps.cap;
  const uint32_t rw = RW > 0 ? (uint32_t)RW : (uint32_t)ps.rw;
  const bool raligned = (reinterpret_cast<uintptr_t>(ps.rec2) & 31u) == 0;
  uint32_t *s_bin = reinterpret_cast<uint32_t *>(s_dyn);
  uint64_t *s_add = reinterpret_cast<uint64_t *>(s_dyn + ((nbw * 4 + 15) & ~15u));
  uint64_t *s_max = s_add + (uint64_t)NSUM * cap;
  if (t < ps.nq) s_qe[t] = ps.qe[t];
  // output descriptors in LDS (read back as uniform values per output)
  __shared__ uint64_t s_out[R_MAX_OUTS][3];  // ROut bits, column base, NULL bitmap base
  const int nouts = ps.nouts;
  if (t < R_MAX_OUTS) {
    ROut r = ps.outs[t < nouts ? t : 0];
    uint64_t rb = 0;
    memcpy(&rb, &r, sizeof r);
    s_out[t][0] = rb;
    s_out[t][1] = reinterpret_cast<uint64_t>(ps.out_base[t < nouts ? t : 0]);
    s_out[t][2] = reinterpret_cast<uint64_t>(ps.out_nbase[t < nouts ? t : 0]);
  }
  const uint32_t km = (1u << ps.s) - 1u;
  const bool kon = (uint32_t)lane <= km;
  const uint64_t lt = (1ull << lane) - 1ull;
  const uint32_t kbin = (uint32_t)lane * NBK;
  uint32_t narm = 0;  // bit j: order-code state NSUM + 1 + j keeps 32-bit codes
#pragma unroll
  for (int j = 0; j < NMAX; ++j) narm |= pr_narrow(ps, NSUM + 1 + j) ? 1u << j : 0u;

  // ---- load: partition P's records -> (key, bucket) bin and rank within it
  // (LDS counters), state contributions (registers); the first result row of
  // (query t, P) for t < nq
  uint32_t nP = 0;
  uint32_t bn[PR_RIPT], rk[PR_RIPT];
  uint64_t ca[PR_RIPT][NSA], cm[PR_RIPT][NMA];
  auto stage_load = [&](uint32_t P) {
    const uint64_t pbase = ps.start2[P];
    const uint64_t capr = ps.start2[P + 1] - pbase;
    const uint32_t c2 = ps.cur2[P];
    nP = (uint32_t)(c2 < capr ? c2 : capr);
    for (uint32_t i = t; i < nbw; i += PR_RBLOCK) s_bin[i] = 0;
    if (nP > cap && t == 0) atomicOr(ps.status, 2u);  // the host sizes partitions from the key histogram
    __syncthreads();
    // every record of the thread is loaded before any is applied, so the
    // loads are in flight together
    uint32_t wr[PR_RIPT][NW];
#pragma unroll
    for (int i = 0; i < PR_RIPT; ++i) {
      const uint32_t j = (uint32_t)t + (uint32_t)i * PR_RBLOCK;
      if (j < nP && j < cap) pr_load_rec(ps.rec2, pbase + j, rw, raligned, wr[i]);
    }
    // the first result row of (query t, P), read while the records load (the
    // previous partition's groups are all written: it is read after the
    // barriers of the sort)
    if (t < ps.nq) {
      const uint64_t *gq = ps.gcount + (uint64_t)t * ps.np;
      s_qrow[t] = ps.row_off[t] + (gq[P] - gq[0]);
    }
#pragma unroll
    for (int i = 0; i < PR_RIPT; ++i) {
      const uint32_t j = (uint32_t)t + (uint32_t)i * PR_RBLOCK;
      rk[i] = ~0u;
      bn[i] = 0;
#pragma unroll
      for (int k = 0; k < NSA; ++k) ca[i][k] = 0;
#pragma unroll
      for (int k = 0; k < NMA; ++k) cm[i][k] = 0;
      if (j >= nP || j >= cap) continue;
#pragma unroll
      for (int s2 = 0; s2 < NSUM; ++s2) ca[i][s2] = pr_contrib(ps, s2, wr[i]);
#pragma unroll
      for (int s2 = 0; s2 < NMAX; ++s2) cm[i][s2] = pr_contrib(ps, NSUM + s2, wr[i]);
      const uint32_t kb = wr[i][0];
      bn[i] = (kb & km) * NBK + (kb >> 16);
      const uint32_t sh = (bn[i] & 1u) * 16u;
      rk[i] = (atomicAdd(&s_bin[bn[i] >> 1], 1u << sh) >> sh) & 0xFFFFu;
    }
    __syncthreads();
  };

  // ---- sort: the loaded partition's contributions in (key, bucket) order,
  // prefix sums of the additive states, and its non-empty (query, key) groups
  uint32_t ngroups = 0;
  auto stage_sort = [&]() {
    // exclusive scan of the bins (packed pairs; thread t owns a word range)
    {
      const uint32_t wpt = (nbw + PR_RBLOCK - 1) / PR_RBLOCK;
      const uint32_t w0 = (uint32_t)t * wpt < nbw ? (uint32_t)t * wpt : nbw;
      const uint32_t w1 = w0 + wpt < nbw ? w0 + wpt : nbw;
      uint32_t sum = 0;
      for (uint32_t w = w0; w < w1; ++w) { const uint32_t x = s_bin[w]; sum += (x & 0xFFFFu) + (x >> 16); }
      uint32_t total;
      uint32_t run = pr_block_exscan_n<PR_RWAVES>(sum, s_sw, &total);
      for (uint32_t w = w0; w < w1; ++w) {
        const uint32_t x = s_bin[w];
        const uint32_t lo = run;
        run += x & 0xFFFFu;
        const uint32_t hi = run;
        run += x >> 16;
        s_bin[w] = lo | (hi << 16);
      }
    }
    __syncthreads();
    // scatter the contributions into (key, bucket) order
#pragma unroll
    for (int i = 0; i < PR_RIPT; ++i) {
      if (rk[i] == ~0u) continue;
      const uint32_t slot = pr_start(s_bin, bn[i]) + rk[i];
#pragma unroll
      for (int j = 0; j < NSUM; ++j) s_add[(uint64_t)j * cap + slot] = ca[i][j];
#pragma unroll
      for (int j = 0; j < NMAX; ++j) {
        if ((narm >> j) & 1u) reinterpret_cast<uint32_t *>(s_max + (uint64_t)j * cap)[slot] = pr_code32(cm[i][j]);
        else s_max[(uint64_t)j * cap + slot] = cm[i][j];
      }
    }
    __syncthreads();
    // inclusive prefix sums of the additive states over the sorted order
    const uint32_t nn = nP < cap ? nP : cap;
#pragma unroll
    for (int j = 0; j < NSUM; ++j) {
      uint64_t *a = s_add + (uint64_t)j * cap;
      const uint32_t pe = (nn + PR_RBLOCK - 1) / PR_RBLOCK;
      const uint32_t i0 = (uint32_t)t * pe < nn ? (uint32_t)t * pe : nn;
      const uint32_t i1 = i0 + pe < nn ? i0 + pe : nn;
      uint64_t sum = 0;
      for (uint32_t i = i0; i < i1; ++i) sum += a[i];
      uint64_t run = pr_block_exscan64_n<PR_RWAVES>(sum, s_sw64);
      for (uint32_t i = i0; i < i1; ++i) { run += a[i]; a[i] = run; }
      __syncthreads();
    }
    // the non-empty (query, key) groups: lanes = keys, one query per wave
    // step, lists each query's keys in ascending order ...
    for (int q = wave; q < ps.nq; q += PR_RWAVES) {
      const uint32_t qe = s_qe[q];
      const uint32_t elo = qe & 0xFFFFu, ehi = qe >> 16;
      bool has = false;
      if (kon && elo < ehi) has = pr_start(s_bin, kbin + elo) < pr_start(s_bin, kbin + ehi);
      const uint64_t m = __ballot(has);
      if (has) s_gk[q * 64 + __builtin_popcountll(m & lt)] = (uint8_t)lane;
      if (lane == 0) s_qcnt[q] = (uint32_t)__builtin_popcountll(m);
    }
    __syncthreads();
    // ... the flat index of each query's first group, and each group's query
    {
      const uint32_t c = t < ps.nq ? s_qcnt[t] : 0u;
      const uint32_t b0 = pr_block_exscan_n<PR_RWAVES>(c, s_sw, &ngroups);
      if (t < ps.nq) s_qbase[t] = b0;
    }
    __syncthreads();
    for (int q = wave; q < ps.nq; q += PR_RWAVES) {
      const uint32_t c = s_qcnt[q];
      if ((uint32_t)lane < c) s_gq[s_qbase[q] + lane] = (uint8_t)q;
    }
    __syncthreads();
  };

  // ---- groups -> registers: lanes = groups, consecutive lanes take
  // consecutive keys of one query, so each result column is written in runs
  uint32_t ga[PR_GPT], gb[PR_GPT], gk[PR_GPT];  // sorted range [ga, gb) (gb = 0: no group), key lane
  uint64_t grow[PR_GPT];                         // result row
  auto decode = [&](uint32_t g0) {
#pragma unroll
    for (int i = 0; i < PR_GPT; ++i) {
      const uint32_t g = g0 + (uint32_t)t + (uint32_t)i * PR_RBLOCK;
      ga[i] = 0;
      gb[i] = 0;
      gk[i] = 0;
      grow[i] = 0;
      if (g < ngroups) {
        const uint32_t q = s_gq[g];
        const uint32_t jq = g - s_qbase[q];
        const uint32_t k = s_gk[q * 64 + jq];
        const uint32_t qe = s_qe[q];
        ga[i] = pr_start(s_bin, k * NBK + (qe & 0xFFFFu));
        gb[i] = pr_start(s_bin, k * NBK + (qe >> 16));  // > ga: the group is non-empty
        gk[i] = k;
        grow[i] = s_qrow[q] + jq;
      }
    }
  };
  // The held groups are finished in pairs: per output, each pair dispatches
  // the output's kind, states and store width once (uniform), and only one
  // pair's values are live at a time.
  constexpr int PW = 2;
  // count state st (COUNT(*) = 0 or an additive COUNT(col)) of held groups
  // h .. h + PW - 1 (no group: 0)
  auto cntw = [&](int st, int h, uint64_t (&r)[PW]) {
    if (st >= 1 && st <= NSUM) {
      const uint64_t *ad = s_add + (uint64_t)(st - 1) * cap;
#pragma unroll
      for (int i = 0; i < PW; ++i)
        r[i] = gb[h + i] ? ad[gb[h + i] - 1] - (ga[h + i] ? ad[ga[h + i] - 1] : 0ull) : 0ull;
    } else {
#pragma unroll
      for (int i = 0; i < PW; ++i) r[i] = st == 0 ? (uint64_t)(gb[h + i] - ga[h + i]) : 0ull;
    }
  };
  // value state st (additive, or an order code: the max over the group's
  // sorted positions [ga, gb)) of held groups h .. h + PW - 1
  auto valw = [&](int st, int h, uint64_t (&r)[PW]) {
    if (NMAX > 0 && st > NSUM && st <= NSUM + NMAX && ((narm >> (st - 1 - NSUM)) & 1u)) {
      const uint32_t *md = reinterpret_cast<const uint32_t *>(s_max + (uint64_t)(st - 1 - NSUM) * cap);
#pragma unroll
      for (int i = 0; i < PW; ++i) {
        uint32_t x0 = 0, x1 = 0;
        uint32_t k = ga[h + i];
        for (; k + 1 < gb[h + i]; k += 2) {
          x0 = x0 > md[k] ? x0 : md[k];
          x1 = x1 > md[k + 1] ? x1 : md[k + 1];
        }
        if (k < gb[h + i]) x0 = x0 > md[k] ? x0 : md[k];
        r[i] = x0 > x1 ? x0 : x1;
      }
    } else if (NMAX > 0 && st > NSUM && st <= NSUM + NMAX) {
      const uint64_t *md = s_max + (uint64_t)(st - 1 - NSUM) * cap;
#pragma unroll
      for (int i = 0; i < PW; ++i) {
        uint64_t x0 = 0, x1 = 0;
        uint32_t k = ga[h + i];
        for (; k + 1 < gb[h + i]; k += 2) {
          x0 = u64max(x0, md[k]);
          x1 = u64max(x1, md[k + 1]);
        }
        if (k < gb[h + i]) x0 = u64max(x0, md[k]);
        r[i] = u64max(x0, x1);
      }
    } else {
      cntw(st, h, r);
    }
  };
  // ---- write: the groups in registers (partition P), one output at a time,
  // so an output's kind and state are uniform and only that state is read.
  // Values as finalize_slot derives them: a group with no counted value is
  // NULL; an all-NaN MIN / MAX keeps the identity.
  auto write = [&](uint32_t P) {
#pragma unroll 1
    for (int o = 0; o < nouts; ++o) {
      const uint64_t rb = rfl64(s_out[o][0]);
      ROut ro;
      memcpy(&ro, &rb, sizeof ro);
      uint8_t *ob = reinterpret_cast<uint8_t *>(rfl64(s_out[o][1]));
      uint32_t *onb = reinterpret_cast<uint32_t *>(rfl64(s_out[o][2]));
      const int lw = pr_out_lw(ro);
#pragma unroll
      for (int h = 0; h < PR_GPT; h += PW) {
        uint64_t val[PW];
        uint32_t nm = 0;  // bit i: held group h + i is NULL
        if (ro.kind == OK_KEY) {
          const int64_t k0 = ps.kmin + (int64_t)((uint64_t)P << ps.s);
#pragma unroll
          for (int i = 0; i < PW; ++i) val[i] = (uint64_t)(k0 + (int64_t)gk[h + i]);
        } else if (ro.kind == OK_COUNT) {
          cntw(ro.cnt_state, h, val);
        } else {
          {
            uint64_t c[PW];
            cntw(ro.cnt_state, h, c);
#pragma unroll
            for (int i = 0; i < PW; ++i) nm |= c[i] == 0 ? 1u << i : 0u;
          }
          valw(ro.state, h, val);
          if (ro.kind == OK_SUM_I64) {
#pragma unroll
            for (int i = 0; i < PW; ++i) val[i] = (nm >> i) & 1u ? 0ull : val[i];
          } else if (ro.kind == OK_AVG_I64) {
            uint64_t c[PW];
            cntw(ro.cnt_state, h, c);
#pragma unroll
            for (int i = 0; i < PW; ++i)
              val[i] = f64_bits((nm >> i) & 1u ? 0.0 : (double)(int64_t)val[i] / (double)c[i]);
          } else {  // OK_MIN / OK_MAX (order codes; MIN keeps ~code)
            const bool mn = ro.kind == OK_MIN;
            if (ro.src_is_float) {
              const bool f32 = ro.out_dtype == DT_F32;
              const uint64_t ident = mn ? 0x7FF0000000000000ull : 0xFFF0000000000000ull;
#pragma unroll
              for (int i = 0; i < PW; ++i) {
                const uint64_t d = (nm >> i) & 1u ? 0ull
                                                   : (val[i] == 0 ? ident : decode_f64_code(mn ? ~val[i] : val[i]));
                val[i] = f32 ? (uint64_t)f32_bits((float)bits_f64(d)) : d;
              }
            } else if (ro.state > NSUM && ro.state <= NSUM + NMAX && ((narm >> (ro.state - 1 - NSUM)) & 1u)) {
#pragma unroll
              for (int i = 0; i < PW; ++i) {
                const uint32_t c32 = (uint32_t)val[i];
                val[i] = (nm >> i) & 1u ? 0ull : (uint64_t)(int64_t)(int32_t)((mn ? ~c32 : c32) ^ 0x80000000u);
              }
            } else {
#pragma unroll
              for (int i = 0; i < PW; ++i)
                val[i] = (nm >> i) & 1u ? 0ull : (uint64_t)decode_i64(mn ? ~val[i] : val[i]);
            }
          }
        }
        if (lw == 3) {
#pragma unroll
          for (int i = 0; i < PW; ++i)
            if (gb[h + i]) *gptr(reinterpret_cast<uint64_t *>(ob + (grow[h + i] << 3))) = val[i];
        } else if (lw == 2) {
#pragma unroll
          for (int i = 0; i < PW; ++i)
            if (gb[h + i]) *gptr(reinterpret_cast<uint32_t *>(ob + (grow[h + i] << 2))) = (uint32_t)val[i];
        } else if (lw == 1) {
#pragma unroll
          for (int i = 0; i < PW; ++i)
            if (gb[h + i]) *gptr(reinterpret_cast<uint16_t *>(ob + (grow[h + i] << 1))) = (uint16_t)val[i];
        } else {
#pragma unroll
          for (int i = 0; i < PW; ++i)
            if (gb[h + i]) *gptr(ob + grow[h + i]) = (uint8_t)val[i];
        }
        if (onb && nm) {
#pragma unroll
          for (int i = 0; i < PW; ++i)
            if (gb[h + i] && ((nm >> i) & 1u)) g_atomic_or(&onb[grow[h + i] >> 5], 1u << (grow[h + i] & 31));
        }
      }
    }
  };

  for (uint32_t P = P0; P < P1; ++P) {
    stage_load(P);
    stage_sort();
    for (uint32_t g0 = 0; g0 < ngroups; g0 += PR_GPT * PR_RBLOCK) {
      decode(g0);
      write(P);
    }
    __syncthreads();  // LDS (bins, sorted states, group lists, first rows) is reused by the next partition
  }
}

}  // namespace qm

using namespace qm;

namespace {

template <int SRC>
const void *pr_scatter_fn(int rw) {
  switch (rw) {
    case 2: return (const void *)pr_scatter_kernel<SRC, 2>;
    case 3: return (const void *)pr_scatter_kernel<SRC, 3>;
    case 4: return (const void *)pr_scatter_kernel<SRC, 4>;
    case 5: return (const void *)pr_scatter_kernel<SRC, 5>;
    case 6: return (const void *)pr_scatter_kernel<SRC, 6>;
    case 7: return (const void *)pr_scatter_kernel<SRC, 7>;
    case 8: return (const void *)pr_scatter_kernel<SRC, 8>;
    default: return nullptr;
  }
}

const void *pr_runs_fn(int nsum, int nmax, int rw) {
#define PR_CASE(a, b) \
  if (nsum == a && nmax == b) return rw == 4 ? (const void *)pr_runs_kernel<a, b, 4> : (const void *)pr_runs_kernel<a, b, 0>;
  PR_CASE(0, 0) PR_CASE(1, 0) PR_CASE(2, 0) PR_CASE(3, 0) PR_CASE(4, 0)
  PR_CASE(0, 1) PR_CASE(1, 1) PR_CASE(2, 1) PR_CASE(3, 1) PR_CASE(4, 1)
  PR_CASE(0, 2) PR_CASE(1, 2) PR_CASE(2, 2) PR_CASE(3, 2)
  PR_CASE(0, 3) PR_CASE(1, 3) PR_CASE(2, 3)
#undef PR_CASE
  return nullptr;
}

}  // namespace

extern "C" uint32_t qmk_pr_tile(int rw) {
  return rw >= 2 && rw <= PR_MAX_RW ? (uint32_t)(PR_STAGE_WORDS / rw / PR_BLOCK) * PR_BLOCK : 0u;
}

extern "C" int qmk_pr_supported(int nsum, int nmax) { return pr_runs_fn(nsum, nmax, 0) != nullptr; }

// dynamic LDS of the write pass for a partition capacity `cap`
extern "C" size_t qmk_pr_runs_lds(uint32_t nbin, int nsum, int nmax, uint32_t cap) {
  return (size_t)(((nbin / 2 + 1) * 4 + 15) & ~15u) + (size_t)(nsum + nmax) * cap * 8;
}
// largest partition capacity that fits the LDS of one workgroup
extern "C" uint32_t qmk_pr_max_cap(uint32_t nbin, int nsum, int nmax) {
  const size_t budget = 118 * 1024;  // of 160 KB: the rest is the kernel's static LDS (query and group lists)
  const size_t fixed = qmk_pr_runs_lds(nbin, 0, 0, 0);
  if (fixed >= budget) return 0;
  uint64_t cap = (uint64_t)PR_RIPT * PR_RBLOCK;
  if (nsum + nmax > 0) {
    const uint64_t c2 = (budget - fixed) / (8ull * (uint64_t)(nsum + nmax));
    cap = c2 < cap ? c2 : cap;
  }
  return (uint32_t)cap;
}

extern "C" hipError_t qmk_pr_scatter(const PRScan *ps, int src, uint32_t grid, hipStream_t st) {
  const void *fn = src == 0 ? pr_scatter_fn<0>(ps->rw) : pr_scatter_fn<1>(ps->rw);
  if (!fn || grid == 0) return hipErrorInvalidValue;
  PRScan a = *ps;
  void *args[] = {&a};
  return hipLaunchKernel(fn, dim3(grid), dim3(PR_BLOCK), args, 0, st);
}

extern "C" hipError_t qmk_pr_count(const PRScan *ps, hipStream_t st) {
  if (ps->np == 0) return hipSuccess;
  PRScan a = *ps;
  void *args[] = {&a};
  return hipLaunchKernel((const void *)pr_count_kernel, dim3(ps->np), dim3(PR_BLOCK), args, 0, st);
}

extern "C" hipError_t qmk_pr_runs(const PRScan *ps, hipStream_t st) {
  const int nmax = ps->nstates - 1 - ps->nsum;
  const void *fn = pr_runs_fn(ps->nsum, nmax, ps->rw);
  if (!fn) return hipErrorInvalidValue;
  if (ps->np == 0) return hipSuccess;
  const uint32_t nbin = (uint32_t)(ps->nbounds + 2) << ps->s;
  const size_t lds = qmk_pr_runs_lds(nbin, ps->nsum, nmax, ps->cap);
  hipError_t e = qmk_func_lds(fn, (size_t)lds);
  if (e != hipSuccess) return e;
  int bpc = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, fn, PR_RBLOCK, lds) != hipSuccess || bpc < 1) bpc = 1;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus < 1)
    cus = 256;
  const uint32_t grid = (uint32_t)std::min<uint64_t>((uint64_t)cus * (uint64_t)bpc, ps->np);
  PRScan a = *ps;
  void *args[] = {&a};
  return hipLaunchKernel(fn, dim3(grid), dim3(PR_RBLOCK), args, lds, st);
}

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
#include "k_common.hip.h"
#include "qm_kernels.h"

namespace qm {

namespace {

static constexpr int PR_WAVES = PR_BLOCK / 64;
static constexpr int PR_CIPT = 8;  // records per thread in the write pass (cap <= PR_CIPT * PR_BLOCK)

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

}  // namespace

// ---------------------------------------------------------------------------
// passes 1 and 2: scatter by a key digit through an LDS-staged counting sort
// ---------------------------------------------------------------------------
template <int SRC, int RW>
__global__ void __launch_bounds__(PR_BLOCK) pr_scatter_kernel(PRScan ps) {
  constexpr int IPT = PR_STAGE_WORDS / RW / PR_BLOCK;
  constexpr uint32_t TILE = (uint32_t)IPT * PR_BLOCK;
  constexpr int NVX = (RW - 1) < PR_MAX_V ? (RW - 1) : PR_MAX_V;  // value columns a record of RW words can hold
  __shared__ __align__(16) uint32_t s_stage[TILE * RW];
  __shared__ uint16_t s_dig[TILE];
  __shared__ uint32_t s_hist[PR_MAX_P];
  __shared__ uint32_t s_lim[PR_MAX_P];
  __shared__ uint64_t s_gofs[PR_MAX_P];
  __shared__ uint64_t s_bounds[SRC == 0 ? R_MAX_BUCKETS : 1];
  __shared__ uint32_t s_brow[SRC == 0 ? R_MAX_BUCKETS : 1];
  __shared__ uint32_t s_sw[PR_WAVES + 1];
  const int t = threadIdx.x;
  const uint32_t ndig = SRC == 0 ? ps.np1 : (1u << (ps.sh1 - ps.s));
  const uint32_t dpt = (ndig + PR_BLOCK - 1) / PR_BLOCK;  // digits per thread in the scan (<= 4)
  const uint32_t nb = (uint32_t)ps.nbounds;
  if constexpr (SRC == 0) {
    const int nbp = (1 << ps.bsteps) - 1;
    for (int i = t; i < nbp; i += PR_BLOCK) s_bounds[i] = ps.bounds[i];
    for (int i = t; i < R_MAX_BUCKETS; i += PR_BLOCK) s_brow[i] = 0;
  }
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
      uint32_t lo = 0, hi = ps.np1;  // last p1 with chunk0[p1] <= u
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (ps.chunk0[mid] <= u) lo = mid; else hi = mid;
      }
      p1 = lo;
      const uint64_t ci = u - ps.chunk0[p1];
      const uint64_t cap1 = ps.start1[p1 + 1] - ps.start1[p1];
      const uint64_t n1 = ps.cur1[p1] < cap1 ? ps.cur1[p1] : cap1;
      src0 = ps.start1[p1] + ci * TILE;
      srcn = ci * TILE < n1 ? (n1 - ci * TILE < TILE ? n1 - ci * TILE : TILE) : 0;
    }
    uint32_t kb[IPT], dg[IPT], rk[IPT], nbits[IPT];
    uint64_t v[IPT][NVX > 0 ? NVX : 1];
    uint32_t w[IPT][RW];
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      const uint64_t it = (uint64_t)t + (uint64_t)i * PR_BLOCK;
      rk[i] = ~0u;
      dg[i] = 0;
      kb[i] = 0;
      nbits[i] = 0;
      if (it >= srcn) continue;
      if constexpr (SRC == 0) {
        const uint64_t row = src0 + it;
        DCol kc;
        kc.data = ps.key; kc.nulls = nullptr; kc.dtype = ps.key_dt; kc.side = 0;
        const uint64_t code = (uint64_t)(load1_i(kc, row) - ps.kmin);
        uint32_t e;
        if (pr_null(ps.pnull, row)) e = nb + 1;
        else e = pr_bucket(s_bounds, ps.bsteps, nb, pr_code(pr_bits(ps.pcol, ps.p_dt, row), ps.p_dt));
        if (!((ps.bucket_on[e >> 5] >> (e & 31)) & 1u)) continue;
        const uint64_t d = code >> ps.sh1;
        if (d >= ndig) { atomicOr(ps.status, 1u); continue; }
        dg[i] = (uint32_t)d;
        kb[i] = ((uint32_t)code & kmask1) | (e << 16);
#pragma unroll
        for (int j = 0; j < NVX; ++j) {
          v[i][j] = 0;
          if (j < ps.nv) {
            v[i][j] = pr_bits(ps.vcol[j], ps.v_dt[j], row);
            if (pr_null(ps.vnull[j], row)) nbits[i] |= 1u << j;
          }
        }
        atomicAdd(&s_brow[e], 1u);
      } else {
        const uint32_t *r = ps.rec1 + (src0 + it) * RW;
#pragma unroll
        for (int k = 0; k < RW; ++k) w[i][k] = r[k];
        dg[i] = (w[i][0] & 0xFFFFu) >> ps.s;
      }
      rk[i] = atomicAdd(&s_hist[dg[i]], 1u);
    }
    __syncthreads();
    // digit starts in the tile, and one reservation per digit in its region
    uint32_t cnt[4] = {0, 0, 0, 0}, mine = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t d = (uint32_t)t * dpt + k;
      if (k < (int)dpt && d < ndig) cnt[k] = s_hist[d];
      mine += cnt[k];
    }
    uint32_t total;
    uint32_t run = pr_block_exscan(mine, s_sw, &total);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
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
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const uint32_t P = blockIdx.x;
  const uint64_t base = ps.start2[P];
  const uint64_t capr = ps.start2[P + 1] - base;
  const uint64_t n = ps.cur2[P] < capr ? ps.cur2[P] : capr;
  for (int i = t; i < 64 * BW; i += PR_BLOCK) s_bm[i] = 0;
  __syncthreads();
  const uint32_t km = (1u << ps.s) - 1u;
  for (uint64_t j = t; j < n; j += PR_BLOCK) {
    const uint32_t kb = ps.rec2[(base + j) * (uint64_t)ps.rw];
    const uint32_t e = kb >> 16;
    atomicOr(&s_bm[(kb & km) * BW + (e >> 5)], 1u << (e & 31));
  }
  __syncthreads();
  const bool kon = (uint32_t)lane <= km;
  const uint32_t *bm = s_bm + lane * BW;
  for (int q = wave; q < ps.nq; q += PR_WAVES) {
    const uint32_t qe = ps.qe[q];
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

__device__ __forceinline__ void pr_store_i(void *out, int32_t dt, uint64_t i, int64_t v) {
  switch (dt) {
    case DT_I8: static_cast<int8_t *>(out)[i] = (int8_t)v; break;
    case DT_BOOL: static_cast<uint8_t *>(out)[i] = (uint8_t)v; break;
    case DT_I16: static_cast<int16_t *>(out)[i] = (int16_t)v; break;
    case DT_I32: static_cast<int32_t *>(out)[i] = (int32_t)v; break;
    default: static_cast<int64_t *>(out)[i] = v; break;
  }
}
__device__ __forceinline__ void pr_store_f(void *out, int32_t dt, uint64_t i, double v) {
  if (dt == DT_F32) static_cast<float *>(out)[i] = (float)v;
  else static_cast<double *>(out)[i] = v;
}

// one output value of one group (the values finalize_slot derives: a group
// with no counted value is NULL; an all-NaN MIN / MAX keeps the identity)
__device__ __forceinline__ void pr_emit(const ROut &ro, void *dst, uint32_t *nbm, uint64_t gg, int64_t key,
                                        uint64_t v, uint64_t cnt) {
  bool isnull = false;
  switch (ro.kind) {
    case OK_KEY: pr_store_i(dst, ro.out_dtype, gg, key); break;
    case OK_COUNT: static_cast<int64_t *>(dst)[gg] = (int64_t)cnt; break;
    case OK_SUM_I64:
      isnull = cnt == 0;
      static_cast<int64_t *>(dst)[gg] = isnull ? 0 : (int64_t)v;
      break;
    case OK_AVG_I64:
      isnull = cnt == 0;
      static_cast<double *>(dst)[gg] = isnull ? 0.0 : (double)(int64_t)v / (double)cnt;
      break;
    default: {  // OK_MIN / OK_MAX
      const bool mn = ro.kind == OK_MIN;
      if (cnt == 0) {
        isnull = true;
        if (ro.src_is_float) pr_store_f(dst, ro.out_dtype, gg, 0.0);
        else pr_store_i(dst, ro.out_dtype, gg, 0);
      } else if (ro.src_is_float) {
        double d;
        if (v == 0) d = mn ? bits_f64(0x7FF0000000000000ull) : bits_f64(0xFFF0000000000000ull);
        else d = bits_f64(decode_f64_code(mn ? ~v : v));
        pr_store_f(dst, ro.out_dtype, gg, d);
      } else {
        pr_store_i(dst, ro.out_dtype, gg, decode_i64(mn ? ~v : v));
      }
      break;
    }
  }
  if (isnull && nbm) atomicOr(&nbm[gg >> 5], 1u << (gg & 31));
}

// state value si of a group: 0 = its row count, then the additive sums, then the max codes
template <int NSUM, int NMAX>
__device__ __forceinline__ uint64_t pr_state(int si, uint64_t cnt, const uint64_t (&sm)[NSUM > 0 ? NSUM : 1],
                                             const uint64_t (&mx)[NMAX > 0 ? NMAX : 1]) {
  uint64_t v = cnt;
#pragma unroll
  for (int j = 0; j < NSUM; ++j) v = si == 1 + j ? sm[j] : v;
#pragma unroll
  for (int j = 0; j < NMAX; ++j) v = si == 1 + NSUM + j ? mx[j] : v;
  return v;
}

template <int NSUM, int NMAX>
__global__ void __launch_bounds__(PR_BLOCK) pr_runs_kernel(PRScan ps) {
  constexpr int NSA = NSUM > 0 ? NSUM : 1, NMA = NMAX > 0 ? NMAX : 1;
  extern __shared__ __align__(16) uint8_t s_dyn[];
  __shared__ uint32_t s_sw[PR_WAVES + 1];
  __shared__ uint64_t s_sw64[PR_WAVES];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const uint32_t P = blockIdx.x;
  const uint64_t base = ps.start2[P];
  const uint64_t capr = ps.start2[P + 1] - base;
  const uint32_t n = (uint32_t)(ps.cur2[P] < capr ? ps.cur2[P] : capr);
  if (n > ps.cap) {  // the host sizes partitions from the key histogram; never expected
    if (t == 0) atomicOr(ps.status, 2u);
    return;
  }
  const uint32_t NBK = (uint32_t)ps.nbounds + 2;
  const uint32_t nbin = (NBK << ps.s);
  const uint32_t nbw = nbin / 2 + 1;              // packed u16 entries 0 .. nbin
  const uint32_t cap = ps.cap;
  uint32_t *s_bin = reinterpret_cast<uint32_t *>(s_dyn);
  uint64_t *s_add = reinterpret_cast<uint64_t *>(s_dyn + ((nbw * 4 + 15) & ~15u));
  uint64_t *s_max = s_add + (uint64_t)NSUM * cap;
  for (uint32_t i = t; i < nbw; i += PR_BLOCK) s_bin[i] = 0;
  __syncthreads();
  // 1. records -> (key, bucket) bin and rank within it; state contributions
  const uint32_t km = (1u << ps.s) - 1u;
  uint32_t bn[PR_CIPT], rk[PR_CIPT];
  uint64_t ca[PR_CIPT][NSA], cm[PR_CIPT][NMA];
#pragma unroll
  for (int i = 0; i < PR_CIPT; ++i) {
    const uint32_t j = (uint32_t)t + (uint32_t)i * PR_BLOCK;
    rk[i] = ~0u;
    bn[i] = 0;
    if (j >= n) continue;
    const uint32_t *r = ps.rec2 + (base + j) * (uint64_t)ps.rw;
    const uint32_t kb = r[0];
    const uint32_t nbits = ps.null_word >= 0 ? r[ps.null_word] : 0u;
#pragma unroll
    for (int s2 = 0; s2 < NSUM + NMAX; ++s2) {
      const DState &st = ps.st[1 + s2];
      const int vj = ps.st_v[1 + s2];
      uint64_t c = 0;
      if (vj < 0) {
        c = st.kind == SK_COUNT ? 1ull : 0ull;  // COUNT(*)
      } else if (!((nbits >> vj) & 1u)) {
        const int32_t dt = ps.v_dt[vj];
        const int wd = ps.v_word[vj];
        const uint64_t bits = pr_wide(dt) ? ((uint64_t)r[wd] | ((uint64_t)r[wd + 1] << 32)) : (uint64_t)r[wd];
        if (st.kind == SK_COUNT) {
          c = 1;
        } else if (st.kind == SK_SUM_I64) {
          c = (uint64_t)pr_ival(bits, dt);
        } else {  // SK_MAX_CODE: MAX keeps code, MIN ~code; NaN never updates either
          bool nan = false;
          uint64_t code;
          if (dt_is_float(dt)) {
            const double fv = pr_fval(bits, dt);
            nan = fv != fv;
            code = code_f64_bits(f64_bits(fv));
          } else {
            code = code_i64(pr_ival(bits, dt));
          }
          c = nan ? 0ull : (st.is_min ? ~code : code);
        }
      }
      if (s2 < NSUM) ca[i][s2 < NSA ? s2 : 0] = c;
      else cm[i][(s2 - NSUM) < NMA ? (s2 - NSUM) : 0] = c;
    }
    const uint32_t e = kb >> 16;
    bn[i] = (kb & km) * NBK + e;
    const uint32_t sh = (bn[i] & 1u) * 16u;
    rk[i] = (atomicAdd(&s_bin[bn[i] >> 1], 1u << sh) >> sh) & 0xFFFFu;
  }
  __syncthreads();
  // 2. exclusive scan of the bins (packed pairs; thread t owns a word range)
  {
    const uint32_t wpt = (nbw + PR_BLOCK - 1) / PR_BLOCK;
    const uint32_t w0 = (uint32_t)t * wpt < nbw ? (uint32_t)t * wpt : nbw;
    const uint32_t w1 = w0 + wpt < nbw ? w0 + wpt : nbw;
    uint32_t sum = 0;
    for (uint32_t w = w0; w < w1; ++w) { const uint32_t x = s_bin[w]; sum += (x & 0xFFFFu) + (x >> 16); }
    uint32_t total;
    uint32_t run = pr_block_exscan(sum, s_sw, &total);
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
  // 3. scatter the contributions into (key, bucket) order
#pragma unroll
  for (int i = 0; i < PR_CIPT; ++i) {
    if (rk[i] == ~0u) continue;
    const uint32_t slot = pr_start(s_bin, bn[i]) + rk[i];
#pragma unroll
    for (int j = 0; j < NSUM; ++j) s_add[(uint64_t)j * cap + slot] = ca[i][j];
#pragma unroll
    for (int j = 0; j < NMAX; ++j) s_max[(uint64_t)j * cap + slot] = cm[i][j];
  }
  __syncthreads();
  // 4. inclusive prefix sums of the additive states over the sorted order
#pragma unroll
  for (int j = 0; j < NSUM; ++j) {
    uint64_t *a = s_add + (uint64_t)j * cap;
    const uint32_t per = (n + PR_BLOCK - 1) / PR_BLOCK;
    const uint32_t i0 = (uint32_t)t * per < n ? (uint32_t)t * per : n;
    const uint32_t i1 = i0 + per < n ? i0 + per : n;
    uint64_t sum = 0;
    for (uint32_t i = i0; i < i1; ++i) sum += a[i];
    uint64_t run = pr_block_exscan64(sum, s_sw64);
    for (uint32_t i = i0; i < i1; ++i) { run += a[i]; a[i] = run; }
    __syncthreads();
  }
  // 5. lanes = keys, one query per wave step
  const bool kon = (uint32_t)lane <= km;
  const int64_t key = ps.kmin + (int64_t)(((uint64_t)P << ps.s) + (uint64_t)lane);
  const uint32_t kbin = (uint32_t)lane * NBK;
  const uint64_t lt = (1ull << lane) - 1ull;
  for (int q = wave; q < ps.nq; q += PR_WAVES) {
    const uint32_t qe = ps.qe[q];
    const uint32_t elo = qe & 0xFFFFu, ehi = qe >> 16;
    if (elo >= ehi) continue;
    uint32_t a = 0, b = 0;
    if (kon) {
      a = pr_start(s_bin, kbin + elo);
      b = pr_start(s_bin, kbin + ehi);
    }
    const bool has = b > a;
    const uint64_t m = __ballot(has);
    if (m == 0) continue;
    if (!has) continue;
    const uint64_t qrow = ps.gcount[(uint64_t)q * ps.np + P] - ps.gcount[(uint64_t)q * ps.np];
    const uint64_t gg = ps.row_off[q] + qrow + (uint64_t)__builtin_popcountll(m & lt);
    const uint64_t cnt = b - a;
    uint64_t sm[NSA], mx[NMA];
#pragma unroll
    for (int j = 0; j < NSUM; ++j) {
      const uint64_t *ad = s_add + (uint64_t)j * cap;
      sm[j] = ad[b - 1] - (a ? ad[a - 1] : 0ull);
    }
#pragma unroll
    for (int j = 0; j < NMAX; ++j) {
      const uint64_t *md = s_max + (uint64_t)j * cap;
      uint64_t x = 0;
      for (uint32_t i = a; i < b; ++i) x = u64max(x, md[i]);
      mx[j] = x;
    }
    for (int o = 0; o < ps.nouts; ++o) {
      const ROut ro = ps.outs[o];
      const uint64_t v = ro.state >= 0 ? pr_state<NSUM, NMAX>(ro.state, cnt, sm, mx) : 0ull;
      const uint64_t c = ro.cnt_state >= 0 ? pr_state<NSUM, NMAX>(ro.cnt_state, cnt, sm, mx) : 0ull;
      pr_emit(ro, ps.out_base[o], ps.out_nbase[o], gg, key, v, c);
    }
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

const void *pr_runs_fn(int nsum, int nmax) {
#define PR_CASE(a, b) if (nsum == a && nmax == b) return (const void *)pr_runs_kernel<a, b>;
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

extern "C" int qmk_pr_supported(int nsum, int nmax) { return pr_runs_fn(nsum, nmax) != nullptr; }

// dynamic LDS of the write pass for a partition capacity `cap`
extern "C" size_t qmk_pr_runs_lds(uint32_t nbin, int nsum, int nmax, uint32_t cap) {
  return (size_t)(((nbin / 2 + 1) * 4 + 15) & ~15u) + (size_t)(nsum + nmax) * cap * 8;
}
// largest partition capacity that fits the LDS of one workgroup
extern "C" uint32_t qmk_pr_max_cap(uint32_t nbin, int nsum, int nmax) {
  const size_t budget = 150 * 1024;
  const size_t fixed = qmk_pr_runs_lds(nbin, 0, 0, 0);
  if (fixed >= budget) return 0;
  uint64_t cap = (uint64_t)PR_CIPT * PR_BLOCK;
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
  const void *fn = pr_runs_fn(ps->nsum, nmax);
  if (!fn) return hipErrorInvalidValue;
  if (ps->np == 0) return hipSuccess;
  const uint32_t nbin = (uint32_t)(ps->nbounds + 2) << ps->s;
  const size_t lds = qmk_pr_runs_lds(nbin, ps->nsum, nmax, ps->cap);
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  PRScan a = *ps;
  void *args[] = {&a};
  return hipLaunchKernel(fn, dim3(ps->np), dim3(PR_BLOCK), args, lds, st);
}

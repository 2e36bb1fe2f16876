// k_runs.hip — many aggregate queries grouped by a sorted integer key.
//
// The table is stored in ascending key order, so a group is a run of
// consecutive rows. Every query of a launch filters by (at most) one range on
// one common column and shares one aggregate list. The launch's sorted
// distinct range bounds cut that column into buckets: bucket(x) = #{bounds <= x}
// (NULLs: bucket nbounds + 1), and a query's range is the bucket range
// [e_lo, e_hi).
//
// A wave owns the runs that start inside its chunk of rows and walks them in
// segments of at most 64 rows (one row per lane):
//   1. each lane finds its row's bucket (binary search over the bounds in LDS);
//   2. a histogram over buckets and its prefix sum give every bucket's first
//      position; scattering the rows there counting-sorts the segment by
//      bucket;
//   3. additive states (COUNT(col), integer SUM) get a prefix sum over the
//      sorted order, order-code states (MIN / MAX) a 6-level sparse max table;
//   4. then lanes switch to queries: a query's rows in the segment are the
//      sorted span [C(e_lo), C(e_hi)), so COUNT is a difference of two prefix
//      entries, a SUM another difference, a MIN / MAX the max of two table
//      entries. The per-(segment, query) cost is a handful of LDS reads,
//      independent of the run length.
// A run's groups are final when the run ends, so each non-empty (run, query)
// is written once, with plain stores, straight into the query's result
// columns. A count pass computes every (chunk, query) group count first; the
// per-query exclusive scan over chunks gives each wave its output positions,
// so results come out in ascending key order (the reference's first-seen
// group order on a table stored in key order).
//
// Aggregate semantics are those of src/partial-agg.ts:38-134 (per group:
// COUNT, exact integer SUM, MIN / MAX by value, NaN never updates MIN / MAX);
// the filter semantics those of src/decode.ts:477-578 (NULL never matches).
#include "k_agg.hip.h"
#include "qm_kernels.h"

namespace qm {

static constexpr int R_WAVES = R_BLOCK / 64;

// order the LDS accesses of the lanes of one wave (the LDS serves a wave's
// requests in issue order; this keeps the compiler from moving them across)
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inclusive wave scans with DPP moves (row shifts within 16-lane rows, then
// row broadcasts of lanes 15 and 31): VALU only, no LDS crossbar traffic.
// Lanes whose source is outside the row (or whose row is masked) add 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp32(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROW_MASK, 0xF, false);
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint64_t dpp64(uint64_t x) {
  const uint32_t lo = dpp32<CTRL, ROW_MASK>((uint32_t)x);
  const uint32_t hi = dpp32<CTRL, ROW_MASK>((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}
template <class T, class F>
__device__ __forceinline__ T wave_incl_scan(T x, F sh) {
  x += sh.template operator()<0x111, 0xF>(x);  // row_shr:1
  x += sh.template operator()<0x112, 0xF>(x);  // row_shr:2
  x += sh.template operator()<0x114, 0xF>(x);  // row_shr:4
  x += sh.template operator()<0x118, 0xF>(x);  // row_shr:8
  x += sh.template operator()<0x142, 0xA>(x);  // row_bcast:15 -> rows 1, 3
  x += sh.template operator()<0x143, 0xC>(x);  // row_bcast:31 -> rows 2, 3
  return x;
}
struct Dpp32 { template <int C, int M> __device__ uint32_t operator()(uint32_t x) const { return dpp32<C, M>(x); } };
struct Dpp64 { template <int C, int M> __device__ uint64_t operator()(uint64_t x) const { return dpp64<C, M>(x); } };
__device__ __forceinline__ uint32_t wave_scan_u32(uint32_t x) { return wave_incl_scan<uint32_t>(x, Dpp32{}); }
__device__ __forceinline__ uint64_t wave_scan_u64(uint64_t x) { return wave_incl_scan<uint64_t>(x, Dpp64{}); }

__device__ __forceinline__ uint64_t rfirst64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ DCol rcol(const void *data, int32_t dt) {
  DCol c;
  c.data = data; c.nulls = nullptr; c.dtype = dt; c.side = 0;
  return c;
}

__device__ __forceinline__ int64_t rkey(const RScan &rs, uint64_t row) { return load1_i(rcol(rs.key, rs.key_dt), row); }

__device__ __forceinline__ bool rnull(const uint8_t *nb, uint64_t row) {
  return nb != nullptr && ((nb[row >> 3] >> (row & 7)) & 1);
}

__device__ __forceinline__ uint64_t rcode(const void *data, int32_t dt, uint64_t row) {
  const DCol c = rcol(data, dt);
  if (dt_is_float(dt)) return code_f64_bits(f64_bits(load1_f(c, row)));
  return code_i64(load1_i(c, row));
}

__device__ __forceinline__ void rstore_i(void *out, int32_t dt, uint64_t i, int64_t v) {
  switch (dt) {
    case DT_I8: static_cast<int8_t *>(out)[i] = (int8_t)v; break;
    case DT_BOOL: static_cast<uint8_t *>(out)[i] = (uint8_t)v; break;
    case DT_I16: static_cast<int16_t *>(out)[i] = (int16_t)v; break;
    case DT_I32: static_cast<int32_t *>(out)[i] = (int32_t)v; break;
    default: static_cast<int64_t *>(out)[i] = v; break;
  }
}
__device__ __forceinline__ void rstore_f(void *out, int32_t dt, uint64_t i, double v) {
  if (dt == DT_F32) static_cast<float *>(out)[i] = (float)v;
  else static_cast<double *>(out)[i] = v;
}

// value of state si of a finished run (0: the row count, then st[si - 1])
template <int NS>
__device__ __forceinline__ uint64_t rstate(int si, uint64_t cnt, const uint64_t (&st)[NS]) {
  uint64_t v = cnt;
#pragma unroll
  for (int j = 0; j < NS; ++j) v = (si == j + 1) ? st[j] : v;
  return v;
}

__device__ __forceinline__ uint64_t umax64(uint64_t a, uint64_t b) { return a > b ? a : b; }

// Raw column words: 8 bytes from the 4-byte-aligned address holding the row's
// value (columns are padded well past their last row), decoded only where the
// value is used, so the loads of a window stay in flight while the previous
// window is processed (a decode at the load would wait for it at once).
// staged values of 4 or 8 bytes (the width is uniform over a launch)
__device__ __forceinline__ void stg_put(void *p, int32_t w, uint64_t i, uint64_t v) {
  if (w == 4) static_cast<uint32_t *>(p)[i] = (uint32_t)v;
  else static_cast<uint64_t *>(p)[i] = v;
}
// one 8-byte load at p + (i << sh) (sh = 2 or 3) either way (the arrays end 64
// bytes past their last entry), so the loads of a window issue back to back
// without branches; stg_dec keeps the entry's bytes and adds the base where
// the value is used, so the loads stay in flight until then
__device__ __forceinline__ uint64_t stg_raw(const void *p, int32_t sh, uint64_t i) {
  const uint32_t *a = reinterpret_cast<const uint32_t *>(static_cast<const uint8_t *>(p) + (i << sh));
  return (uint64_t)a[0] | ((uint64_t)a[1] << 32);
}
__device__ __forceinline__ uint64_t stg_dec(uint64_t raw, int32_t sh, uint64_t base) {
  return (sh == 2 ? (raw & 0xFFFFFFFFull) : raw) + base;
}
__device__ __forceinline__ int rbytes(int32_t dt) {
  return (dt == DT_F64 || dt == DT_I64) ? 8 : ((dt == DT_I32 || dt == DT_F32) ? 4 : (dt == DT_I16 ? 2 : 1));
}
// (the word addresses are formed by pointer arithmetic from the column
// pointer, so the compiler keeps them global: global loads count only
// against vmcnt, where flat loads would also hold every LDS wait)
__device__ __forceinline__ uint64_t rraw(const void *data, int32_t dt, uint64_t row) {
  const uint8_t *b = static_cast<const uint8_t *>(data) + row * (uint64_t)rbytes(dt);
  const uint32_t *w = reinterpret_cast<const uint32_t *>(b - (reinterpret_cast<uintptr_t>(b) & 3));
  return (uint64_t)w[0] | ((uint64_t)w[1] << 32);
}
__device__ __forceinline__ uint32_t rraw_null(const uint8_t *nb, uint64_t row) {
  if (!nb) return 0u;
  const uint8_t *b = nb + (row >> 3);
  return *reinterpret_cast<const uint32_t *>(b - (reinterpret_cast<uintptr_t>(b) & 3));
}
__device__ __forceinline__ bool rdec_null(uint32_t raw, uint64_t row) {
  return (raw >> ((((row >> 3) & 3) << 3) + (row & 7))) & 1u;
}
__device__ __forceinline__ int64_t rdec_i(uint64_t raw, int32_t dt, uint64_t row) {
  const int sh = (int)(((row * (uint64_t)rbytes(dt)) & 3) << 3);
  switch (dt) {
    case DT_I8: return (int8_t)(raw >> sh);
    case DT_BOOL: return (uint8_t)(raw >> sh);
    case DT_I16: return (int16_t)(raw >> sh);
    case DT_I32: return (int32_t)(uint32_t)raw;
    default: return (int64_t)raw;
  }
}
__device__ __forceinline__ double rdec_f(uint64_t raw, int32_t dt) {
  return dt == DT_F32 ? (double)__uint_as_float((uint32_t)raw) : bits_f64(raw);
}
__device__ __forceinline__ uint64_t rdec_code(uint64_t raw, int32_t dt, uint64_t row) {
  if (dt_is_float(dt)) return code_f64_bits(f64_bits(rdec_f(raw, dt)));
  return code_i64(rdec_i(raw, dt, row));
}

// One 64-row window (one row per lane, rows p + lane): raw words of the key,
// the filter column (+ NULL word) and (write pass) the state columns.
template <int NS>
struct RWin {
  uint64_t p;
  bool ok;        // row < nrows
  uint64_t kr, xr;
  uint32_t xn;
  uint64_t cr[NS > 0 ? NS : 1];
  uint32_t cn[NS > 0 ? NS : 1];
};

template <int PASS, int NS>
__device__ __forceinline__ void rload(const RScan &rs, uint64_t p, int lane, RWin<NS> &w) {
  const uint64_t row = p + (uint64_t)lane;
  w.p = p;
  w.ok = row < rs.nrows;
  w.kr = w.xr = 0;
  w.xn = 0;
#pragma unroll
  for (int j = 0; j < (NS > 0 ? NS : 1); ++j) { w.cr[j] = 0; w.cn[j] = 0; }
  if (w.ok) {
    w.kr = rraw(rs.key, rs.key_dt, row);
    w.xr = rraw(rs.pcol, rs.p_dt, row);
    w.xn = rraw_null(rs.pnull, row);
    if constexpr (PASS == 1) {
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        if (rs.scol[1 + j]) w.cr[j] = rraw(rs.scol[1 + j], rs.s_dt[1 + j], row);
        w.cn[j] = rraw_null(rs.snull[1 + j], row);
      }
    }
  }
}

// contribution of the window row to state s = 1 + j (0: the state's identity)
// The run-sorted launches hold only COUNT / SUM_I64 (additive, states
// 1..nsum) and MAX_CODE states (the rest), so the write pass takes the two
// kinds apart; the values equal contrib_of's for those kinds.
template <int NS>
__device__ __forceinline__ uint64_t radd_w(const RScan &rs, const RWin<NS> &w, int j, uint64_t row) {
  const int s = 1 + j;
  if (rs.snull[s] && rdec_null(w.cn[j], row)) return 0;
  if (rs.st[s].kind == SK_COUNT) return 1;
  return (uint64_t)rdec_i(w.cr[j], rs.s_dt[s], row);  // SK_SUM_I64
}
template <int NS>
__device__ __forceinline__ uint64_t rmax_w(const RScan &rs, const RWin<NS> &w, int j, uint64_t row) {
  const int s = 1 + j;
  if (rs.snull[s] && rdec_null(w.cn[j], row)) return 0;
  uint64_t code;
  if (dt_is_float(rs.s_dt[s])) {
    const double fv = rdec_f(w.cr[j], rs.s_dt[s]);
    if (fv != fv) return 0;  // NaN never updates MIN / MAX
    code = code_f64_bits(f64_bits(fv));
  } else {
    code = code_i64(rdec_i(w.cr[j], rs.s_dt[s], row));
  }
  return rs.st[s].is_min ? ~code : code;
}

// PASS 0: groups per (chunk, query) into rs.counts.
// PASS 1: the groups themselves, at rs.offsets.
template <int PASS, int NSUM, int NMAX, int QPL>
__global__ void __launch_bounds__(R_BLOCK) runs_kernel(RScan rs) {
  constexpr int NS = NSUM + NMAX;
  constexpr int NSA = NS > 0 ? NS : 1;
  __shared__ uint64_t s_bounds[R_MAX_BUCKETS];
  __shared__ __align__(16) uint32_t s_cnt[R_WAVES][R_MAX_BUCKETS];
  __shared__ uint64_t s_sum[R_WAVES][NSUM > 0 ? NSUM : 1][64];
  __shared__ uint64_t s_tab[R_WAVES][NMAX > 0 ? NMAX : 1][6][64];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nbp = (1 << rs.bsteps) - 1;  // padded bound count
  for (int i = threadIdx.x; i < nbp; i += R_BLOCK) s_bounds[i] = rs.bounds[i];
  __syncthreads();

  const uint64_t chunk = (uint64_t)blockIdx.x * R_WAVES + wave;
  if (chunk >= rs.nchunks) return;  // no block barrier below
  uint32_t *cnt = s_cnt[wave];
  const uint32_t nb = (uint32_t)rs.nbounds;

  bool qon[QPL];
  uint32_t elo[QPL], ehi[QPL], groups[QPL];
  uint32_t opos[QPL], nsel[QPL];
  uint64_t acnt[QPL], ast[QPL][NSA];
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const int q = k * 64 + lane;
    qon[k] = q < rs.nq;
    const uint32_t w = qon[k] ? rs.qe[q] : 0u;
    elo[k] = w & 0xFFFFu;
    ehi[k] = w >> 16;
    groups[k] = 0;
    nsel[k] = 0;
    opos[k] = (PASS == 1 && qon[k]) ? (uint32_t)rs.offsets[chunk * (uint64_t)rs.nq + q] : 0u;
  }

  // staging range of this lane's query block in each slot k (the 16 lanes
  // of a DPP row share one block)
  const int qbw = rs.qbw;  // 0: one block (every slot of the wave appends to one range)
  uint64_t sbase[QPL];
  uint32_t sdone[QPL];
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const int qb = qbw ? (k * 64 + lane) / qbw : 0;
    const bool inb = PASS == 1 && qb < rs.nqb;
    sbase[k] = inb ? rs.stg_base[chunk * (uint64_t)rs.nqb + qb] : 0;
    sdone[k] = 0;
  }
  const uint64_t c0 = chunk * rs.chunk_rows;
  const uint64_t c1 = (c0 + rs.chunk_rows < rs.nrows) ? c0 + rs.chunk_rows : rs.nrows;
  uint64_t s = c0;
  if (c0 > 0 && c0 < c1) {
    // skip the rows of the run that started in the previous chunk
    const int64_t prev = rkey(rs, c0 - 1);
    for (;;) {
      const uint64_t row = s + lane;
      const bool diff = row >= rs.nrows || rkey(rs, row) != prev;
      const uint64_t m = __ballot(diff);
      if (m) { s += (uint64_t)__builtin_ctzll(m); break; }
      s += 64;
    }
  }

  // Windows of 64 rows, each loaded while the previous one is processed. A
  // window's leading rows with the current run key are the run's segment.
  uint64_t p = s;
  RWin<NS> cur, nxt;
  if (p < c1) rload<PASS, NS>(rs, p, lane, cur);
  bool fresh = true;  // `cur` starts a new run
  int64_t rk = 0;
  while (true) {
    if (fresh) {
      if (p >= c1) break;  // runs starting at or after the chunk end belong to the next chunk
      rk = (int64_t)rfirst64((uint64_t)rdec_i(cur.kr, rs.key_dt, cur.p + lane));
#pragma unroll
      for (int k = 0; k < QPL; ++k) {
        acnt[k] = 0;
#pragma unroll
        for (int j = 0; j < NSA; ++j) ast[k][j] = 0;
      }
      fresh = false;
    }
    bool run_end;
    {
      const uint64_t row = cur.p + (uint64_t)lane;
      const bool inr = cur.ok && rdec_i(cur.kr, rs.key_dt, row) == rk;
      const uint64_t m = __ballot(inr);
      const int n = (~m == 0ull) ? 64 : __builtin_ctzll(~m);  // the run's rows: a prefix of the window
      run_end = n < 64;
      const uint64_t pn = p + (uint64_t)n;
      if (n > 0) rload<PASS, NS>(rs, pn, lane, nxt);
      const bool valid = lane < n;
      if (n > 0) {
      // 1. bucket of the row's filter value
      uint32_t e = 0;
      if (valid) {
        if (rs.pnull && rdec_null(cur.xn, row)) {
          e = nb + 1;
        } else {
          const uint64_t x = rdec_code(cur.xr, rs.p_dt, row);
          uint32_t f = 0;
          for (int step = (nbp + 1) >> 1; step > 0; step >>= 1)
            if (s_bounds[f + step - 1] <= x) f += step;
          e = f < nb ? f : nb;
        }
      }
      // 2. histogram and prefix over buckets (lane l owns buckets 8l .. 8l+7)
      uint4 *c4 = reinterpret_cast<uint4 *>(cnt) + 2 * lane;
      c4[0] = make_uint4(0, 0, 0, 0);
      c4[1] = make_uint4(0, 0, 0, 0);
      wave_lds_sync();
      if (valid) atomicAdd(&cnt[e], 1u);
      wave_lds_sync();
      {
        const uint4 a = c4[0], b = c4[1];
        uint32_t v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        uint32_t t = 0;
#pragma unroll
        for (int g = 0; g < 8; ++g) {
          const uint32_t c = v[g];
          v[g] = (PASS == 0) ? t + c : t;  // inclusive (count pass) / exclusive (scatter turns it inclusive)
          t += c;
        }
        const uint32_t inc = wave_scan_u32(t);
        const uint32_t ex = inc - t;
        c4[0] = make_uint4(v[0] + ex, v[1] + ex, v[2] + ex, v[3] + ex);
        c4[1] = make_uint4(v[4] + ex, v[5] + ex, v[6] + ex, v[7] + ex);
      }
      wave_lds_sync();
      if constexpr (PASS == 1) {
        // counting sort: each row to its position in bucket order
        if (valid) {
          const uint32_t pos = atomicAdd(&cnt[e], 1u);
#pragma unroll
          for (int j = 0; j < NSUM; ++j) s_sum[wave][j][pos] = radd_w<NS>(rs, cur, j, row);
#pragma unroll
          for (int j = 0; j < NMAX; ++j) s_tab[wave][j][0][pos] = rmax_w<NS>(rs, cur, NSUM + j, row);
        } else {
          // positions n..63 of level 0 hold the identity (the sorted rows fill 0..n-1)
#pragma unroll
          for (int j = 0; j < NMAX; ++j) s_tab[wave][j][0][lane] = 0ull;
        }
        wave_lds_sync();
        // 3. prefix sums and sparse max tables over the sorted order
#pragma unroll
        for (int j = 0; j < NSUM; ++j) {
          const uint64_t x = wave_scan_u64(valid ? s_sum[wave][j][lane] : 0ull);
          if (valid) s_sum[wave][j][lane] = x;
        }
#pragma unroll
        for (int j = 0; j < NMAX; ++j) {
          // level l at i = max of level l-1 at i and i + 2^(l-1) (lanes past 63 keep their value)
          uint64_t t = s_tab[wave][j][0][lane];
#pragma unroll
          for (int l = 1; l < 6; ++l) {
            const int src = lane + (1 << (l - 1));
            const uint64_t u = s_tab[wave][j][l - 1][src < 64 ? src : lane];
            t = umax64(t, u);
            s_tab[wave][j][l][lane] = t;
            wave_lds_sync();
          }
        }
        wave_lds_sync();
      }
      // 4. lanes = queries: the query's span [a, b) of the sorted segment
#pragma unroll
      for (int k = 0; k < QPL; ++k) {
        if (!qon[k]) continue;
        const uint32_t a = elo[k] ? cnt[elo[k] - 1] : 0u;
        const uint32_t b = ehi[k] ? cnt[ehi[k] - 1] : 0u;
        if (b <= a) continue;
        acnt[k] += b - a;
        if constexpr (PASS == 1) {
#pragma unroll
          for (int j = 0; j < NSUM; ++j) ast[k][j] += s_sum[wave][j][b - 1] - (a ? s_sum[wave][j][a - 1] : 0ull);
          if constexpr (NMAX > 0) {
            // two windows of 2^l rows cover [a, b); level 5 (32 rows) covers a full segment
            const int l0 = 31 - __builtin_clz(b - a);
            const int l = l0 < 5 ? l0 : 5;
            const uint32_t b2 = b - (1u << l);
#pragma unroll
            for (int j = 0; j < NMAX; ++j)
              ast[k][NSUM + j] = umax64(ast[k][NSUM + j], umax64(s_tab[wave][j][l][a], s_tab[wave][j][l][b2]));
          }
        }
      }
      wave_lds_sync();
      }  // n > 0
      if (n > 0) {
        p = pn;
        cur = nxt;
      }
    }
    if (run_end) {
      // the run is complete: one group per query that selected a row of it
#pragma unroll
      for (int k = 0; k < QPL; ++k) {
        const bool has = qon[k] && acnt[k] != 0;
        if constexpr (PASS == 0) {
          groups[k] += has ? 1u : 0u;
        } else {
          // records of this run: consecutive in the (chunk, query block) range
          const uint64_t m = __ballot(has);
          // this lane's block within the wave: all 64 lanes (qbw 0 or 64) or a row of qbw lanes
          const int bw = (qbw == 0 || qbw >= 64) ? 64 : qbw;
          const uint64_t bmask = bw == 64 ? ~0ull : ((1ull << bw) - 1ull);
          const int b0 = lane & ~(bw - 1);
          const uint64_t gm = (m >> b0) & bmask;
          const uint32_t nrec = (uint32_t)__builtin_popcountll(gm);
          if (has) {
            const uint64_t below = (lane - b0) ? (gm & ((1ull << (lane - b0)) - 1ull)) : 0ull;
            const uint64_t i = sbase[k] + sdone[k] + (uint64_t)__builtin_popcountll(below);
            if (i < rs.stg_total) {  // the count pass sized every range
              rs.stg_q[i] = (uint16_t)(k * 64 + lane);
              stg_put(rs.stg_key, rs.key_w, i, (uint64_t)rk - (uint64_t)rs.kbase);
              rs.stg_pos[i] = opos[k];
              stg_put(rs.stg_val[0], rs.val_w[0], i, acnt[k]);
#pragma unroll
              for (int j = 0; j < NS; ++j) stg_put(rs.stg_val[1 + j], rs.val_w[1 + j], i, ast[k][j] - rs.vbase[1 + j]);
            }
            ++opos[k];
            nsel[k] += (uint32_t)acnt[k];
          }
          if (qbw == 0) {
#pragma unroll
            for (int k2 = 0; k2 < QPL; ++k2) sdone[k2] += nrec;  // one range for every slot
          } else {
            sdone[k] += nrec;
          }
        }
      }
      fresh = true;
    }
  }
  uint64_t gsum = 0;
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const int q = k * 64 + lane;
    if constexpr (PASS == 0) {
      if (qon[k]) rs.counts[chunk * (uint64_t)rs.nq + q] = groups[k];
      // groups of the query block: qbw lanes (qbw 0: accumulated over the wave and its slots)
      const int bw = (qbw == 0 || qbw >= 64) ? 64 : qbw;
      uint32_t t = qon[k] ? groups[k] : 0u;
      for (int off = bw >> 1; off > 0; off >>= 1) t += __shfl_xor(t, off, bw);
      if (qbw == 0) {
        gsum += t;
        if (k == QPL - 1 && lane == 0) rs.chunk_groups[chunk] = gsum;
      } else {
        const int qb = (k * 64 + lane) / qbw;
        if ((lane & (bw - 1)) == 0 && qb < rs.nqb) rs.chunk_groups[chunk * (uint64_t)rs.nqb + qb] = t;
      }
    } else {
      if (qon[k] && nsel[k]) atomicAdd((unsigned long long *)&rs.selected[q], (unsigned long long)nsel[k]);
    }
  }
}

// Place pass: a workgroup moves one chunk's staged records into the result
// columns, one window of R_WIN records at a time: coalesced loads of the
// window into LDS slots ordered by query (a counting sort), then each thread
// writes one slot, so consecutive threads write consecutive rows of one
// query's columns.
static constexpr int R_PBLOCK = 512;

template <int NS, int R_WIN>
__global__ void __launch_bounds__(R_PBLOCK) runs_place_kernel(RScan rs) {
  constexpr int NV = NS + 1;  // count + states
  constexpr int PER = R_WIN / R_PBLOCK;
  extern __shared__ __align__(16) uint8_t s_dyn[];
  uint64_t *s_key = reinterpret_cast<uint64_t *>(s_dyn);  // [R_WIN]
  uint64_t *s_val = s_key + R_WIN;                           // [NV][R_WIN]
  uint32_t *s_row = reinterpret_cast<uint32_t *>(s_val + NV * R_WIN);  // [R_WIN] row within the query's result
  uint16_t *s_qs = reinterpret_cast<uint16_t *>(s_row + R_WIN);       // [R_WIN] query of each slot
  __shared__ uint32_t s_cnt[R_MAX_QUERIES];
  __shared__ uint32_t s_wsum[R_PBLOCK / 64];
  __shared__ uint64_t s_roff[R_MAX_QUERIES], s_rend[R_MAX_QUERIES];  // each query's result rows [roff, rend)
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  // workgroups in query-block-major order: the ones running together write
  // the columns of the same few queries
  const uint64_t qb = blockIdx.x / rs.nchunks, c = blockIdx.x % rs.nchunks;
  const uint64_t b0 = rs.stg_base[c * rs.nqb + qb], b1 = rs.stg_base[c * rs.nqb + qb + 1];
  const uint32_t nq = (uint32_t)rs.nq;
  if (tid < R_MAX_QUERIES) {
    s_roff[tid] = tid < (int)nq ? rs.row_off[tid] : 0;
    s_rend[tid] = tid < (int)nq ? rs.row_off[tid] + rs.totals[tid] : 0;
  }
  // the records of one window, one window ahead (loads in flight while the
  // previous window is sorted and written)
  // raw staged words (decoded at the scatter into LDS)
  // (loaded unconditionally from a clamped index, `in` = the slot holds a
  // record: a fixed number of loads per window lets each buffer's loads be
  // waited for exactly where they are used)
  struct Rec { uint64_t key, v[NV]; uint32_t pos; uint16_t q; bool in; };
  Rec cur[PER], nxt[PER];
  const int32_t ksh = rs.key_w == 4 ? 2 : 3;
  int32_t vsh[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) vsh[k] = rs.val_w[k] == 4 ? 2 : 3;
  auto load = [&](uint64_t w0, Rec (&r)[PER]) {
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const uint64_t i0 = w0 + (uint64_t)(tid + t * R_PBLOCK);
      r[t].in = i0 < b1;
      const uint64_t i = r[t].in ? i0 : b1 - 1;  // b1 > w0 >= b0: a staged record
      r[t].q = rs.stg_q[i];
      r[t].key = stg_raw(rs.stg_key, ksh, i);
      r[t].pos = rs.stg_pos[i];
#pragma unroll
      for (int k = 0; k < NV; ++k) r[t].v[k] = stg_raw(rs.stg_val[k], vsh[k], i);
    }
  };
  // one window: counting sort of its records by query in LDS, then the
  // result columns; windows alternate between two record buffers, so each
  // buffer's loads stay in flight for a whole window (no copy between them)
  auto process = [&](uint64_t w0, Rec (&cur)[PER]) {
    const int n = (int)((b1 - w0 < (uint64_t)R_WIN) ? (b1 - w0) : (uint64_t)R_WIN);
    // every lane takes this buffer's words here, loaded one window ago (the
    // uses below are conditional; without this the compiler cannot tell
    // they completed and waits for all loads, the next window's included,
    // where their registers are reused)
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      __asm__ volatile("" ::"v"(cur[t].q), "v"(cur[t].pos), "v"(cur[t].key));
#pragma unroll
      for (int k = 0; k < NV; ++k) __asm__ volatile("" ::"v"(cur[t].v[k]));
    }
    if (tid < R_MAX_QUERIES) s_cnt[tid] = 0;
    __syncthreads();
#pragma unroll
    for (int t = 0; t < PER; ++t)
      if (cur[t].in) atomicAdd(&s_cnt[cur[t].q < nq ? cur[t].q : 0u], 1u);
    __syncthreads();
    {  // exclusive scan over the queries (thread t < R_MAX_QUERIES owns query t)
      const uint32_t v = tid < R_MAX_QUERIES ? s_cnt[tid] : 0u;
      const uint32_t inc = wave_scan_u32(v);
      if (lane == 63) s_wsum[wave] = inc;
      __syncthreads();
      uint32_t base = 0;
      for (int w = 0; w < wave; ++w) base += s_wsum[w];
      if (tid < R_MAX_QUERIES) s_cnt[tid] = base + inc - v;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      if (!cur[t].in) continue;
      const uint32_t q = cur[t].q < nq ? cur[t].q : 0u;
      const uint32_t slot = atomicAdd(&s_cnt[q], 1u);
      s_qs[slot] = (uint16_t)q;
      s_key[slot] = stg_dec(cur[t].key, ksh, (uint64_t)rs.kbase);
      s_row[slot] = cur[t].pos;
#pragma unroll
      for (int k = 0; k < NV; ++k) s_val[k * R_WIN + slot] = stg_dec(cur[t].v[k], vsh[k], rs.vbase[k]);
    }
    __syncthreads();
    // one output column at a time over this thread's slots (the output's
    // kind and type are decided once per column, not per row)
    int qq[PER];
    uint64_t g[PER];
    bool ok[PER];
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int j = tid + t * R_PBLOCK;
      ok[t] = j < n;
      qq[t] = ok[t] ? s_qs[j] : 0;
      const uint64_t pos = ok[t] ? s_row[j] : 0;  // row within the query's result
      g[t] = s_roff[qq[t]] + pos;
      ok[t] = ok[t] && g[t] < s_rend[qq[t]];  // the write pass staged every record
    }
    for (int o = 0; o < rs.nouts; ++o) {
      const ROut ro = rs.outs[o];
      void *dst = rs.out_base[o];
      uint32_t *const nbm = rs.out_nbase[o];
      const int vs = ro.state, cs = ro.cnt_state;
#pragma unroll
      for (int t = 0; t < PER; ++t) {
        if (!ok[t]) continue;
        const int j = tid + t * R_PBLOCK;
        const uint64_t cnt = cs < 0 ? 0 : s_val[cs * R_WIN + j];
        const uint64_t v = s_val[vs * R_WIN + j];
        const uint64_t gg = g[t];
        bool isnull = false;
        switch (ro.kind) {
          case OK_KEY: rstore_i(dst, ro.out_dtype, gg, (int64_t)s_key[j]); break;
          case OK_COUNT: static_cast<int64_t *>(dst)[gg] = (int64_t)cnt; break;
          case OK_SUM_I64:
            isnull = cnt == 0;
            static_cast<int64_t *>(dst)[gg] = isnull ? 0 : (int64_t)v;
            break;
          case OK_AVG_I64:
            isnull = cnt == 0;
            static_cast<double *>(dst)[gg] = isnull ? 0.0 : (double)(int64_t)v / (double)cnt;
            break;
          default: {  // OK_MIN / OK_MAX (the values finalize_slot derives)
            const bool mn = ro.kind == OK_MIN;
            if (cnt == 0) {
              isnull = true;
              if (ro.src_is_float) rstore_f(dst, ro.out_dtype, gg, 0.0);
              else rstore_i(dst, ro.out_dtype, gg, 0);
            } else if (ro.src_is_float) {
              double d;
              if (v == 0) d = mn ? bits_f64(0x7FF0000000000000ull) : bits_f64(0xFFF0000000000000ull);
              else d = bits_f64(decode_f64_code(mn ? ~v : v));
              rstore_f(dst, ro.out_dtype, gg, d);
            } else {
              rstore_i(dst, ro.out_dtype, gg, decode_i64(mn ? ~v : v));
            }
            break;
          }
        }
        if (isnull && nbm) atomicOr(&nbm[gg >> 5], 1u << (gg & 31));
      }
    }
    __syncthreads();
  };
  // the next window is always loaded (past the end: the last record again,
  // unused), so every path issues the same loads
  if (b0 >= b1) return;
  load(b0, cur);
  for (uint64_t w0 = b0; w0 < b1; w0 += 2 * (uint64_t)R_WIN) {
    load(w0 + R_WIN, nxt);
    process(w0, cur);
    if (w0 + R_WIN >= b1) break;
    load(w0 + 2 * (uint64_t)R_WIN, cur);
    process(w0 + R_WIN, nxt);
  }
}

static size_t runs_place_lds(int ns, int win) { return (size_t)win * (8 * (1 + ns + 1) + 4 + 2); }

// exclusive scan over chunks, per query: a block serves 64 queries (lanes),
// its 16 waves one slice of the chunks each
__global__ void __launch_bounds__(1024) runs_offsets_kernel(const uint32_t *__restrict__ counts, uint32_t nchunks,
                                                            uint32_t nq, uint64_t *offsets, uint64_t *totals) {
  __shared__ uint64_t part[16][64];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const uint32_t q = blockIdx.x * 64 + lane;
  const uint32_t per = (nchunks + 15) / 16;
  const uint32_t a = w * per < nchunks ? w * per : nchunks;
  const uint32_t b = a + per < nchunks ? a + per : nchunks;
  uint64_t s = 0;
  if (q < nq)
    for (uint32_t c = a; c < b; ++c) s += counts[(uint64_t)c * nq + q];
  part[w][lane] = s;
  __syncthreads();
  uint64_t base = 0;
  for (int i = 0; i < w; ++i) base += part[i][lane];
  if (q < nq) {
    for (uint32_t c = a; c < b; ++c) {
      offsets[(uint64_t)c * nq + q] = base;
      base += counts[(uint64_t)c * nq + q];
    }
    if (w == 15) totals[q] = base;
  }
}

// ---------------------------------------------------------------------------
// Lean shared-interval pass (SScan). Each workgroup owns a contiguous row
// range; a thread takes RPL rows of it per step (each a row per lane,
// coalesced), with the next step's raw words in flight while the current
// rows are applied: interval of the range value (binary search over the
// bounds in LDS, the rows' searches interleaved), then each state's
// contribution into the (interval, key) entry with LDS atomics.
// Rows whose range value is NULL match no query of the group.
template <int NSV>
struct SRow {
  bool ok;
  uint64_t kr;
  uint64_t vr[NSV];
  uint32_t vn[NSV];
};

template <int NSV>
__device__ __forceinline__ void sload(const SScan &ss, uint64_t row, uint64_t r1, SRow<NSV> &w) {
  w.ok = row < r1 && !page_bit(ss.page_skip, row);  // a page no query can match is not read
  w.kr = 0;
#pragma unroll
  for (int j = 0; j < NSV; ++j) { w.vr[j] = 0; w.vn[j] = 0; }
  if (w.ok) {
    if (ss.key) w.kr = rraw(ss.key, ss.key_dt, row);
#pragma unroll
    for (int j = 0; j < NSV; ++j) {
      if (j >= ss.nv) break;
      w.vr[j] = rraw(ss.vcol[j], ss.v_dt[j], row);
      w.vn[j] = rraw_null(ss.vnull[j], row);
    }
  }
}

// RPL rows per thread per step: 2 where the bound search is deep (many
// ranges; the entry table in LDS then bounds the occupancy anyway), else 1
// (the registers of a second row would cost waves)
template <int NSV, int RPL>
__global__ void __launch_bounds__(S_BLOCK) shared_kernel(SScan ss) {
  extern __shared__ __align__(16) uint8_t s_dyn[];
  const int nbp = (1 << ss.bsteps) - 1;
  uint64_t *sb = reinterpret_cast<uint64_t *>(s_dyn);     // padded bounds
  lds_u64 *tab = (lds_u64 *)(sb + nbp + 1);               // (interval, key) entries
  const uint32_t W = (uint32_t)ss.entry_words;
  const uint32_t nent = (uint32_t)(ss.nbounds + 1) * ss.dom;
  for (int i = threadIdx.x; i < nbp; i += S_BLOCK) sb[i] = ss.bounds[i];
  for (uint32_t i = threadIdx.x; i < nent * W; i += S_BLOCK) tab[i] = 0;
  __syncthreads();
  const uint64_t r0 = (uint64_t)blockIdx.x * ss.rows_per_block;
  const uint64_t r1 = (r0 + ss.rows_per_block < ss.nrows) ? r0 + ss.rows_per_block : ss.nrows;
  const uint32_t nb = (uint32_t)ss.nbounds;
  // a thread takes RPL rows per step (S_BLOCK apart, so each load is
  // coalesced): their bound searches interleave, and each state's kind is
  // dispatched once for all of them
  constexpr uint64_t STEP = (uint64_t)S_BLOCK * RPL;
  SRow<NSV> cur[RPL], nxt[RPL];
  uint64_t row = r0 + threadIdx.x;
#pragma unroll
  for (int j = 0; j < RPL; ++j) sload<NSV>(ss, row + (uint64_t)j * S_BLOCK, r1, cur[j]);
  for (; row < r1; row += STEP) {
#pragma unroll
    for (int j = 0; j < RPL; ++j) sload<NSV>(ss, row + STEP + (uint64_t)j * S_BLOCK, r1, nxt[j]);
    bool ok[RPL];
    uint64_t x[RPL];
    uint32_t f[RPL];
#pragma unroll
    for (int j = 0; j < RPL; ++j) {
      const uint64_t rj = row + (uint64_t)j * S_BLOCK;
      ok[j] = cur[j].ok && !(ss.vnull[0] && rdec_null(cur[j].vn[0], rj));
      x[j] = ok[j] ? rdec_code(cur[j].vr[0], ss.v_dt[0], rj) : 0ull;
      f[j] = 0;
    }
    for (int step = (nbp + 1) >> 1; step > 0; step >>= 1) {
#pragma unroll
      for (int j = 0; j < RPL; ++j)
        if (sb[f[j] + step - 1] <= x[j]) f[j] += step;
    }
    lds_u64 *ent[RPL];
#pragma unroll
    for (int j = 0; j < RPL; ++j) {
      const uint64_t rj = row + (uint64_t)j * S_BLOCK;
      const uint32_t e = f[j] < nb ? f[j] : nb;
      uint32_t d = 0;
      if (ss.key && ok[j]) {
        const int64_t kv = rdec_i(cur[j].kr, ss.key_dt, rj) - ss.kmin;
        d = (kv >= 0 && kv < (int64_t)ss.dom) ? (uint32_t)kv : 0u;  // the key's statistics bound it
      }
      ent[j] = tab + ((uint64_t)e * ss.dom + d) * W;
    }
#pragma unroll
    for (int si = 0; si < B_MAX_SHARED_STATES; ++si) {
      if (si >= ss.nstates) break;
      const DState &st = ss.st[si];
      const int vi = ss.st_v[si];
#pragma unroll
      for (int j = 0; j < RPL; ++j) {
        if (!ok[j]) continue;
        const uint64_t rj = row + (uint64_t)j * S_BLOCK;
        if (si == 0 && st.kind == SK_COUNT && vi < 0) {
          // COUNT(*) (always state 0): a 32-bit add into the low half of its
          // entry word — a workgroup counts fewer than 2^32 rows
          // (qmk_shared_scan), so the high half stays 0
          __hip_atomic_fetch_add(reinterpret_cast<__attribute__((address_space(3))) uint32_t *>(ent[j] + st.word), 1u,
                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          continue;
        }
        Contrib k;
        k.a = 1; k.b = k.c = 0; k.skip = false;
        if (vi >= 0) {
          uint64_t raw = 0;
          uint32_t rn = 0;
          int32_t dt = 0;
#pragma unroll
          for (int v = 0; v < NSV; ++v)
            if (vi == v) { raw = cur[j].vr[v]; rn = cur[j].vn[v]; dt = ss.v_dt[v]; }
          if (ss.vnull[vi] && rdec_null(rn, rj)) continue;
          if (st.kind != SK_COUNT) {
            if (dt_is_float(dt)) k = contrib_of(st, true, 0, rdec_f(raw, dt));
            else k = contrib_of(st, false, rdec_i(raw, dt, rj), 0.0);
            if (k.skip) continue;
          }
        }
        apply_lds(st, ent[j], k);
      }
    }
#pragma unroll
    for (int j = 0; j < RPL; ++j) cur[j] = nxt[j];
  }
  __syncthreads();
  // merge into the global table: one entry per thread, its states in order
  for (uint32_t en = threadIdx.x; en < nent; en += S_BLOCK) {
    const lds_u64 *src = tab + (uint64_t)en * W;
    if (src[ss.st[0].word] == 0) continue;  // state 0 is COUNT(*): no row reached the entry
    uint64_t *dst = ss.table + (uint64_t)en * W;
    for (int si = 0; si < ss.nstates; ++si) {
      const DState &st = ss.st[si];
      const uint32_t w = (uint32_t)st.word;
      switch (st.kind) {
        case SK_COUNT: case SK_SUM_I64: if (src[w]) g_atomic_add(dst + w, src[w]); break;
        case SK_SUM_128:
          g_add128(dst + w, src[w], src[w + 1]);
          if (st.fixed && src[w + 2]) g_atomic_or(dst + w + 2, src[w + 2]);
          break;
        default: if (src[w]) g_atomic_max(dst + w, src[w]); break;
      }
    }
  }
}

// rows per thread per step for a launch (deep bound searches: 2)
static int shared_rpl(const SScan *ss) { return ss->bsteps >= 4 ? 2 : 1; }

static const void *shared_fn(int nv, int rpl) {
  switch (nv * 2 + (rpl == 2 ? 1 : 0)) {
    case 2: return (const void *)shared_kernel<1, 1>;
    case 3: return (const void *)shared_kernel<1, 2>;
    case 4: return (const void *)shared_kernel<2, 1>;
    case 5: return (const void *)shared_kernel<2, 2>;
    case 6: return (const void *)shared_kernel<3, 1>;
    case 7: return (const void *)shared_kernel<3, 2>;
    case 8: return (const void *)shared_kernel<4, 1>;
    case 9: return (const void *)shared_kernel<4, 2>;
  }
  return nullptr;
}

// ---------------------------------------------------------------------------
// One query grouped by a sorted key (S1Scan).
struct S1Win {
  uint64_t p;
  bool ok;
  uint64_t kr;
  uint64_t pr[B_MAX_SLOTS];
  uint32_t pn[B_MAX_SLOTS];
  uint64_t sr[R_MAX_STATES];
  uint32_t sn[R_MAX_STATES];
};

__device__ __forceinline__ void s1load(const S1Scan &q, uint64_t p, int lane, S1Win &w) {
  const uint64_t row = p + (uint64_t)lane;
  w.p = p;
  w.ok = row < q.nrows;
  w.kr = 0;
#pragma unroll
  for (int j = 0; j < B_MAX_SLOTS; ++j) { w.pr[j] = 0; w.pn[j] = 0; }
#pragma unroll
  for (int j = 0; j < R_MAX_STATES; ++j) { w.sr[j] = 0; w.sn[j] = 0; }
  if (!w.ok) return;
  w.kr = rraw(q.key, q.key_dt, row);
#pragma unroll
  for (int j = 0; j < B_MAX_SLOTS; ++j) {
    if (j >= q.nslots) break;
    w.pr[j] = rraw(q.pcol[j], q.p_dt[j], row);
    w.pn[j] = rraw_null(q.pnull[j], row);
  }
#pragma unroll
  for (int j = 1; j < R_MAX_STATES; ++j) {
    if (j >= q.nstates) break;
    if (q.scol[j]) w.sr[j] = rraw(q.scol[j], q.s_dt[j], row);
    w.sn[j] = rraw_null(q.snull[j], row);
  }
}

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t x) {
  // inclusive max scan (out-of-row sources read 0, the identity), lane 63 holds the max
  x = umax64(x, dpp64<0x111, 0xF>(x));
  x = umax64(x, dpp64<0x112, 0xF>(x));
  x = umax64(x, dpp64<0x114, 0xF>(x));
  x = umax64(x, dpp64<0x118, 0xF>(x));
  x = umax64(x, dpp64<0x142, 0xA>(x));
  x = umax64(x, dpp64<0x143, 0xC>(x));
  return (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, 63) |
         ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), 63) << 32);
}
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x) {
  x = wave_scan_u64(x);
  return (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, 63) |
         ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), 63) << 32);
}

__global__ void __launch_bounds__(R_BLOCK) sorted1_kernel(S1Scan q) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t chunk = (uint64_t)blockIdx.x * R_WAVES + wave;
  if (chunk >= q.nchunks || q.nslots < 0) return;
  const uint64_t c0 = chunk * q.chunk_rows;
  const uint64_t c1 = (c0 + q.chunk_rows < q.nrows) ? c0 + q.chunk_rows : q.nrows;
  uint64_t s = c0;
  if (c0 > 0 && c0 < c1) {  // skip the run that started in the previous chunk
    const int64_t prev = rdec_i(rraw(q.key, q.key_dt, c0 - 1), q.key_dt, c0 - 1);
    for (;;) {
      const uint64_t row = s + lane;
      const bool diff = row >= q.nrows || rdec_i(rraw(q.key, q.key_dt, row), q.key_dt, row) != prev;
      const uint64_t m = __ballot(diff);
      if (m) { s += (uint64_t)__builtin_ctzll(m); break; }
      s += 64;
    }
  }
  const int W = q.entry_words;
  uint64_t p = s, nsel = 0;
  S1Win cur, nxt;
  if (p < c1) s1load(q, p, lane, cur);
  bool fresh = true;
  int64_t rk = 0;
  uint64_t acc[R_MAX_STATES];  // wave-uniform run accumulators
  while (true) {
    if (fresh) {
      if (p >= c1) break;
      rk = (int64_t)rfirst64((uint64_t)rdec_i(cur.kr, q.key_dt, cur.p + lane));
#pragma unroll
      for (int j = 0; j < R_MAX_STATES; ++j) acc[j] = 0;
      fresh = false;
    }
    const uint64_t row = cur.p + (uint64_t)lane;
    const bool inr = cur.ok && rdec_i(cur.kr, q.key_dt, row) == rk;
    const uint64_t m = __ballot(inr);
    const int n = (~m == 0ull) ? 64 : __builtin_ctzll(~m);
    const bool run_end = n < 64;
    if (n > 0) {
      const uint64_t pn = p + (uint64_t)n;
      s1load(q, pn, lane, nxt);
      bool sel = lane < n;
#pragma unroll
      for (int j = 0; j < B_MAX_SLOTS; ++j) {
        if (j >= q.nslots) break;
        if (q.pnull[j] && rdec_null(cur.pn[j], row)) sel = false;
        sel = sel && (rdec_code(cur.pr[j], q.p_dt[j], row) - q.lo[j]) <= q.span[j];
      }
      const uint64_t sm = __ballot(sel);
      if (sm) {
        acc[0] += (uint64_t)__builtin_popcountll(sm);
#pragma unroll
        for (int j = 1; j < R_MAX_STATES; ++j) {
          if (j >= q.nstates) break;
          const DState &st = q.st[j];
          uint64_t v = 0;
          if (sel && !(q.snull[j] && rdec_null(cur.sn[j], row))) {
            if (st.kind == SK_COUNT) {
              v = 1;
            } else {
              Contrib k;
              if (dt_is_float(q.s_dt[j])) k = contrib_of(st, true, 0, rdec_f(cur.sr[j], q.s_dt[j]));
              else k = contrib_of(st, false, rdec_i(cur.sr[j], q.s_dt[j], row), 0.0);
              v = k.skip ? 0 : k.a;
            }
          }
          if (st.kind == SK_MAX_CODE) acc[j] = umax64(acc[j], wave_max_u64(v));
          else acc[j] += wave_sum_u64(v);
        }
      }
      p = pn;
      cur = nxt;
    }
    if (run_end) {
      if (acc[0]) {
        nsel += acc[0];
        if (lane < q.nstates) {  // one state word per lane, plain stores (this wave owns the run)
          uint64_t v = acc[0];
#pragma unroll
          for (int j = 1; j < R_MAX_STATES; ++j) v = lane == j ? acc[j] : v;
          q.table[(uint64_t)(rk - q.kmin) * (uint64_t)W + (uint64_t)q.st[lane].word] = v;
        }
      }
      fresh = true;
    }
  }
  if (lane == 0 && nsel) atomicAdd((unsigned long long *)q.selected, (unsigned long long)nsel);
}

#define R_FN(P, S, M)                                                                   \
  case (S) * 4 + (M):                                                                   \
    return qpl == 1 ? (const void *)runs_kernel<P, S, M, 1>                             \
                    : (qpl == 2 ? (const void *)runs_kernel<P, S, M, 2> : (const void *)runs_kernel<P, S, M, 4>);

static const void *runs_write_fn(int nsum, int nmax, int qpl) {
  switch (nsum * 4 + nmax) {
    R_FN(1, 0, 0)
    R_FN(1, 0, 1)
    R_FN(1, 0, 2)
    R_FN(1, 0, 3)
    R_FN(1, 1, 0)
    R_FN(1, 1, 1)
    R_FN(1, 1, 2)
    R_FN(1, 2, 0)
    R_FN(1, 2, 1)
    R_FN(1, 3, 0)
  }
  return nullptr;
}
#undef R_FN

static int runs_qpl(int nq) { return nq <= 64 ? 1 : (nq <= 128 ? 2 : 4); }

template <int WIN>
static const void *runs_place_fn_w(int ns) {
  switch (ns) {
    case 0: return (const void *)runs_place_kernel<0, WIN>;
    case 1: return (const void *)runs_place_kernel<1, WIN>;
    case 2: return (const void *)runs_place_kernel<2, WIN>;
    case 3: return (const void *)runs_place_kernel<3, WIN>;
  }
  return nullptr;
}
static const void *runs_place_fn(int ns, int win) {
  return win >= 2048 ? runs_place_fn_w<2048>(ns) : runs_place_fn_w<1024>(ns);
}

static bool runs_args_ok(const RScan *rs) {
  return rs->nq >= 1 && rs->nq <= R_MAX_QUERIES && rs->nbounds >= 0 && rs->nbounds + 2 <= R_MAX_BUCKETS &&
         rs->bsteps >= 0 && rs->bsteps <= 9 && ((1 << rs->bsteps) - 1) >= rs->nbounds && rs->nchunks >= 1 &&
         rs->chunk_rows >= 1 && rs->nsum >= 0 && rs->nsum <= rs->nstates - 1 && rs->nstates >= 1 &&
         rs->nstates <= R_MAX_STATES && rs->nouts >= 1 && rs->nouts <= R_MAX_OUTS &&
         (uint64_t)rs->nchunks * rs->chunk_rows >= rs->nrows &&
         (rs->qbw == 0 || rs->qbw == 16 || rs->qbw == 32 || rs->qbw == 64) &&
         rs->nqb == (rs->qbw ? (rs->nq + rs->qbw - 1) / rs->qbw : 1);
}

}  // namespace qm

using namespace qm;

extern "C" int qmk_runs_supported(int nsum, int nmax) {
  return nsum >= 0 && nmax >= 0 && 1 + nsum + nmax <= R_MAX_STATES && runs_write_fn(nsum, nmax, 1) != nullptr;
}

extern "C" hipError_t qmk_runs_count(const RScan *rs, hipStream_t st) {
  if (!runs_args_ok(rs)) return hipErrorInvalidValue;
  const int qpl = runs_qpl(rs->nq);
  const void *fn = qpl == 1 ? (const void *)runs_kernel<0, 0, 0, 1>
                            : (qpl == 2 ? (const void *)runs_kernel<0, 0, 0, 2> : (const void *)runs_kernel<0, 0, 0, 4>);
  RScan a = *rs;
  void *args[] = {&a};
  const uint32_t grid = (rs->nchunks + R_WAVES - 1) / R_WAVES;
  return hipLaunchKernel(fn, dim3(grid), dim3(R_BLOCK), args, 0, st);
}

extern "C" hipError_t qmk_runs_offsets(const uint32_t *counts, uint32_t nchunks, uint32_t nq, uint64_t *offsets,
                                       uint64_t *totals, hipStream_t st) {
  if (nq == 0 || nchunks == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(runs_offsets_kernel, dim3((nq + 63) / 64), dim3(1024), 0, st, counts, nchunks, nq, offsets,
                     totals);
  return hipGetLastError();
}

// lds_pad: extra dynamic LDS per workgroup, which caps the waves in flight
// (fewer concurrent output streams per cache)
extern "C" hipError_t qmk_runs_write(const RScan *rs, uint32_t lds_pad, hipStream_t st) {
  if (!runs_args_ok(rs)) return hipErrorInvalidValue;
  const void *fn = runs_write_fn(rs->nsum, rs->nstates - 1 - rs->nsum, runs_qpl(rs->nq));
  if (!fn) return hipErrorInvalidValue;
  if (lds_pad) {
    hipError_t e = qmk_func_lds(fn, (size_t)lds_pad);
    if (e != hipSuccess) return e;
  }
  RScan a = *rs;
  void *args[] = {&a};
  const uint32_t grid = (rs->nchunks + R_WAVES - 1) / R_WAVES;
  return hipLaunchKernel(fn, dim3(grid), dim3(R_BLOCK), args, lds_pad, st);
}

// win: records per window (1024 or 2048; larger windows give longer
// contiguous runs per query, smaller ones more workgroups per CU)
extern "C" hipError_t qmk_runs_place(const RScan *rs, int win, hipStream_t st) {
  if (!runs_args_ok(rs)) return hipErrorInvalidValue;
  const int ns = rs->nstates - 1;
  win = win >= 2048 ? 2048 : 1024;
  const void *fn = runs_place_fn(ns, win);
  if (!fn) return hipErrorInvalidValue;
  const size_t lds = runs_place_lds(ns, win);
  hipError_t e = qmk_func_lds(fn, (size_t)lds);
  if (e != hipSuccess) return e;
  RScan a = *rs;
  void *args[] = {&a};
  return hipLaunchKernel(fn, dim3(rs->nchunks * (uint32_t)rs->nqb), dim3(R_PBLOCK), args, lds, st);
}

extern "C" size_t qmk_shared_lds_bytes(const SScan *ss) {
  const uint64_t nbp = (1ull << ss->bsteps) - 1;
  return (size_t)(nbp + 1 + (uint64_t)(ss->nbounds + 1) * ss->dom * (uint64_t)ss->entry_words) * 8;
}

extern "C" int qmk_shared_blocks_per_cu(const SScan *ss) {
  const void *fn = shared_fn(ss->nv, shared_rpl(ss));
  if (!fn) return 0;
  const size_t lds = qmk_shared_lds_bytes(ss);
  if (lds > 64 * 1024) (void)qmk_func_lds(fn, (size_t)lds);
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, S_BLOCK, lds) != hipSuccess) nb = 1;
  return nb < 1 ? 1 : nb;
}

extern "C" hipError_t qmk_shared_scan(const SScan *ss, uint32_t grid, hipStream_t st) {
  const void *fn = shared_fn(ss->nv, shared_rpl(ss));
  if (!fn || grid == 0 || ss->nstates < 1 || ss->nstates > B_MAX_SHARED_STATES || ss->dom < 1 ||
      ss->bsteps < 0 || ss->bsteps > 10 || ((1 << ss->bsteps) - 1) < ss->nbounds ||
      (uint64_t)grid * ss->rows_per_block < ss->nrows || ss->rows_per_block >= (1ull << 32))  // 32-bit COUNT(*) in LDS
    return hipErrorInvalidValue;
  for (int i = 0; i < ss->nstates; ++i)
    if (ss->st_v[i] >= ss->nv) return hipErrorInvalidValue;
  const size_t lds = qmk_shared_lds_bytes(ss);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  if (lds > 64 * 1024) {
    hipError_t e = qmk_func_lds(fn, (size_t)lds);
    if (e != hipSuccess) return e;
  }
  SScan a = *ss;
  void *args[] = {&a};
  return hipLaunchKernel(fn, dim3(grid), dim3(S_BLOCK), args, lds, st);
}

extern "C" hipError_t qmk_sorted1_scan(const S1Scan *q, hipStream_t st) {
  if (q->nchunks < 1 || q->nstates < 1 || q->nstates > R_MAX_STATES || q->nslots > B_MAX_SLOTS ||
      (uint64_t)q->nchunks * q->chunk_rows < q->nrows || q->entry_words < q->nstates)
    return hipErrorInvalidValue;
  for (int j = 0; j < q->nstates; ++j)
    if (q->st[j].kind == SK_SUM_128 || q->st[j].word < 0 || q->st[j].word >= q->entry_words) return hipErrorInvalidValue;
  S1Scan a = *q;
  void *args[] = {&a};
  return hipLaunchKernel((const void *)sorted1_kernel, dim3((q->nchunks + R_WAVES - 1) / R_WAVES), dim3(R_BLOCK),
                         args, 0, st);
}

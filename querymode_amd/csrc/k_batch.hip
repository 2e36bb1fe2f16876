// k_batch.hip — many aggregate queries over one resident table in one pass.
//
// One persistent launch reads every referenced column once. A wave owns a
// contiguous range of 512-row tiles (64 lanes x 8 consecutive rows); per tile
// it turns each loaded value into its order-preserving unsigned code, then
// serves two kinds of queries:
//
//  * shared-interval groups: queries whose filter is one range on the same
//    column and whose key is a small dense domain (or absent). The rows are
//    aggregated once by (elementary interval of that column, key) in an LDS
//    table; each query's result is afterwards the merge of the intervals its
//    range covers (SUM / COUNT / MIN / MAX decompose exactly over disjoint
//    intervals). The per-row cost does not grow with the number of queries.
//
//  * sorted-key queries: GROUP BY an integer column stored in ascending
//    order. Each query evaluates its AND-ed ranges (one unsigned compare
//    each), then reduces the selected rows over the key's runs: runs inside a
//    lane are reduced in registers, runs across lanes by a segmented scan,
//    runs across tiles through a per-wave carry in LDS. A run closed inside a
//    wave's range is written once with plain stores into the query's DIRECT
//    table; only the first and the last run of a range (which the
//    neighbouring range may share) are merged with atomics.
//
// The reference runs one operator pipeline per query (src/operators.ts:
// 3073-3179); the aggregate semantics are those of src/partial-agg.ts.
#include "k_agg.hip.h"
#include "qm_kernels.h"

namespace qm {

static constexpr int B_TILE = 64 * RPT;  // rows per wave tile

// ---- order codes of 8 rows of one column ----
__device__ __forceinline__ void load8_codes(const void *data, int32_t dtype, uint64_t row0, uint64_t v[RPT]) {
  DCol c;
  c.data = data; c.nulls = nullptr; c.dtype = dtype; c.side = 0;
  if (dt_is_float(dtype)) {
    double f[RPT];
    load8_f(c, row0, f);
#pragma unroll
    for (int r = 0; r < RPT; ++r) v[r] = code_f64_bits(f64_bits(f[r]));
  } else {
    int64_t x[RPT];
    load8_i(c, row0, x);
#pragma unroll
    for (int r = 0; r < RPT; ++r) v[r] = code_i64(x[r]);
  }
}

__device__ __forceinline__ int64_t code_int(uint64_t c) { return decode_i64(c); }
__device__ __forceinline__ double code_flt(uint64_t c) { return bits_f64(decode_f64_code(c)); }

// ---- register accumulators of one run (states of a sorted-key query) ----
struct Acc {
  uint64_t lo[B_MAX_STATES];
  uint64_t hi[B_MAX_STATES];
  uint64_t fl;  // SK_SUM_128 fixed: NaN / +Inf / -Inf flags, 8 bits per state
};

__device__ __forceinline__ void acc_zero(Acc &a) {
#pragma unroll
  for (int s = 0; s < B_MAX_STATES; ++s) { a.lo[s] = 0; a.hi[s] = 0; }
  a.fl = 0;
}

__device__ __forceinline__ void acc_merge(Acc &a, const Acc &b, const BQuery &q) {
#pragma unroll
  for (int s = 0; s < B_MAX_STATES; ++s) {
    if (s >= q.nstates) break;
    const uint8_t k = q.st[s].kind;
    if (k == SK_MAX_CODE) {
      a.lo[s] = a.lo[s] > b.lo[s] ? a.lo[s] : b.lo[s];
    } else if (k == SK_SUM_128) {
      uint64_t lo = a.lo[s] + b.lo[s];
      a.hi[s] = a.hi[s] + b.hi[s] + (lo < a.lo[s] ? 1 : 0);
      a.lo[s] = lo;
    } else {
      a.lo[s] += b.lo[s];
    }
  }
  a.fl |= b.fl;
}

// add one row (column codes c[] of the row) to an accumulator
template <int NC>
__device__ __forceinline__ void acc_row(Acc &a, const BQuery &q, const uint64_t (&c)[NC], const int32_t *dtype,
                                        uint32_t nullbits) {
#pragma unroll
  for (int s = 0; s < B_MAX_STATES; ++s) {
    if (s >= q.nstates) break;
    const DState &st = q.st[s];
    const uint8_t col = q.st_col[s];
    if (col == 0xFF) { a.lo[s] += 1; continue; }  // COUNT(*)
    uint64_t code = 0;
    bool isf = false;
#pragma unroll
    for (int j = 0; j < NC; ++j)
      if (col == j) { code = c[j]; isf = dt_is_float(dtype[j]); }
    if ((nullbits >> col) & 1) continue;
    Contrib k = contrib_of(st, isf, isf ? 0 : code_int(code), isf ? code_flt(code) : 0.0);
    if (k.skip) continue;
    if (st.kind == SK_MAX_CODE) {
      a.lo[s] = a.lo[s] > k.a ? a.lo[s] : k.a;
    } else if (st.kind == SK_SUM_128) {
      uint64_t lo = a.lo[s] + k.a;
      a.hi[s] = a.hi[s] + k.b + (lo < a.lo[s] ? 1 : 0);
      a.lo[s] = lo;
      a.fl |= k.c << (8 * s);
    } else {
      a.lo[s] += k.a;
    }
  }
}

__device__ __forceinline__ Acc acc_shfl_up(const Acc &a, int off, const BQuery &q) {
  Acc r;
#pragma unroll
  for (int s = 0; s < B_MAX_STATES; ++s) {
    r.lo[s] = 0; r.hi[s] = 0;
    if (s < q.nstates) {
      r.lo[s] = __shfl_up(a.lo[s], off, 64);
      if (q.st[s].kind == SK_SUM_128) r.hi[s] = __shfl_up(a.hi[s], off, 64);
    }
  }
  r.fl = __shfl_up(a.fl, off, 64);
  return r;
}

__device__ __forceinline__ Acc acc_shfl(const Acc &a, int src, const BQuery &q) {
  Acc r;
#pragma unroll
  for (int s = 0; s < B_MAX_STATES; ++s) {
    r.lo[s] = 0; r.hi[s] = 0;
    if (s < q.nstates) {
      r.lo[s] = __shfl(a.lo[s], src, 64);
      if (q.st[s].kind == SK_SUM_128) r.hi[s] = __shfl(a.hi[s], src, 64);
    }
  }
  r.fl = __shfl(a.fl, src, 64);
  return r;
}

// write a closed run of key code `kc` into the query's DIRECT table
__device__ __forceinline__ void acc_flush(const BQuery &q, uint64_t kc, const Acc &a, bool atomic) {
  if (a.lo[0] == 0) return;  // state 0 is COUNT(*): no selected row in the run
  uint64_t *e = q.table + (uint64_t)(code_int(kc) - q.kmin) * (uint64_t)q.entry_words;
#pragma unroll
  for (int s = 0; s < B_MAX_STATES; ++s) {
    if (s >= q.nstates) break;
    const DState &st = q.st[s];
    uint64_t *w = e + st.word;
    if (st.kind == SK_SUM_128) {
      const uint64_t f = (a.fl >> (8 * s)) & 0xFF;
      if (atomic) {
        g_add128(w, a.lo[s], a.hi[s]);
        if (st.fixed && f) g_atomic_or(w + 2, f);
      } else {
        w[0] = a.lo[s];
        w[1] = a.hi[s];
        if (st.fixed) w[2] = f;
      }
    } else if (st.kind == SK_MAX_CODE) {
      if (atomic) { if (a.lo[s]) g_atomic_max(w, a.lo[s]); }
      else w[0] = a.lo[s];
    } else {
      if (atomic) { if (a.lo[s]) g_atomic_add(w, a.lo[s]); }
      else w[0] = a.lo[s];
    }
  }
}

// per-wave carry of one query in LDS: [key code, flags, selected, lo[], hi[], fl]
//   flags bit 0: a run is open (valid), bit 1: that run is the range's first
static constexpr int CARRY_WORDS = 4 + 2 * B_MAX_STATES;

__device__ __forceinline__ void carry_load(const uint64_t *cw, uint64_t &key, uint32_t &flags, uint64_t &sel, Acc &a) {
  key = cw[0]; flags = (uint32_t)cw[1]; sel = cw[2];
#pragma unroll
  for (int s = 0; s < B_MAX_STATES; ++s) { a.lo[s] = cw[3 + s]; a.hi[s] = cw[3 + B_MAX_STATES + s]; }
  a.fl = cw[3 + 2 * B_MAX_STATES];
}

__device__ __forceinline__ void carry_store(uint64_t *cw, uint64_t key, uint32_t flags, uint64_t sel, const Acc &a) {
  cw[0] = key; cw[1] = flags; cw[2] = sel;
#pragma unroll
  for (int s = 0; s < B_MAX_STATES; ++s) { cw[3 + s] = a.lo[s]; cw[3 + B_MAX_STATES + s] = a.hi[s]; }
  cw[3 + 2 * B_MAX_STATES] = a.fl;
}

// ---- shared-interval group: interval index of a code (upper bound) ----
__device__ __forceinline__ uint32_t interval_of(const uint64_t *bounds, int n, uint64_t x) {
  int lo = 0, hi = n;  // first bound > x
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (bounds[mid] <= x) lo = mid + 1; else hi = mid;
  }
  return (uint32_t)lo;
}

template <int NC>
__global__ void __launch_bounds__(B_BLOCK) batch_kernel(BScan bs, const BQuery *__restrict__ qs) {
  extern __shared__ uint64_t lds[];
  // LDS layout (word offsets from the host): per shared group a bounds copy
  // and a table; then the per-wave carries of the sorted-key queries
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  for (int g = 0; g < bs.nshared; ++g) {
    const BShared &G = bs.sh[g];
    for (int i = threadIdx.x; i < G.nbounds; i += blockDim.x) lds[G.bnd_off + i] = G.bounds[i];
    const uint32_t words = (uint32_t)(G.nbounds + 1) * G.dom * (uint32_t)G.entry_words;
    for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) lds[G.lds_off + i] = 0;
  }
  uint64_t *my_carry = lds + bs.carry_off + (uint64_t)wave * bs.nq * CARRY_WORDS;
  for (int i = lane; i < bs.nq * CARRY_WORDS; i += 64) my_carry[i] = 0;
  __syncthreads();

  int32_t dtype[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) dtype[j] = bs.dtype[j];

  const uint64_t ntiles = (bs.nrows + B_TILE - 1) / B_TILE;
  const uint64_t wid = (uint64_t)blockIdx.x * nwaves + wave;
  const uint64_t t0 = wid * bs.tiles_per_wave;
  const uint64_t t1 = (t0 + bs.tiles_per_wave < ntiles) ? t0 + bs.tiles_per_wave : ntiles;

  for (uint64_t t = t0; t < t1; ++t) {
    const uint64_t row0 = t * B_TILE + (uint64_t)lane * RPT;
    uint32_t valid = 0;
    if (row0 < bs.nrows) {
      uint64_t left = bs.nrows - row0;
      valid = left >= RPT ? 0xFFu : ((1u << left) - 1);
    }
    uint64_t code[NC][RPT];
    uint32_t nul[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) {
#pragma unroll
      for (int r = 0; r < RPT; ++r) code[j][r] = 0;
      nul[j] = 0;
    }
    if (valid) {
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        load8_codes(bs.data[j], dtype[j], row0, code[j]);
        nul[j] = bs.nulls[j] ? (uint32_t)bs.nulls[j][row0 >> 3] : 0u;
      }
    }
    // per-row null bits across columns (bit j = column j is NULL)
    uint32_t rown[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      rown[r] = 0;
#pragma unroll
      for (int j = 0; j < NC; ++j) rown[r] |= ((nul[j] >> r) & 1u) << j;
    }

    // ---- shared-interval groups ----
    for (int g = 0; g < bs.nshared; ++g) {
      const BShared &G = bs.sh[g];
      const uint64_t *gb = lds + G.bnd_off;
      lds_u64 *tab = (lds_u64 *)(lds + G.lds_off);
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        if (!((valid >> r) & 1)) continue;
        uint64_t x = 0, kc = 0;
#pragma unroll
        for (int j = 0; j < NC; ++j) {
          if (G.col == j) x = code[j][r];
          if (G.key_col == j) kc = code[j][r];
        }
        if ((rown[r] >> G.col) & 1) continue;  // NULL never matches a range
        const uint32_t e = interval_of(gb, G.nbounds, x);
        uint32_t d = 0;
        if (G.key_col != 0xFF) {
          if ((rown[r] >> G.key_col) & 1) continue;  // small-domain keys are non-nullable here
          d = (uint32_t)(code_int(kc) - G.kmin);
        }
        lds_u64 *ent = tab + ((uint64_t)e * G.dom + d) * G.entry_words;
        for (int s = 0; s < G.nstates; ++s) {
          const DState &st = G.st[s];
          const uint8_t col = G.st_col[s];
          Contrib k;
          if (col == 0xFF) {
            k.a = 1; k.b = k.c = 0; k.skip = false;
          } else {
            if ((rown[r] >> col) & 1) continue;
            uint64_t c = 0;
            bool isf = false;
#pragma unroll
            for (int j = 0; j < NC; ++j)
              if (col == j) { c = code[j][r]; isf = dt_is_float(dtype[j]); }
            k = contrib_of(st, isf, isf ? 0 : code_int(c), isf ? code_flt(c) : 0.0);
            if (k.skip) continue;
          }
          apply_lds(st, ent, k);
        }
      }
    }

    // ---- sorted-key queries ----
    if (bs.nq == 0) continue;
    // key runs of this tile (all sorted-key queries of a launch share one key column)
    const int kcol = qs[0].key_col;
    uint64_t kk[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      uint64_t v = 0;
#pragma unroll
      for (int j = 0; j < NC; ++j) if (kcol == j) v = code[j][r];
      kk[r] = ((valid >> r) & 1) ? v : ~0ull;  // rows past the end form one empty run
    }
    const uint64_t kfirst = kk[0], klast = kk[RPT - 1];
    bool single = true;
#pragma unroll
    for (int r = 1; r < RPT; ++r) if (kk[r] != kk[r - 1]) single = false;
    const uint64_t prev_last = __shfl_up(klast, 1, 64);
    const uint64_t next_first = __shfl_down(kfirst, 1, 64);
    const bool head = lane == 0 || !single || kfirst != prev_last;
    const uint64_t heads = __ballot(head);
    const uint64_t tile_first = __shfl(kfirst, 0, 64);
    const uint64_t tile_last = __shfl(klast, 63, 64);
    const bool single0 = __shfl((int)single, 0, 64) != 0;
    // lane L's tail chain starts at lane 0 (no head in lanes 1..L)
    const uint64_t upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1);
    const bool chain0 = single0 && ((heads & upto & ~1ull) == 0);
    const bool chain0_prev = __shfl_up((int)chain0, 1, 64) != 0;

    for (int qi = 0; qi < bs.nq; ++qi) {
      const BQuery &q = qs[qi];
      // nslots > B_MAX_SLOTS marks a query whose filter can never hold
      uint32_t m = q.nslots > B_MAX_SLOTS ? 0u : valid;
      const int nsl = q.nslots > B_MAX_SLOTS ? 0 : q.nslots;
      for (int sl = 0; sl < nsl; ++sl) {
        const uint8_t c = q.slot_col[sl];
        const uint64_t lo = q.lo[sl], span = q.span[sl];
        uint32_t bits = 0;
#pragma unroll
        for (int j = 0; j < NC; ++j) {
          if (c == j) {
#pragma unroll
            for (int r = 0; r < RPT; ++r) bits |= ((code[j][r] - lo) <= span ? 1u : 0u) << r;
            bits &= ~nul[j];
          }
        }
        m &= bits;
      }
      uint64_t *cw = my_carry + qi * CARRY_WORDS;
      uint64_t ck;
      uint32_t cflags;
      uint64_t csel;
      Acc ca;
      carry_load(cw, ck, cflags, csel, ca);
      const bool cvalid = cflags & 1;
      const bool cfirst = (cflags & 2) != 0 || !cvalid;  // the open run (or the next) is the range's first
      const uint64_t any = __ballot(m != 0);
      if (!any) {
        // no selected row: only the run structure moves on
        if (cvalid && ck == tile_first && ck == tile_last) continue;  // the open run spans the tile
        if (lane == 0) {
          if (cvalid && ck != tile_first) acc_flush(q, ck, ca, cfirst);  // open run ends at the tile start
          // a tile holds at least two runs unless tile_first == tile_last
          const bool still_first = cfirst && (!cvalid || ck == tile_first) && tile_first == tile_last;
          Acc z;
          acc_zero(z);
          if (cvalid && ck == tile_first && tile_first != tile_last) {
            // the open run ends inside this tile with no new rows: close it
            acc_flush(q, ck, ca, cfirst);
          }
          carry_store(cw, tile_last, 1u | (still_first ? 2u : 0u), csel, z);
        }
        continue;
      }
      // per-lane runs: head run H, tail run T, interior runs flushed here (plain)
      Acc H, T;
      acc_zero(H);
      acc_zero(T);
      bool split = false;
      uint32_t nsel = 0;
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        if (r > 0 && kk[r] != kk[r - 1]) {
          if (!split) { H = T; split = true; }
          else acc_flush(q, kk[r - 1], T, false);
          acc_zero(T);
        }
        if ((m >> r) & 1) {
          uint64_t rc[NC];
#pragma unroll
          for (int j = 0; j < NC; ++j) rc[j] = code[j][r];
          acc_row<NC>(T, q, rc, dtype, rown[r]);
          ++nsel;
        }
      }
      // carry-in (lane 0)
      bool carry_closed_first = false;
      if (lane == 0 && cvalid) {
        if (ck == kfirst) {
          if (split) acc_merge(H, ca, q); else acc_merge(T, ca, q);
        } else {
          acc_flush(q, ck, ca, cfirst);  // the open run ended at the tile start
          carry_closed_first = cfirst;
        }
      }
      // the run containing this tile's row 0 is the range's first run iff
      // the open run continued into it (or this is the range's first tile)
      const bool rf0 = cfirst && (!cvalid || ck == tile_first);
      (void)carry_closed_first;
      // segmented inclusive scan of the tail partials over lanes
      Acc S = T;
      bool f = head;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        Acc y = acc_shfl_up(S, off, q);
        const bool fy = __shfl_up((int)f, off, 64) != 0;
        if (lane >= off) {
          if (!f) acc_merge(S, y, q);
          f = f || fy;
        }
      }
      const Acc Sprev = acc_shfl_up(S, 1, q);
      // head closure: a lane with several runs closes its first run
      if (split) {
        Acc hh = H;
        bool rf;
        if (lane > 0 && kfirst == prev_last) {
          acc_merge(hh, Sprev, q);
          rf = rf0 && chain0_prev;
        } else {
          rf = rf0 && lane == 0;
        }
        acc_flush(q, kfirst, hh, rf);
      }
      // tail closure: the chain ends where the next lane starts another key
      if (lane < 63 && next_first != klast) acc_flush(q, klast, S, rf0 && chain0);
      // lane 63's chain stays open into the next tile
      const Acc S63 = acc_shfl(S, 63, q);
      // selected rows of the tile
      uint32_t tot = nsel;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) tot += __shfl_xor(tot, off, 64);
      const bool open_first = rf0 && __shfl((int)chain0, 63, 64) != 0;
      if (lane == 0) carry_store(cw, tile_last, 1u | (open_first ? 2u : 0u), csel + tot, S63);
    }
  }
  // range end: flush each query's open run (the next range may share it)
  if (bs.nq) {
    for (int qi = lane; qi < bs.nq; qi += 64) {
      const BQuery &q = qs[qi];
      uint64_t *cw = my_carry + qi * CARRY_WORDS;
      uint64_t ck;
      uint32_t cflags;
      uint64_t csel;
      Acc ca;
      carry_load(cw, ck, cflags, csel, ca);
      if (cflags & 1) acc_flush(q, ck, ca, true);
      if (csel) g_atomic_add(q.selected, csel);
    }
  }
  // shared groups: merge the workgroup tables into the global tables
  __syncthreads();
  for (int g = 0; g < bs.nshared; ++g) {
    const BShared &G = bs.sh[g];
    const uint32_t nent = (uint32_t)(G.nbounds + 1) * G.dom;
    for (uint32_t i = threadIdx.x; i < nent; i += blockDim.x) {
      const uint64_t *src = lds + G.lds_off + (uint64_t)i * G.entry_words;
      if (src[0] == 0) continue;  // state 0 counts rows
      uint64_t *dst = G.table + (uint64_t)i * G.entry_words;
      for (int s = 0; s < G.nstates; ++s) {
        const DState &st = G.st[s];
        const uint64_t *w = src + st.word;
        uint64_t *d = dst + st.word;
        if (st.kind == SK_SUM_128) {
          g_add128(d, w[0], w[1]);
          if (st.fixed && w[2]) g_atomic_or(d + 2, w[2]);
        } else if (st.kind == SK_MAX_CODE) {
          if (w[0]) g_atomic_max(d, w[0]);
        } else if (w[0]) {
          g_atomic_add(d, w[0]);
        }
      }
    }
  }
}

// ---- shared-interval results: merge intervals [e_lo, e_hi) of each key into
// the query's DIRECT table (its own state layout) ----
__global__ void batch_combine_kernel(const BCombine *__restrict__ cb, uint32_t n) {
  const uint32_t qi = blockIdx.y;
  if (qi >= n) return;
  const BCombine &c = cb[qi];
  for (uint32_t d = blockIdx.x * blockDim.x + threadIdx.x; d < c.dom; d += gridDim.x * blockDim.x) {
    uint64_t acc[24];
    for (int w = 0; w < c.dst_words; ++w) acc[w] = 0;
    for (uint32_t e = c.e_lo; e < c.e_hi; ++e) {
      const uint64_t *src = c.src + ((uint64_t)e * c.dom + d) * c.src_words;
      if (src[0] == 0) continue;
      for (int s = 0; s < c.nstates; ++s) {
        const uint64_t *w = src + c.src_word[s];
        uint64_t *a = acc + c.dst_word[s];
        const uint8_t k = c.kind[s];
        if (k == SK_SUM_128) {
          uint64_t lo = a[0] + w[0];
          a[1] = a[1] + w[1] + (lo < a[0] ? 1 : 0);
          a[0] = lo;
          if (c.fixed[s]) a[2] |= w[2];
        } else if (k == SK_MAX_CODE) {
          a[0] = a[0] > w[0] ? a[0] : w[0];
        } else {
          a[0] += w[0];
        }
      }
    }
    uint64_t *dst = c.dst + (uint64_t)d * c.dst_words;
    for (int w = 0; w < c.dst_words; ++w) dst[w] = acc[w];
    if (d == 0 && c.selected) {
      // rows selected by this query: COUNT(*) over its intervals and keys
      uint64_t tot = 0;
      for (uint32_t dd = 0; dd < c.dom; ++dd)
        for (uint32_t e = c.e_lo; e < c.e_hi; ++e) tot += c.src[((uint64_t)e * c.dom + dd) * c.src_words];
      *c.selected = tot;
    }
  }
}

}  // namespace qm

using namespace qm;

static const void *batch_fn(int nc) {
  switch (nc) {
    case 1: return (const void *)batch_kernel<1>;
    case 2: return (const void *)batch_kernel<2>;
    case 3: return (const void *)batch_kernel<3>;
    default: return (const void *)batch_kernel<4>;
  }
}

extern "C" uint32_t qmk_batch_carry_words() { return CARRY_WORDS; }

extern "C" int qmk_batch_waves_per_cu(int nc, size_t lds_bytes) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_fn(nc), B_BLOCK, lds_bytes) != hipSuccess) nb = 1;
  return (nb < 1 ? 1 : nb) * (B_BLOCK / 64);
}

extern "C" hipError_t qmk_batch_scan(const BScan *bs, const BQuery *d_q, uint32_t grid, size_t lds_bytes,
                                     hipStream_t st) {
  if (bs->ncols < 1 || bs->ncols > B_MAX_COLS) return hipErrorInvalidValue;
  const void *fn = batch_fn(bs->ncols);
  if (lds_bytes > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
  }
  BScan a = *bs;
  const BQuery *q = d_q;
  void *args[] = {&a, &q};
  return hipLaunchKernel(fn, dim3(grid), dim3(B_BLOCK), args, lds_bytes, st);
}

extern "C" hipError_t qmk_batch_combine(const BCombine *d_cb, uint32_t n, uint32_t max_dom, hipStream_t st) {
  if (n == 0) return hipSuccess;
  uint32_t bx = (max_dom + 255) / 256;
  if (bx > 1024) bx = 1024;
  hipLaunchKernelGGL(batch_combine_kernel, dim3(bx, n), dim3(256), 0, st, d_cb, n);
  return hipGetLastError();
}

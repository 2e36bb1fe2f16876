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
// W (wide): the launch has 128-bit sums (int64 SUM, exact f64 SUM); without
// them the high words and flags are compiled out.
template <bool W>
struct AccT {
  uint64_t lo[B_MAX_STATES];
  uint64_t hi[W ? B_MAX_STATES : 1];
  uint64_t fl;  // SK_SUM_128 fixed: NaN / +Inf / -Inf flags, 8 bits per state
};

template <bool W>
__device__ __forceinline__ void acc_zero(AccT<W> &a) {
#pragma unroll
  for (int s = 0; s < B_MAX_STATES; ++s) a.lo[s] = 0;
  if constexpr (W) {
#pragma unroll
    for (int s = 0; s < B_MAX_STATES; ++s) a.hi[s] = 0;
  }
  a.fl = 0;
}

// add one contribution to state s
template <bool W>
__device__ __forceinline__ void acc_add(AccT<W> &a, int s, uint8_t kind, const Contrib &k) {
  if (kind == SK_MAX_CODE) {
    a.lo[s] = a.lo[s] > k.a ? a.lo[s] : k.a;
  } else if (kind == SK_SUM_128) {
    if constexpr (W) {
      uint64_t lo = a.lo[s] + k.a;
      a.hi[s] = a.hi[s] + k.b + (lo < a.lo[s] ? 1 : 0);
      a.lo[s] = lo;
      a.fl |= k.c << (8 * s);
    }
  } else {
    a.lo[s] += k.a;
  }
}

template <bool W, class Q>
__device__ __forceinline__ void acc_merge(AccT<W> &a, const AccT<W> &b, const Q &q) {
#pragma unroll
  for (int s = 0; s < B_MAX_STATES; ++s) {
    if (s >= q.nstates) break;
    const uint8_t k = q.st[s].kind;
    if (k == SK_MAX_CODE) {
      a.lo[s] = a.lo[s] > b.lo[s] ? a.lo[s] : b.lo[s];
    } else if (k == SK_SUM_128) {
      if constexpr (W) {
        uint64_t lo = a.lo[s] + b.lo[s];
        a.hi[s] = a.hi[s] + b.hi[s] + (lo < a.lo[s] ? 1 : 0);
        a.lo[s] = lo;
      }
    } else {
      a.lo[s] += b.lo[s];
    }
  }
  if constexpr (W) a.fl |= b.fl;
}

// add one row (column codes c[] of the row) to an accumulator
template <int NC, bool W, class Q>
__device__ __forceinline__ void acc_row(AccT<W> &a, const Q &q, const uint64_t (&c)[NC], const int32_t *dtype,
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
    acc_add<W>(a, s, st.kind, k);
  }
}

template <bool W, class Q>
__device__ __forceinline__ AccT<W> acc_shfl_up(const AccT<W> &a, int off, const Q &q) {
  AccT<W> r;
  acc_zero(r);
#pragma unroll
  for (int s = 0; s < B_MAX_STATES; ++s) {
    if (s < q.nstates) {
      r.lo[s] = __shfl_up(a.lo[s], off, 64);
      if constexpr (W) if (q.st[s].kind == SK_SUM_128) r.hi[s] = __shfl_up(a.hi[s], off, 64);
    }
  }
  if constexpr (W) r.fl = __shfl_up(a.fl, off, 64);
  return r;
}

template <bool W, class Q>
__device__ __forceinline__ AccT<W> acc_shfl(const AccT<W> &a, int src, const Q &q) {
  AccT<W> r;
  acc_zero(r);
#pragma unroll
  for (int s = 0; s < B_MAX_STATES; ++s) {
    if (s < q.nstates) {
      r.lo[s] = __shfl(a.lo[s], src, 64);
      if constexpr (W) if (q.st[s].kind == SK_SUM_128) r.hi[s] = __shfl(a.hi[s], src, 64);
    }
  }
  if constexpr (W) r.fl = __shfl(a.fl, src, 64);
  return r;
}

// write a closed run of key code `kc` into the query's DIRECT table
template <bool W, class Q>
__device__ __forceinline__ void acc_flush(const Q &q, uint64_t kc, const AccT<W> &a, bool atomic) {
  if (a.lo[0] == 0) return;  // state 0 is COUNT(*): no selected row in the run
  uint64_t *e = q.table + (uint64_t)(code_int(kc) - q.kmin) * (uint64_t)q.entry_words;
#pragma unroll
  for (int s = 0; s < B_MAX_STATES; ++s) {
    if (s >= q.nstates) break;
    const DState &st = q.st[s];
    uint64_t *w = e + st.word;
    if (st.kind == SK_SUM_128) {
      if constexpr (W) {
        const uint64_t f = (a.fl >> (8 * s)) & 0xFF;
        if (atomic) {
          g_add128(w, a.lo[s], a.hi[s]);
          if (st.fixed && f) g_atomic_or(w + 2, f);
        } else {
          gptr(w)[0] = a.lo[s];
          gptr(w)[1] = a.hi[s];
          if (st.fixed) gptr(w)[2] = f;
        }
      }
    } else if (st.kind == SK_MAX_CODE) {
      if (atomic) { if (a.lo[s]) g_atomic_max(w, a.lo[s]); }
      else *gptr(w) = a.lo[s];
    } else {
      if (atomic) { if (a.lo[s]) g_atomic_add(w, a.lo[s]); }
      else *gptr(w) = a.lo[s];
    }
  }
}


// The fields of a sorted-key query the tile loop reads, pinned to scalar
// registers (the descriptors live in LDS: re-reading them from global memory
// for every tile thrashes the scalar cache once a batch holds many queries).
struct QHot {
  int nslots;  // > B_MAX_SLOTS: the filter never holds
  uint8_t slot_col[B_MAX_SLOTS];
  uint64_t lo[B_MAX_SLOTS], span[B_MAX_SLOTS];
  int nstates;
  DState st[B_MAX_STATES];
  uint8_t st_col[B_MAX_STATES];
  int64_t kmin;
  int32_t entry_words;
  uint64_t *table;
};

__device__ __forceinline__ uint32_t rfl(uint32_t x) { return (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)x); }
__device__ __forceinline__ uint64_t rfl64(uint64_t x) {
  return ((uint64_t)rfl((uint32_t)(x >> 32)) << 32) | rfl((uint32_t)x);
}

__device__ __forceinline__ QHot qhot_of(const BQuery &d) {
  QHot q;
  q.nslots = (int)rfl(d.nslots);
#pragma unroll
  for (int i = 0; i < B_MAX_SLOTS; ++i) {
    q.slot_col[i] = (uint8_t)rfl(d.slot_col[i]);
    q.lo[i] = rfl64(d.lo[i]);
    q.span[i] = rfl64(d.span[i]);
  }
  q.nstates = (int)rfl(d.nstates);
#pragma unroll
  for (int i = 0; i < B_MAX_STATES; ++i) {
    q.st[i] = d.st[i];
    q.st[i].kind = (uint8_t)rfl(d.st[i].kind);
    q.st[i].is_min = (uint8_t)rfl(d.st[i].is_min);
    q.st[i].fixed = (uint8_t)rfl(d.st[i].fixed);
    q.st[i].word = (int16_t)rfl((uint32_t)(int32_t)d.st[i].word);
    q.st[i].scale = (int32_t)rfl((uint32_t)d.st[i].scale);
    q.st_col[i] = (uint8_t)rfl(d.st_col[i]);
  }
  q.kmin = (int64_t)rfl64((uint64_t)d.kmin);
  q.entry_words = (int32_t)rfl((uint32_t)d.entry_words);
  q.table = (uint64_t *)rfl64((uint64_t)(uintptr_t)d.table);
  return q;
}

// contributions of one state over the rows of masks mH / mT (fast path: the
// state's column is chosen once, not per row)
template <int NC, bool W>
__device__ __forceinline__ void state_masked(const QHot &q, int s, const uint64_t (&code)[NC][RPT],
                                             const int32_t *dtype, const uint32_t *nul, uint32_t mH, uint32_t mT,
                                             AccT<W> &H, AccT<W> &T) {
  const DState &st = q.st[s];
  const uint8_t col = q.st_col[s];
  if (col == 0xFF) {  // COUNT(*)
    H.lo[s] += __builtin_popcount(mH);
    T.lo[s] += __builtin_popcount(mT);
    return;
  }
#pragma unroll
  for (int j = 0; j < NC; ++j) {
    if (col != j) continue;
    const bool isf = dt_is_float(dtype[j]);
    const uint32_t mh = mH & ~nul[j], mt = mT & ~nul[j];
    if (st.kind == SK_COUNT) {
      H.lo[s] += __builtin_popcount(mh);
      T.lo[s] += __builtin_popcount(mt);
      continue;
    }
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      const bool inh = (mh >> r) & 1, int_ = (mt >> r) & 1;
      if (!(inh || int_)) continue;
      const uint64_t c = code[j][r];
      Contrib k = contrib_of(st, isf, isf ? 0 : code_int(c), isf ? code_flt(c) : 0.0);
      if (k.skip) continue;
      if (inh) acc_add<W>(H, s, st.kind, k);
      else acc_add<W>(T, s, st.kind, k);
    }
  }
}

// per-wave carry of one query in LDS: [key code, flags, selected, lo[], hi[], fl]
//   flags bit 0: a run is open (valid), bit 1: that run is the range's first
static constexpr int CARRY_WORDS = 4 + 2 * B_MAX_STATES;

template <bool W>
__device__ __forceinline__ void carry_load(const uint64_t *cw, uint64_t &key, uint32_t &flags, uint64_t &sel,
                                           AccT<W> &a) {
  key = cw[0]; flags = (uint32_t)cw[1]; sel = cw[2];
#pragma unroll
  for (int s = 0; s < B_MAX_STATES; ++s) a.lo[s] = cw[3 + s];
  if constexpr (W) {
#pragma unroll
    for (int s = 0; s < B_MAX_STATES; ++s) a.hi[s] = cw[3 + B_MAX_STATES + s];
    a.fl = cw[3 + 2 * B_MAX_STATES];
  } else {
    a.fl = 0;
  }
}

template <bool W>
__device__ __forceinline__ void carry_store(uint64_t *cw, uint64_t key, uint32_t flags, uint64_t sel,
                                            const AccT<W> &a) {
  cw[0] = key; cw[1] = flags; cw[2] = sel;
#pragma unroll
  for (int s = 0; s < B_MAX_STATES; ++s) cw[3 + s] = a.lo[s];
  if constexpr (W) {
#pragma unroll
    for (int s = 0; s < B_MAX_STATES; ++s) cw[3 + B_MAX_STATES + s] = a.hi[s];
    cw[3 + 2 * B_MAX_STATES] = a.fl;
  }
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

// MODE 1: shared-interval groups only; MODE 2: sorted-key queries only (the
// two paths run as separate launches so each keeps its registers).
template <int NC, bool W, int MODE>
__global__ void __launch_bounds__(B_BLOCK) batch_kernel(BScan bs, const BQuery *__restrict__ qs) {
  using Acc = AccT<W>;
  extern __shared__ uint64_t lds[];
  // LDS layout (word offsets from the host): per shared group a bounds copy
  // and a table; then the per-wave carries of the sorted-key queries
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
#pragma unroll
  for (int g = 0; g < B_MAX_SHARED; ++g) {
    if (g >= bs.nshared) break;
    const BShared &G = bs.sh[g];
    for (int i = threadIdx.x; i < G.nbounds; i += blockDim.x) lds[G.bnd_off + i] = G.bounds[i];
    const uint32_t words = (uint32_t)(G.nbounds + 1) * G.dom * (uint32_t)G.entry_words;
    for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) lds[G.lds_off + i] = 0;
  }
  __shared__ DState sst[B_MAX_SHARED * B_MAX_SHARED_STATES];
  __shared__ uint8_t scol[B_MAX_SHARED * B_MAX_SHARED_STATES];
  if (threadIdx.x < B_MAX_SHARED * B_MAX_SHARED_STATES) {
    const int g = threadIdx.x / B_MAX_SHARED_STATES, k = threadIdx.x % B_MAX_SHARED_STATES;
    sst[threadIdx.x] = bs.sh[g].st[k];
    scol[threadIdx.x] = bs.sh[g].st_col[k];
  }
  uint64_t *my_carry = lds + bs.carry_off + (uint64_t)wave * bs.nq * CARRY_WORDS;
  for (int i = lane; i < bs.nq * CARRY_WORDS; i += 64) my_carry[i] = 0;
  BQuery *qlds = reinterpret_cast<BQuery *>(lds + bs.qdesc_off);
  {
    const uint64_t *src = reinterpret_cast<const uint64_t *>(qs);
    uint64_t *dst = lds + bs.qdesc_off;
    const uint32_t words = (uint32_t)(bs.nq * sizeof(BQuery) / 8);
    for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
  }
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
    if constexpr (MODE == 1) {
#pragma unroll
    for (int g = 0; g < B_MAX_SHARED; ++g) {
      if (g >= bs.nshared) break;
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
        const DState *gst = sst + g * B_MAX_SHARED_STATES;
        const uint8_t *gcol = scol + g * B_MAX_SHARED_STATES;
#pragma unroll 1
        for (int s = 0; s < G.nstates; ++s) {
          const DState st = gst[s];
          const uint8_t col = gcol[s];
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

    }  // MODE 1

    // ---- sorted-key queries ----
    if constexpr (MODE == 2) {
    if (bs.nq == 0) continue;
    // key runs of this tile: the launch's key column is batch column 0; rows
    // past the end form one empty run
#pragma unroll
    for (int r = 0; r < RPT; ++r)
      if (!((valid >> r) & 1)) code[0][r] = ~0ull;
    const uint64_t (&kk)[RPT] = code[0];
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

    // run structure of the lane's 8 rows (query independent): head run rows
    // [0, fc), tail run rows [lc, 8), interior runs in between
    int fc = RPT, lc = 0;
#pragma unroll
    for (int r = 1; r < RPT; ++r)
      if (kk[r] != kk[r - 1]) { if (fc == RPT) fc = r; lc = r; }
    const bool lane_split = fc < RPT;
    const uint32_t hmask = lane_split ? ((1u << fc) - 1) : 0u;
    const uint32_t tmask = 0xFFu & ~((1u << lc) - 1);
    const uint32_t imask = 0xFFu & ~hmask & ~tmask;

    for (int qi = 0; qi < bs.nq; ++qi) {
      const QHot q = qhot_of(qlds[qi]);
      // nslots > B_MAX_SLOTS marks a query whose filter can never hold
      uint32_t m = q.nslots > B_MAX_SLOTS ? 0u : valid;
      const int nsl = q.nslots > B_MAX_SLOTS ? 0 : q.nslots;
#pragma unroll
      for (int sl = 0; sl < B_MAX_SLOTS; ++sl) {
        if (sl >= nsl) break;
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
      // per-lane runs: head run H, tail run T (interior runs are flushed here)
      Acc H, T;
      acc_zero(H);
      acc_zero(T);
      const bool split = lane_split;
      const uint32_t nsel = __builtin_popcount(m);
      if (__ballot((m & imask) != 0) == 0) {
        // no selected row inside an interior run: masked sums per state
        const uint32_t mH = m & hmask, mT = m & tmask;
#pragma unroll
        for (int st = 0; st < B_MAX_STATES; ++st) {
          if (st >= q.nstates) break;
          state_masked<NC, W>(q, st, code, dtype, nul, mH, mT, H, T);
        }
      } else {
        bool sp = false;
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
          if (r > 0 && kk[r] != kk[r - 1]) {
            if (!sp) { H = T; sp = true; }
            else acc_flush(q, kk[r - 1], T, false);
            acc_zero(T);
          }
          if ((m >> r) & 1) {
            uint64_t rc[NC];
#pragma unroll
            for (int j = 0; j < NC; ++j) rc[j] = code[j][r];
            acc_row<NC, W>(T, q, rc, dtype, rown[r]);
          }
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
    }  // MODE 2
  }
  // range end: flush each query's open run (the next range may share it)
  if (bs.nq) {
    for (int qi = lane; qi < bs.nq; qi += 64) {
      const BQuery &q = qlds[qi];
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
#pragma unroll
  for (int g = 0; g < B_MAX_SHARED; ++g) {
    if (g >= bs.nshared) break;
    const BShared &G = bs.sh[g];
    const uint32_t nent = (uint32_t)(G.nbounds + 1) * G.dom;
    for (uint32_t i = threadIdx.x; i < nent; i += blockDim.x) {
      const uint64_t *src = lds + G.lds_off + (uint64_t)i * G.entry_words;
      if (src[0] == 0) continue;  // state 0 counts rows
      uint64_t *dst = G.table + (uint64_t)i * G.entry_words;
#pragma unroll
      for (int s = 0; s < B_MAX_SHARED_STATES; ++s) {
        if (s >= G.nstates) break;
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


// ---- broadcast helpers of the query-parallel path ----
__device__ __forceinline__ uint64_t rl64(uint64_t v, int r) {
  uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)v, r);
  uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)(v >> 32), r);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t load1_code(const void *data, int32_t dtype, uint64_t row) {
  DCol c;
  c.data = data; c.nulls = nullptr; c.dtype = dtype; c.side = 0;
  if (dt_is_float(dtype)) return code_f64_bits(f64_bits(load1_f(c, row)));
  return code_i64(load1_i(c, row));
}

// UNI: every query's filter is (at most) one range on column bs.qp_pred_col,
// so the compared value is one uniform select per row.
// ---- lean query-parallel kernel: single-range queries on one column with an
// aggregate list of NADD additive states (COUNT(*), COUNT(col), integer SUM;
// state 0 is COUNT(*)) followed by NMAX order-code states (MIN / MAX). Each
// 64-row chunk computes every state's per-row contribution once, with rows
// across lanes; the row loop then broadcasts them and each lane only selects
// and accumulates. bs.sig_st / sig_col hold the states in that order. ----
template <int NADD, int NMAX, int QPL>
__global__ void __launch_bounds__(B_BLOCK) batch_qpf_kernel(BScan bs, const BQuery *__restrict__ qs) {
  constexpr int NS = NADD + NMAX;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const uint64_t wid = (uint64_t)blockIdx.x * (blockDim.x >> 6) + wave;
  const uint64_t span_rows = bs.tiles_per_wave * (uint64_t)B_TILE;
  const uint64_t r0 = wid * span_rows;
  const uint64_t r1 = (r0 + span_rows < bs.nrows) ? r0 + span_rows : bs.nrows;
  if (r0 >= r1) return;
  const int pc = bs.qp_pred_col;
  bool qon[QPL];
  uint64_t slo[QPL], sspan[QPL];
  bool needs_value[QPL];  // the query has a range (else: IS NOT NULL-free, every row)
  uint64_t *tab[QPL];
  int64_t kmin = 0;
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    const int qi = k * 64 + lane;
    qon[k] = qi < bs.nq;
    const BQuery &q = qs[qon[k] ? qi : 0];
    if (q.nslots > 1) qon[k] = false;
    needs_value[k] = q.nslots == 1;
    slo[k] = q.nslots == 1 ? q.lo[0] : 0;
    sspan[k] = q.nslots == 1 ? q.span[0] : ~0ull;
    tab[k] = q.table;
    kmin = q.kmin;
  }
  kmin = (int64_t)rfl64((uint64_t)kmin);
  const int W_ = bs.sig_entry_words;
  uint64_t acc[QPL][NS];
  uint32_t nsel[QPL];
#pragma unroll
  for (int k = 0; k < QPL; ++k) {
    nsel[k] = 0;
#pragma unroll
    for (int s = 0; s < NS; ++s) acc[k][s] = 0;
  }
  auto flush = [&](uint64_t kc, bool atomic) {
#pragma unroll
    for (int k = 0; k < QPL; ++k) {
      if (!qon[k] || acc[k][0] == 0) continue;
      uint64_t *e = tab[k] + (uint64_t)(code_int(kc) - kmin) * (uint64_t)W_;
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        uint64_t *w = e + bs.sig_st[s].word;
        if (atomic) {
          if (s < NADD) g_atomic_add(w, acc[k][s]);
          else if (acc[k][s]) g_atomic_max(w, acc[k][s]);
        } else {
          *gptr(w) = acc[k][s];
        }
        acc[k][s] = 0;
      }
    }
  };
  const DCol kcol = DCol{bs.data[0], nullptr, bs.dtype[0], 0};
  uint64_t cur = 0;
  bool have = false, first = true;
  for (uint64_t c0 = r0; c0 < r1; c0 += 64) {
    const uint64_t row = c0 + lane;
    const bool ok = row < r1;
    // per-row (lane) values: key, predicate code (+ NULL), state contributions
    uint64_t kv = 0, pv = 0, cv[NS];
    bool pn = false;
    if (ok) {
      kv = code_i64(load1_i(kcol, row));
      pv = load1_code(bs.data[pc], bs.dtype[pc], row);
      pn = bs.nulls[pc] && ((bs.nulls[pc][row >> 3] >> (row & 7)) & 1);
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      cv[s] = 0;
      const uint8_t col = bs.sig_col[s];
      if (col == 0xFF) { cv[s] = ok ? 1 : 0; continue; }
      if (!ok) continue;
      if (bs.nulls[col] && ((bs.nulls[col][row >> 3] >> (row & 7)) & 1)) continue;
      const bool isf = dt_is_float(bs.dtype[col]);
      const uint64_t c = load1_code(bs.data[col], bs.dtype[col], row);
      Contrib kx = contrib_of(bs.sig_st[s], isf, isf ? 0 : code_int(c), isf ? code_flt(c) : 0.0);
      if (!kx.skip) cv[s] = kx.a;
    }
    const uint32_t pnm = (uint32_t)__ballot(pn);  // lanes 0..31 only matter below via readlane
    (void)pnm;
    const int nr = (int)((r1 - c0 < 64) ? (r1 - c0) : 64);
    for (int r = 0; r < nr; ++r) {
      const uint64_t key = rl64(kv, r);
      if (!have || key != cur) {
        if (have) {
          flush(cur, first);
          first = false;
        }
        cur = key;
        have = true;
      }
      const uint64_t x = rl64(pv, r);
      const bool xnull = __builtin_amdgcn_readlane((int)pn, r) != 0;
      uint64_t c[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) c[s] = rl64(cv[s], r);
#pragma unroll
      for (int k = 0; k < QPL; ++k) {
        const bool sel = qon[k] && (x - slo[k]) <= sspan[k] && !(xnull && needs_value[k]);
        nsel[k] += sel ? 1u : 0u;
#pragma unroll
        for (int s = 0; s < NADD; ++s) acc[k][s] += sel ? c[s] : 0ull;
#pragma unroll
        for (int s = NADD; s < NS; ++s) {
          const uint64_t v = sel ? c[s] : 0ull;
          acc[k][s] = acc[k][s] > v ? acc[k][s] : v;
        }
      }
    }
  }
  if (have) flush(cur, true);
#pragma unroll
  for (int k = 0; k < QPL; ++k)
    if (qon[k] && nsel[k]) g_atomic_add(qs[k * 64 + lane].selected, (uint64_t)nsel[k]);
}

}  // namespace qm

using namespace qm;

template <int MODE>
static const void *batch_fn_m(int nc, int wide) {
  if (wide) {
    switch (nc) {
      case 1: return (const void *)batch_kernel<1, true, MODE>;
      case 2: return (const void *)batch_kernel<2, true, MODE>;
      case 3: return (const void *)batch_kernel<3, true, MODE>;
      default: return (const void *)batch_kernel<4, true, MODE>;
    }
  }
  switch (nc) {
    case 1: return (const void *)batch_kernel<1, false, MODE>;
    case 2: return (const void *)batch_kernel<2, false, MODE>;
    case 3: return (const void *)batch_kernel<3, false, MODE>;
    default: return (const void *)batch_kernel<4, false, MODE>;
  }
}

// the shared path never needs the wide accumulators (it keeps no registers per run)
static const void *batch_fn_shared(int nc) {
  switch (nc) {
    case 1: return (const void *)batch_kernel<1, false, 1>;
    case 2: return (const void *)batch_kernel<2, false, 1>;
    case 3: return (const void *)batch_kernel<3, false, 1>;
    default: return (const void *)batch_kernel<4, false, 1>;
  }
}

static const void *batch_fn(int nc, int wide, int mode) {
  return mode == 1 ? batch_fn_shared(nc) : batch_fn_m<2>(nc, wide);
}

extern "C" uint32_t qmk_batch_carry_words() { return CARRY_WORDS; }

template <int QPL>
static const void *qpf_fn_q(int nadd, int nmax) {
  switch (nadd * 8 + nmax) {
    case 1 * 8 + 0: return (const void *)batch_qpf_kernel<1, 0, QPL>;
    case 1 * 8 + 1: return (const void *)batch_qpf_kernel<1, 1, QPL>;
    case 1 * 8 + 2: return (const void *)batch_qpf_kernel<1, 2, QPL>;
    case 1 * 8 + 3: return (const void *)batch_qpf_kernel<1, 3, QPL>;
    case 2 * 8 + 0: return (const void *)batch_qpf_kernel<2, 0, QPL>;
    case 2 * 8 + 1: return (const void *)batch_qpf_kernel<2, 1, QPL>;
    case 2 * 8 + 2: return (const void *)batch_qpf_kernel<2, 2, QPL>;
    case 3 * 8 + 0: return (const void *)batch_qpf_kernel<3, 0, QPL>;
    case 3 * 8 + 1: return (const void *)batch_qpf_kernel<3, 1, QPL>;
    case 4 * 8 + 0: return (const void *)batch_qpf_kernel<4, 0, QPL>;
  }
  return nullptr;
}

static const void *qpf_fn(int nadd, int nmax, int qpl) {
  switch (qpl) {
    case 1: return qpf_fn_q<1>(nadd, nmax);
    case 2: return qpf_fn_q<2>(nadd, nmax);
    default: return qpf_fn_q<4>(nadd, nmax);
  }
}

extern "C" int qmk_batch_qpf_supported(int nadd, int nmax) {
  return nadd >= 1 && nmax >= 0 && nadd + nmax <= B_MAX_STATES && qpf_fn(nadd, nmax, 1) != nullptr;
}

extern "C" int qmk_batch_qpf_waves_per_cu(int nadd, int nmax, int qpl) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, qpf_fn(nadd, nmax, qpl), B_BLOCK, 0) != hipSuccess) nb = 1;
  return (nb < 1 ? 1 : nb) * (B_BLOCK / 64);
}

// lean query-parallel launch: states ordered additive first, then order codes
extern "C" hipError_t qmk_batch_qpf_scan(const BScan *bs, const BQuery *d_q, uint32_t grid, int nadd, int nmax,
                                         int qpl, hipStream_t st) {
  const void *fn = qpf_fn(nadd, nmax, qpl);
  if (!fn || bs->nq > 64 * qpl || bs->nshared) return hipErrorInvalidValue;
  BScan a = *bs;
  const BQuery *q = d_q;
  void *args[] = {&a, &q};
  return hipLaunchKernel(fn, dim3(grid), dim3(B_BLOCK), args, 0, st);
}

extern "C" int qmk_batch_waves_per_cu(int nc, int wide, int mode, size_t lds_bytes) {
  int nb = 0;
  const void *fn = batch_fn(nc, wide, mode);
  if (lds_bytes > 64 * 1024)
    (void)qmk_func_lds(fn, lds_bytes);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, B_BLOCK, lds_bytes) != hipSuccess) nb = 1;
  return (nb < 1 ? 1 : nb) * (B_BLOCK / 64);
}

extern "C" hipError_t qmk_batch_scan(const BScan *bs, const BQuery *d_q, uint32_t grid, size_t lds_bytes,
                                     int wide, hipStream_t st) {
  if (bs->ncols < 1 || bs->ncols > B_MAX_COLS) return hipErrorInvalidValue;
  if ((bs->nshared > 0) == (bs->nq > 0)) return hipErrorInvalidValue;  // one kind of query per launch
  const void *fn = batch_fn(bs->ncols, wide, bs->nshared > 0 ? 1 : 2);
  if (lds_bytes > 64 * 1024) {
    hipError_t e = qmk_func_lds(fn, lds_bytes);
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

// qm_common.h — value encodings shared by the host planner and the gfx950
// kernels: exact fixed-point FP64 summation, order-preserving integer codes
// for MIN/MAX, the hash-table key code and the 64-bit mixer.
//
// Everything here is integer arithmetic on IEEE-754 bit patterns, so the
// device and host compilations give identical results.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define QM_HD __host__ __device__ __forceinline__
#else
#define QM_HD static inline
#endif

namespace qm {

QM_HD uint64_t f64_bits(double x) {
  union { double d; uint64_t u; } v; v.d = x; return v.u;
}
QM_HD double bits_f64(uint64_t u) {
  union { double d; uint64_t u; } v; v.u = u; return v.d;
}
QM_HD uint32_t f32_bits(float x) {
  union { float f; uint32_t u; } v; v.f = x; return v.u;
}

QM_HD bool f64_is_nan_bits(uint64_t b) {
  return ((b >> 52) & 0x7FF) == 0x7FF && (b & 0xFFFFFFFFFFFFFull) != 0;
}
QM_HD bool f64_is_inf_bits(uint64_t b) {
  return ((b >> 52) & 0x7FF) == 0x7FF && (b & 0xFFFFFFFFFFFFFull) == 0;
}

// ---------------------------------------------------------------------------
// Order-preserving unsigned codes. MIN and MAX are both kept as an unsigned
// MAX so that an all-zero table is the identity state:
//   max slot = max(code(v)),  min slot = max(~code(v)).
// A slot that is still 0 has seen no (non-NaN) value.
// ---------------------------------------------------------------------------
QM_HD uint64_t code_i64(int64_t v) { return (uint64_t)v ^ 0x8000000000000000ull; }
QM_HD int64_t decode_i64(uint64_t c) { return (int64_t)(c ^ 0x8000000000000000ull); }

// f64: sign-magnitude -> two's-complement-like unsigned order. -0 is folded
// onto +0 (they compare equal in the reference's `<` / `>` updates).
QM_HD uint64_t code_f64_bits(uint64_t b) {
  if (b == 0x8000000000000000ull) b = 0;
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
QM_HD uint64_t decode_f64_code(uint64_t c) {
  return (c >> 63) ? (c & 0x7FFFFFFFFFFFFFFFull) : ~c;
}

// ---------------------------------------------------------------------------
// Hash-table key code. Stored key 0 marks an empty slot; a raw 64-bit key k
// is stored as k + KEY_SALT (mod 2^64). The single raw key that would store
// as 0 (KEY_UNREPRESENTABLE) lives in a reserved side entry.
// ---------------------------------------------------------------------------
static constexpr uint64_t KEY_SALT = 0x9E3779B97F4A7C15ull;
static constexpr uint64_t KEY_UNREPRESENTABLE = 0ull - KEY_SALT;

// MurmurHash3 64-bit finalizer (the mixer the reference's vectorized engine
// uses for its LinearHashTable, wasm/src/query/vector_engine.zig:241-249).
QM_HD uint64_t mix64(uint64_t k) {
  k ^= k >> 33; k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33; return k;
}

// Group-key canonicalisation for f64 keys: the reference groups by
// String(v) (src/types.ts:10-18), so -0 and +0 share a group and every NaN
// shares one group.
QM_HD uint64_t canon_f64_key_bits(uint64_t b) {
  if (b == 0x8000000000000000ull) return 0;
  if (f64_is_nan_bits(b)) return 0x7FF8000000000000ull;
  return b;
}

// ---------------------------------------------------------------------------
// Exact FP64 summation by 128-bit fixed point.
// A finite double x with |x| < 2^E is mapped to round(x * 2^(97-E)) as a
// two's-complement 128-bit integer (|q| < 2^97). Sums of up to 2^30 such
// terms cannot overflow 128 bits and are independent of the order of
// addition; values with |x| >= 2^(E-44) convert exactly, smaller ones carry
// an error below 2^(E-98) each. `scale` = 97 - E.
// ---------------------------------------------------------------------------
struct u128 { uint64_t lo, hi; };

QM_HD u128 u128_neg(u128 a) {
  u128 r; r.lo = ~a.lo + 1; r.hi = ~a.hi + (r.lo == 0 ? 1 : 0); return r;
}
QM_HD u128 u128_add(u128 a, u128 b) {
  u128 r; r.lo = a.lo + b.lo; r.hi = a.hi + b.hi + (r.lo < a.lo ? 1 : 0); return r;
}

// x must be finite. Rounds half away from zero.
QM_HD u128 f64_to_fixed(double x, int scale) {
  uint64_t b = f64_bits(x);
  int neg = (int)(b >> 63);
  int ex = (int)((b >> 52) & 0x7FF);
  uint64_t m = b & 0xFFFFFFFFFFFFFull;
  u128 q; q.lo = 0; q.hi = 0;
  if (ex == 0) { if (m == 0) return q; ex = 1; } else { m |= 0x10000000000000ull; }
  int sh = ex - 1075 + scale;  // x = m * 2^(ex-1075)
  if (sh >= 0) {
    if (sh >= 128) return q;   // cannot happen when |x| < 2^E
    if (sh == 0) { q.lo = m; }
    else if (sh < 64) { q.lo = m << sh; q.hi = m >> (64 - sh); }
    else { q.hi = m << (sh - 64); }
  } else {
    int r = -sh;
    if (r > 54) return q;      // |x * 2^scale| < 1/2 -> 0
    q.lo = (m + (1ull << (r - 1))) >> r;
  }
  return neg ? u128_neg(q) : q;
}

QM_HD int clz64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return x ? __clzll((long long)x) : 64;
#else
  return x ? __builtin_clzll(x) : 64;
#endif
}

// Multiply by 2^e exactly (result assumed normal or zero/inf).
QM_HD double pow2_scale(double v, int e) {
  // split to stay inside the exponent range of each factor
  while (e > 1000) { v *= bits_f64((uint64_t)(1023 + 1000) << 52); e -= 1000; }
  while (e < -1000) { v *= bits_f64((uint64_t)(1023 - 1000) << 52); e += 1000; }
  return v * bits_f64((uint64_t)(1023 + e) << 52);
}

// Correctly rounded (to nearest, ties to even) conversion of the 128-bit
// fixed-point sum back to a double: q * 2^-scale.
QM_HD double fixed_to_f64(u128 q, int scale) {
  int neg = (int)(q.hi >> 63);
  if (neg) q = u128_neg(q);
  if (q.hi == 0 && q.lo == 0) return 0.0;
  int nbits = q.hi ? 128 - clz64(q.hi) : 64 - clz64(q.lo);
  uint64_t mant; int shift = 0;
  if (nbits <= 53) {
    mant = q.lo;
  } else {
    shift = nbits - 53;
    // mant = q >> shift; rem = q & (2^shift - 1)
    // shift is in [1, 75] because nbits <= 128
    uint64_t rem_hi, rem_lo, half_hi, half_lo;
    if (shift >= 64) {
      mant = q.hi >> (shift - 64);
      rem_hi = (shift == 64) ? 0 : (q.hi & ((1ull << (shift - 64)) - 1));
      rem_lo = q.lo;
    } else {
      mant = (q.lo >> shift) | (q.hi << (64 - shift));
      rem_hi = 0;
      rem_lo = q.lo & ((1ull << shift) - 1);
    }
    if (shift - 1 >= 64) { half_hi = 1ull << (shift - 65); half_lo = 0; }
    else { half_hi = 0; half_lo = 1ull << (shift - 1); }
    bool gt = (rem_hi > half_hi) || (rem_hi == half_hi && rem_lo > half_lo);
    bool eq = (rem_hi == half_hi && rem_lo == half_lo);
    if (gt || (eq && (mant & 1))) {
      mant += 1;
      if (mant == (1ull << 53)) { mant >>= 1; shift += 1; }
    }
  }
  double r = (double)mant;  // exact: mant <= 2^53
  r = pow2_scale(r, shift - scale);
  return neg ? -r : r;
}

// Sum-state flags for values the fixed point cannot hold.
static constexpr uint64_t SUMF_NAN = 1, SUMF_PINF = 2, SUMF_NINF = 4;

QM_HD double finish_f64_sum(u128 q, uint64_t flags, int scale) {
  if ((flags & SUMF_NAN) || ((flags & SUMF_PINF) && (flags & SUMF_NINF)))
    return bits_f64(0x7FF8000000000000ull);
  if (flags & SUMF_PINF) return bits_f64(0x7FF0000000000000ull);
  if (flags & SUMF_NINF) return bits_f64(0xFFF0000000000000ull);
  return fixed_to_f64(q, scale);
}

// The fixed-point scale for a column whose finite values satisfy
// |x| <= max_abs: smallest E with max_abs < 2^E, scale = 97 - E.
QM_HD int fixed_scale_for(double max_abs) {
  if (!(max_abs > 0)) return 97;  // all zeros (or empty): any scale works
  uint64_t b = f64_bits(max_abs);
  int ex = (int)((b >> 52) & 0x7FF);
  int E = (ex == 0) ? -1022 : ex - 1023 + 1;  // max_abs < 2^E
  return 97 - E;
}

// Reference AVG for int64 (bigint) columns, src/partial-agg.ts:38-44:
// Number(S / n) + Number(S % n) / n with truncating BigInt division.
// S is a 128-bit two's-complement sum, n > 0.
QM_HD double bigint_avg(u128 s, uint64_t n) {
  int neg = (int)(s.hi >> 63);
  u128 a = neg ? u128_neg(s) : s;
  // long division of the 128-bit magnitude by n (bit by bit; n < 2^63)
  u128 quo; quo.lo = 0; quo.hi = 0;
  uint64_t rem = 0;
  for (int i = 127; i >= 0; --i) {
    uint64_t bit = (i >= 64) ? ((a.hi >> (i - 64)) & 1) : ((a.lo >> i) & 1);
    // rem < n < 2^63, so rem*2+bit does not overflow
    rem = (rem << 1) | bit;
    if (rem >= n) {
      rem -= n;
      if (i >= 64) quo.hi |= (1ull << (i - 64)); else quo.lo |= (1ull << i);
    }
  }
  double qd = fixed_to_f64(quo, 0);
  double rd = (double)rem;  // exact: rem < n <= 2^53 in practice
  if (neg) { qd = -qd; rd = -rd; }
  return qd + rd / (double)n;
}

// ---------------------------------------------------------------------------
// Exact FP64 summation over any exponent range (SK_SUM_WIDE): signed 32-bit
// digits held in 64-bit words, value = sum_i w_i * 2^(base + 32 i). A finite
// x = +-m * 2^e (m < 2^53) adds +-(m << o) split into three digits starting
// at word (e - base) >> 5, o = (e - base) & 31 (base <= the exponent of the
// lowest bit of every value of the column). Words never carry while summing
// (|w_i| < 2^63 for up to 2^31 terms), so the sum is independent of the
// order of addition; carries are propagated once at the end, and the exact
// integer is rounded once (to nearest, ties to even).
// ---------------------------------------------------------------------------
// *idx = first word (-1: x == 0), d[3] = signed digits
QM_HD void f64_wide_digits(double x, int base, int *idx, int64_t d[3]) {
  const uint64_t b = f64_bits(x);
  const int neg = (int)(b >> 63);
  const int ex = (int)((b >> 52) & 0x7FF);
  uint64_t m = b & 0xFFFFFFFFFFFFFull;
  int e;
  if (ex == 0) e = -1074; else { m |= 0x10000000000000ull; e = ex - 1075; }
  if (m == 0) { *idx = -1; d[0] = d[1] = d[2] = 0; return; }
  const int p = e - base;
  const int o = p & 31;
  *idx = p >> 5;
  const uint64_t lo = m << o;
  const uint64_t hi = o ? (m >> (64 - o)) : 0ull;
  d[0] = (int64_t)(lo & 0xFFFFFFFFull);
  d[1] = (int64_t)(lo >> 32);
  d[2] = (int64_t)hi;
  if (neg) { d[0] = -d[0]; d[1] = -d[1]; d[2] = -d[2]; }
}

// Correctly rounded double of the digit words (flags: SUMF_* of the
// non-finite values).
QM_HD double wide_to_f64(const uint64_t *w, int nw, int base, uint64_t flags) {
  if ((flags & SUMF_NAN) || ((flags & SUMF_PINF) && (flags & SUMF_NINF))) return bits_f64(0x7FF8000000000000ull);
  if (flags & SUMF_PINF) return bits_f64(0x7FF0000000000000ull);
  if (flags & SUMF_NINF) return bits_f64(0xFFF0000000000000ull);
  // sign: the carry out of the top word
  int64_t c = 0;
  for (int i = 0; i < nw; ++i) c = ((int64_t)w[i] + c) >> 32;
  const bool neg = c < 0;
  // magnitude digits (of -sum when negative): highest nonzero digit
  int top = -1;
  c = 0;
  for (int i = 0; i < nw; ++i) {
    const int64_t t = (neg ? -(int64_t)w[i] : (int64_t)w[i]) + c;
    if ((uint32_t)t) top = i;
    c = t >> 32;
  }
  if (top < 0) return 0.0;
  // digits top, top - 1, top - 2 (96 bits, >= 65 significant) and a sticky bit for the rest
  uint64_t d2 = 0, d1 = 0, d0 = 0;
  bool sticky = false;
  c = 0;
  for (int i = 0; i <= top; ++i) {
    const int64_t t = (neg ? -(int64_t)w[i] : (int64_t)w[i]) + c;
    const uint64_t dig = (uint32_t)t;
    c = t >> 32;
    if (i == top) d2 = dig;
    else if (i == top - 1) d1 = dig;
    else if (i == top - 2) d0 = dig;
    else if (dig) sticky = true;
  }
  u128 q;
  q.hi = d2;
  q.lo = (d1 << 32) | d0 | (sticky ? 1ull : 0ull);
  const double v = fixed_to_f64(q, -(base + 32 * (top - 2)));
  return neg ? -v : v;
}

// Digit words for a column whose nonzero finite |x| lie in [min_abs, max_abs]:
// base = exponent of the lowest bit min_abs can have; room for the largest
// value's top digit, 2^31 terms and the sign.
QM_HD void wide_layout(double min_abs, double max_abs, int *base, int *nw) {
  const uint64_t bl = f64_bits(min_abs), bh = f64_bits(max_abs);
  const int el = (int)((bl >> 52) & 0x7FF), eh = (int)((bh >> 52) & 0x7FF);
  *base = (el == 0 ? -1074 : el - 1075);
  const int top = (eh == 0 ? -1074 : eh - 1075) + 53;  // |x| < 2^top
  *nw = ((top - *base) >> 5) + 3;                      // value digits + 2^31 terms + sign
}

}  // namespace qm

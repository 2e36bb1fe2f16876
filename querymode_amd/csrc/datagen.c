/* datagen.c — deterministic synthetic columns for the benchmark configs.
 *
 * Every value is a pure function of (seed, stream, row): a splitmix64 hash
 * of the three, so any row range can be generated independently (OpenMP),
 * and the GPU path and the oracle consume the same bytes.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline uint64_t mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

static inline uint64_t rng(uint64_t seed, uint32_t stream, uint64_t row) {
  uint64_t k = mix(seed * 0x9E3779B97F4A7C15ull + (uint64_t)stream * 0xD1B54A32D192ED03ull + 0x632BE59BD9B4E019ull);
  return mix(k + row * 0x9E3779B97F4A7C15ull);
}

static inline double u01(uint64_t r) { return (double)(r >> 11) * 0x1.0p-53; }
static inline uint64_t below(uint64_t r, uint64_t span) {
  return (uint64_t)(((unsigned __int128)r * span) >> 64);
}

void qg_fill_i64(uint64_t seed, uint32_t stream, uint64_t row0, uint64_t n, int64_t lo, uint64_t span, int64_t *out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i)
    out[i] = (int64_t)((uint64_t)lo + below(rng(seed, stream, row0 + i), span));
}

void qg_fill_i32(uint64_t seed, uint32_t stream, uint64_t row0, uint64_t n, int64_t lo, uint64_t span, int32_t *out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i)
    out[i] = (int32_t)(int64_t)((uint64_t)lo + below(rng(seed, stream, row0 + i), span));
}

void qg_fill_i8(uint64_t seed, uint32_t stream, uint64_t row0, uint64_t n, int64_t lo, uint64_t span, int8_t *out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i)
    out[i] = (int8_t)(int64_t)((uint64_t)lo + below(rng(seed, stream, row0 + i), span));
}

void qg_fill_uniform_f64(uint64_t seed, uint32_t stream, uint64_t row0, uint64_t n, double lo, double hi,
                         double *out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) out[i] = lo + (hi - lo) * u01(rng(seed, stream, row0 + i));
}

/* float32 in [0,1) scaled: 24 random bits, exact in float */
void qg_fill_uniform_f32(uint64_t seed, uint32_t stream, uint64_t row0, uint64_t n, float lo, float hi, float *out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    float u = (float)(rng(seed, stream, row0 + i) >> 40) * 0x1.0p-24f;
    out[i] = lo + (hi - lo) * u;
  }
}

static inline double normal_at(uint64_t seed, uint32_t stream, uint64_t row) {
  uint64_t a = rng(seed, stream, 2 * row), b = rng(seed, stream, 2 * row + 1);
  double u1 = ((double)(a >> 11) + 1.0) * 0x1.0p-53; /* (0, 1] */
  double u2 = u01(b);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

void qg_fill_normal_f64(uint64_t seed, uint32_t stream, uint64_t row0, uint64_t n, double mu, double sigma,
                        double *out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) out[i] = mu + sigma * normal_at(seed, stream, row0 + i);
}

void qg_fill_lognormal_f64(uint64_t seed, uint32_t stream, uint64_t row0, uint64_t n, double mu, double sigma,
                           double *out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) out[i] = exp(mu + sigma * normal_at(seed, stream, row0 + i));
}

void qg_fill_bernoulli_i8(uint64_t seed, uint32_t stream, uint64_t row0, uint64_t n, double p, int8_t *out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) out[i] = u01(rng(seed, stream, row0 + i)) < p ? 1 : 0;
}

/* Zipf(s) over ranks 1..nvals by inverse-CDF binary search. map != 0 turns a
 * rank into a sparse int64 id through a bijective 64-bit mix. */
void qg_fill_zipf_i64(uint64_t seed, uint32_t stream, uint64_t row0, uint64_t n, double s, uint64_t nvals, int map,
                      int64_t *out) {
  double *cdf = (double *)malloc(nvals * sizeof(double));
  if (!cdf) return;
  double acc = 0;
  for (uint64_t k = 0; k < nvals; ++k) { acc += pow((double)(k + 1), -s); cdf[k] = acc; }
  for (uint64_t k = 0; k < nvals; ++k) cdf[k] /= acc;
  cdf[nvals - 1] = 1.0;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    double u = u01(rng(seed, stream, row0 + i));
    uint64_t lo = 0, hi = nvals - 1;
    while (lo < hi) {
      uint64_t mid = (lo + hi) >> 1;
      if (cdf[mid] > u) hi = mid; else lo = mid + 1;
    }
    uint64_t rank = lo + 1;
    out[i] = map ? (int64_t)mix(rank ^ 0x5DEECE66Dull) : (int64_t)rank;
  }
  free(cdf);
}

/* A random permutation of base .. base+n-1 (Fisher-Yates). */
void qg_permutation_i64(uint64_t seed, uint32_t stream, uint64_t n, int64_t base, int64_t *out) {
  for (uint64_t i = 0; i < n; ++i) out[i] = base + (int64_t)i;
  for (uint64_t i = n; i > 1; --i) {
    uint64_t j = below(rng(seed, stream, i), i);
    int64_t t = out[i - 1]; out[i - 1] = out[j]; out[j] = t;
  }
}

/* Foreign keys: with probability match_rate an id of [0, nids) (the id set of
 * a permutation of 0..nids-1), otherwise a value of [nids, 2*nids). */
void qg_fill_fk_i64(uint64_t seed, uint32_t stream, uint64_t row0, uint64_t n, uint64_t nids, double match_rate,
                    int64_t *out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    uint64_t a = rng(seed, stream, 2 * (row0 + i)), b = rng(seed, stream, 2 * (row0 + i) + 1);
    uint64_t id = below(b, nids);
    out[i] = (int64_t)(u01(a) < match_rate ? id : nids + id);
  }
}

// Sorted sample of n uniform integers over [lo, lo + span): the order
// statistics of n uniform draws, from normalised partial sums of exponential
// spacings (E_i = -log u_i; U_(k) = (E_1 + ... + E_k) / (E_1 + ... + E_{n+1})).
// Deterministic in (seed, stream, n); parallel over chunks of rows.
void qg_sorted_uniform_i64(uint64_t seed, uint32_t stream, uint64_t n, int64_t lo, uint64_t span, int64_t *out) {
  const uint64_t chunk = 1ull << 20;
  const uint64_t nch = (n + 1 + chunk - 1) / chunk;
  double *sums = (double *)malloc((nch + 1) * sizeof(double));
  if (!sums) return;
#pragma omp parallel for schedule(static)
  for (uint64_t c = 0; c < nch; ++c) {
    uint64_t a = c * chunk, b = a + chunk < n + 1 ? a + chunk : n + 1;
    double s = 0;
    for (uint64_t i = a; i < b; ++i) s += -log(1.0 - u01(rng(seed, stream, i)));
    sums[c + 1] = s;
  }
  sums[0] = 0;
  for (uint64_t c = 0; c < nch; ++c) sums[c + 1] += sums[c];
  const double total = sums[nch];
#pragma omp parallel for schedule(static)
  for (uint64_t c = 0; c < nch; ++c) {
    uint64_t a = c * chunk, b = a + chunk < n ? a + chunk : n;
    double s = sums[c];
    for (uint64_t i = a; i < b; ++i) {
      s += -log(1.0 - u01(rng(seed, stream, i)));
      uint64_t v = (uint64_t)(s / total * (double)span);
      out[i] = lo + (int64_t)(v < span ? v : span - 1);
    }
  }
  free(sums);
}

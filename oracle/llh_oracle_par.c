/*
 * Multi-threaded oracle and seeded workload generator.
 *
 * TEST INFRASTRUCTURE ONLY: the checker that makes the full-size configurations
 * (n = 16384, 32768) checkable against the reference's own operation order, and
 * the CPU restatement of the synthetic inputs.  Only tests, the fixture script
 * and the benchmark's checks load it; the product never does.
 *
 * orc_par_solvec_d is bit-identical to orc_solvec_inplace_d (the serial
 * restatement of cholesky.rs:136-144 with triangular.rs:95-144) on a contiguous
 * row-major matrix.  The reference fixes every value by its formula:
 *   L[j,k] = (A[j,k] - s_jk) / L[k,k],  s_jk = sum_{i<k} L[k,i] L[j,i]  (i order, from +0)
 *   d_j = sum_{k<j} L[j,k]^2 (k order, from +0),  L[j,j] = sqrt(A[j,j] - d_j),
 * and the sweeps fix every b[r] as b[r] + (-coeff_i) x_r,i in the order the
 * reference visits i.  Any schedule that adds each sum's terms in that order,
 * one separately rounded multiply and add at a time, reproduces every bit; this
 * file keeps the partial sums s_jk in a separate array and applies a block of 64
 * columns at a time to all later rows (rows in parallel, columns vectorised),
 * each element still receiving its terms in increasing i.  Compiled with
 * -ffp-contract=off -fno-fast-math, like the serial oracle.
 *
 * The generator: element (i, j) of a rows x cols matrix drawn with (seed,
 * stream) is value(key, i * cols + j), key = splitmix64 output 0 of state
 * seed + stream * 0x9E3779B97F4A7C15, and the draws are splitmix64 outputs of
 * the state key (the SplitMix64 generator of Steele, Lea and Flood, OOPSLA
 * 2014, as published by Vigna: state += 0x9E3779B97F4A7C15, then the two
 * multiply-xorshift rounds).  Uniform draws take the top 53 bits; normal draws
 * use Marsaglia's polar method (v1, v2 uniform in [-1, 1), accept
 * 0 < s = v1^2 + v2^2 < 1, return v1 sqrt(-2 ln(s) / s)), with ln computed by
 * the atanh series below from exactly rounded operations only, so the device
 * kernel of the product reproduces every value bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "llh_oracle.h"

#define GOLDEN 0x9E3779B97F4A7C15ull

static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint64_t orc_splitmix64(uint64_t state, uint64_t c) { return mix64(state + GOLDEN * (c + 1)); }

uint64_t orc_gen_key(uint64_t seed, uint64_t stream) { return orc_splitmix64(seed + GOLDEN * stream, 0); }

static inline double u01(uint64_t u) { return (double)(u >> 11) * 0x1.0p-53; }

/* ln(s) for a positive normal s: s = m 2^e with m in [sqrt(1/2), sqrt(2)),
 * ln(m) = 2 atanh(f), f = (m - 1) / (m + 1), |f| <= 0.1716, by the series
 * 2 f (1 + w/3 + w^2/5 + ... + w^12/25), w = f^2 (the first omitted term is
 * below 2^-60 relative), then e ln 2 + ln(m). */
double orc_log(double s) {
  uint64_t bits;
  memcpy(&bits, &s, 8);
  int64_t e = (int64_t)((bits >> 52) & 0x7ff) - 1023;
  bits = (bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
  double m;
  memcpy(&m, &bits, 8);
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e += 1;
  }
  const double f = (m - 1.0) / (m + 1.0);
  const double w = f * f;
  double p = 1.0 / 25.0;
  p = p * w + 1.0 / 23.0;
  p = p * w + 1.0 / 21.0;
  p = p * w + 1.0 / 19.0;
  p = p * w + 1.0 / 17.0;
  p = p * w + 1.0 / 15.0;
  p = p * w + 1.0 / 13.0;
  p = p * w + 1.0 / 11.0;
  p = p * w + 1.0 / 9.0;
  p = p * w + 1.0 / 7.0;
  p = p * w + 1.0 / 5.0;
  p = p * w + 1.0 / 3.0;
  p = p * w + 1.0;
  const double lm = (2.0 * f) * p;
  return (double)e * 0.6931471805599453 + lm;
}

/* One value of distribution dist (ORC_GEN_NORMAL: N(0, 1); ORC_GEN_UNIFORM:
 * U[-1, 1)) at logical index idx. */
double orc_gen_value(int dist, uint64_t key, uint64_t idx) {
  if (dist == ORC_GEN_UNIFORM) return 2.0 * u01(orc_splitmix64(key, idx)) - 1.0;
  for (uint64_t t = 0; t < 64; ++t) {
    const uint64_t c = 2 * (idx * 64 + t);
    const double v1 = 2.0 * u01(orc_splitmix64(key, c)) - 1.0;
    const double v2 = 2.0 * u01(orc_splitmix64(key, c + 1)) - 1.0;
    const double s = v1 * v1 + v2 * v2;
    if (s < 1.0 && s != 0.0) {
      const double l = orc_log(s);
      const double q = (-2.0 * l) / s;
      return v1 * sqrt(q);
    }
  }
  return 0.0; /* 64 rejections in a row: probability below 1e-42 */
}

void orc_gen_matrix_d(int dist, uint64_t seed, uint64_t stream, int64_t rows, int64_t cols, double* x,
                      int64_t ldx) {
  const uint64_t key = orc_gen_key(seed, stream);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < rows; ++i)
    for (int64_t j = 0; j < cols; ++j) x[i * ldx + j] = orc_gen_value(dist, key, (uint64_t)(i * cols + j));
}

/* Rows [row0, row0 + nrows) of the rows x cols draw (any rows >= row0 + nrows),
 * into x (nrows x cols, contiguous). */
void orc_gen_rows_d(int dist, uint64_t seed, uint64_t stream, int64_t cols, int64_t row0, int64_t nrows,
                    double* x) {
  const uint64_t key = orc_gen_key(seed, stream);
  for (int64_t i = 0; i < nrows; ++i)
    for (int64_t j = 0; j < cols; ++j)
      x[i * cols + j] = orc_gen_value(dist, key, (uint64_t)((row0 + i) * cols + j));
}

/* A = B B^T / n + lam I with B n x n drawn with (seed, stream 0):
 * a_ij = (sum_k b_ik b_jk, k in order from +0) / n, plus lam on the diagonal
 * (one more rounded add); the upper triangle mirrors the lower.  b_work: n x n. */
void orc_gen_spd_d(int dist, uint64_t seed, int64_t n, double lam, double* a, int64_t lda, double* b_work) {
  orc_gen_matrix_d(dist, seed, 0, n, n, b_work, n);
  const double dn = (double)n;
  const int64_t TB = 32, KC = 256;
  const int64_t nt = (n + TB - 1) / TB;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t t = 0; t < nt * (nt + 1) / 2; ++t) {
    /* lower tile (ti, tj), tj <= ti */
    int64_t ti = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) / 2.0);
    while (ti * (ti + 1) / 2 > t) --ti;
    while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
    const int64_t tj = t - ti * (ti + 1) / 2;
    const int64_t i0 = ti * TB, j0 = tj * TB;
    const int64_t ie = i0 + TB < n ? i0 + TB : n, je = j0 + TB < n ? j0 + TB : n;
    double acc[32][32];
    double bt[256][32]; /* bt[k][jj] = b[j0 + jj][k0 + k] */
    memset(acc, 0, sizeof(acc));
    for (int64_t k0 = 0; k0 < n; k0 += KC) {
      const int64_t kc = k0 + KC < n ? KC : n - k0;
      for (int64_t jj = 0; jj < 32; ++jj)
        for (int64_t k = 0; k < kc; ++k) bt[k][jj] = jj < je - j0 ? b_work[(j0 + jj) * n + k0 + k] : 0.0;
      for (int64_t ii = 0; ii < ie - i0; ++ii) {
        const double* bi = b_work + (i0 + ii) * n + k0;
        double* ac = acc[ii];
        for (int64_t k = 0; k < kc; ++k) {
          const double x = bi[k];
          const double* br = bt[k];
          for (int64_t jj = 0; jj < 32; ++jj) ac[jj] = ac[jj] + x * br[jj];
        }
      }
    }
    for (int64_t i = i0; i < ie; ++i)
      for (int64_t j = j0; j < je && j <= i; ++j) {
        double v = acc[i - i0][j - j0] / dn;
        if (i == j) v = v + lam;
        a[i * lda + j] = v;
        a[j * lda + i] = v;
      }
  }
}

/* The dirty factor in place (row-major, lda), bit-identical to the serial
 * oracle's cholesky_inplace_dirty.  On failure *fail = the first row whose
 * pivot fails `d <= 0` and rows below it hold partial results; else -1. */
int orc_par_cholesky_dirty_d(int64_t n, double* a, int64_t lda, int64_t* fail) {
  const int64_t NB = 64, KC = 512;
  *fail = -1;
  if (n <= 0) return ORC_OK;
  double* S = (double*)calloc((size_t)n * (size_t)n, sizeof(double)); /* s_jk for k < j */
  double* d = (double*)calloc((size_t)n, sizeof(double));
  double* lt = (double*)malloc((size_t)NB * (size_t)n * sizeof(double)); /* lt[i][k'] = L[k'][K0 + i] */
  if (!S || !d || !lt) {
    free(S);
    free(d);
    free(lt);
    return -1;
  }
  int status = ORC_OK;
  for (int64_t K0 = 0; K0 < n && status == ORC_OK; K0 += NB) {
    const int64_t K1 = K0 + NB < n ? K0 + NB : n;
    /* rows of the block itself, in order: columns K0..j-1 of the block, then the pivot */
    for (int64_t j = K0; j < K1; ++j) {
      double* sj = S + j * n;
      for (int64_t k = K0; k < j; ++k) {
        const double l = (a[j * lda + k] - sj[k]) / a[k * lda + k];
        a[j * lda + k] = l;
        d[j] = d[j] + l * l;
        for (int64_t k2 = k + 1; k2 < j; ++k2) sj[k2] = sj[k2] + a[k2 * lda + k] * l;
      }
      const double dj = a[j * lda + j] - d[j];
      if (dj <= 0.0) {
        *fail = j;
        status = ORC_NOT_POSITIVE_DEFINITE;
        break;
      }
      a[j * lda + j] = sqrt(dj);
    }
    if (status != ORC_OK) break;
    /* rows below the block: the block's columns, each row on its own */
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t j = K1; j < n; ++j) {
      double* sj = S + j * n;
      double* aj = a + j * lda;
      for (int64_t k = K0; k < K1; ++k) {
        const double l = (aj[k] - sj[k]) / a[k * lda + k];
        aj[k] = l;
        d[j] = d[j] + l * l;
        for (int64_t k2 = k + 1; k2 < K1; ++k2) sj[k2] = sj[k2] + a[k2 * lda + k] * l;
      }
    }
    /* the block's terms of every later sum: S[j][k'] += L[k'][i] L[j][i], i in order */
    const int64_t w = K1 - K0;
#pragma omp parallel for schedule(static)
    for (int64_t k2 = K1; k2 < n; ++k2)
      for (int64_t i = 0; i < w; ++i) lt[i * n + k2] = a[k2 * lda + K0 + i];
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t j = K1 + 1; j < n; ++j) {
      double* sj = S + j * n;
      const double* aj = a + j * lda + K0;
      for (int64_t c0 = K1; c0 < j; c0 += KC) {
        const int64_t c1 = c0 + KC < j ? c0 + KC : j;
        for (int64_t i = 0; i < w; ++i) {
          const double lji = aj[i];
          const double* li = lt + i * n;
          for (int64_t k2 = c0; k2 < c1; ++k2) sj[k2] = sj[k2] + li[k2] * lji;
        }
      }
    }
  }
  free(S);
  free(d);
  free(lt);
  return status;
}

/* L y = b then L^T x = y for b n x m row-major (ldb), in place, each element
 * receiving its terms in the reference's order (triangular.rs:113-125: rows
 * 0..n for the lower solve, n-1..0 for the upper solve on L.t()). */
void orc_par_solve_lower_upper_d(int64_t n, const double* a, int64_t lda, int64_t m, double* b, int64_t ldb) {
  const int64_t NB = 64;
  double* coeff = (double*)malloc((size_t)NB * (size_t)(m > 0 ? m : 1) * sizeof(double));
  if (!coeff) return;
  /* forward: block [I0, I1) final, then every later row gets the block's terms in i order */
  for (int64_t I0 = 0; I0 < n; I0 += NB) {
    const int64_t I1 = I0 + NB < n ? I0 + NB : n;
    for (int64_t i = I0; i < I1; ++i) {
      for (int64_t c = 0; c < m; ++c) {
        const double cf = b[i * ldb + c] / a[i * lda + i];
        b[i * ldb + c] = cf;
        coeff[(i - I0) * m + c] = -cf;
      }
      for (int64_t r = i + 1; r < I1; ++r)
        for (int64_t c = 0; c < m; ++c) b[r * ldb + c] = b[r * ldb + c] + coeff[(i - I0) * m + c] * a[r * lda + i];
    }
#pragma omp parallel for schedule(static)
    for (int64_t r = I1; r < n; ++r)
      for (int64_t i = I0; i < I1; ++i) {
        const double x = a[r * lda + i];
        const double* cf = coeff + (i - I0) * m;
        for (int64_t c = 0; c < m; ++c) b[r * ldb + c] = b[r * ldb + c] + cf[c] * x;
      }
  }
  /* backward on U = L^T: U[r][i] = L[i][r]; rows n-1..0 */
  for (int64_t I1 = n; I1 > 0; I1 -= NB) {
    const int64_t I0 = I1 - NB > 0 ? I1 - NB : 0;
    for (int64_t i = I1 - 1; i >= I0; --i) {
      for (int64_t c = 0; c < m; ++c) {
        const double cf = b[i * ldb + c] / a[i * lda + i];
        b[i * ldb + c] = cf;
        coeff[(i - I0) * m + c] = -cf;
      }
      for (int64_t r = I0; r < i; ++r)
        for (int64_t c = 0; c < m; ++c) b[r * ldb + c] = b[r * ldb + c] + coeff[(i - I0) * m + c] * a[i * lda + r];
    }
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < I0; ++r)
      for (int64_t i = I1 - 1; i >= I0; --i) {
        const double x = a[i * lda + r];
        const double* cf = coeff + (i - I0) * m;
        for (int64_t c = 0; c < m; ++c) b[r * ldb + c] = b[r * ldb + c] + cf[c] * x;
      }
  }
  free(coeff);
}

/* solvec_inplace (cholesky.rs:136-144) on a contiguous row-major A (lda) and
 * B (n x m, ldb); A is left holding the dirty factor. */
int orc_par_solvec_d(int64_t n, double* a, int64_t lda, int64_t m, double* b, int64_t ldb, int64_t* fail) {
  const int st = orc_par_cholesky_dirty_d(n, a, lda, fail);
  if (st != ORC_OK) return st;
  orc_par_solve_lower_upper_d(n, a, lda, m, b, ldb);
  return ORC_OK;
}

/*
 * CPU oracle for linfa-linalg 0.2.1's Cholesky / triangular-solve path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU baseline
 * timed beside it.  Only tests, the smoke check and the benchmark's CPU-baseline
 * leg load this library; the product never does.
 *
 * Every function restates the reference operation for operation (loop order,
 * operand order, comparison) so that, compiled with -ffp-contract=off
 * -fno-fast-math on an SSE2 target, it rounds exactly as the Rust code does:
 * rustc never contracts a*b+c into an FMA and the crate never calls mul_add.
 * Element (i, j) of a view lives at p[i*rs + j*cs]; strides may be negative.
 */
#include "llh_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define AT(p, i, j, rs, cs) (p)[(int64_t)(i) * (rs) + (int64_t)(j) * (cs)]

/* ORC_MUTANT = k > 0 builds a copy with one plausible mistake planted (a dropped
 * term, a wrong sign, a transposed operand, an off-by-one pivot test ...).  The
 * oracle's own tests build every mutant and require the known-answer checks to
 * reject each one; the faithful oracle is ORC_MUTANT = 0. */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif
#define MUT(k) (ORC_MUTANT == (k))

static void set_err(orc_err* e, int code, int64_t a, int64_t b, int64_t col) {
  if (!e) return;
  e->code = code;
  e->a = a;
  e->b = b;
  e->column = col;
}

#define ORC_IMPL(T, sfx, SQRT)                                                         \
  /* check_square, lib.rs:64-71 */                                                     \
  static int check_square_##sfx(int64_t rows, int64_t cols, orc_err* e) {              \
    if (rows != cols) {                                                                \
      set_err(e, ORC_NOT_SQUARE, rows, cols, -1);                                      \
      return ORC_NOT_SQUARE;                                                           \
    }                                                                                  \
    return ORC_OK;                                                                     \
  }                                                                                    \
                                                                                       \
  /* CholeskyInplace::cholesky_inplace_dirty, cholesky.rs:51-76.  Row-wise            \
   * Cholesky-Banachiewicz: s accumulates L[k,i]*L[j,i] for i = 0..k-1, then          \
   * L[j,k] = (A[j,k] - s) / L[k,k]; d accumulates L[j,k]^2 in k order; the pivot     \
   * test is `d <= 0` (a NaN pivot passes).  On failure at row j, rows < j hold      \
   * L, row j holds L[j,0..j-1] and the original A[j,j], later rows are untouched.   \
   */                                                                                  \
  int orc_cholesky_inplace_dirty_##sfx(int64_t rows, int64_t cols, T* a, int64_t rs,   \
                                       int64_t cs, orc_err* e) {                       \
    set_err(e, ORC_OK, 0, 0, -1);                                                      \
    if (check_square_##sfx(rows, cols, e)) return ORC_NOT_SQUARE;                      \
    const int64_t n = rows;                                                            \
    for (int64_t j = 0; j < n; ++j) {                                                  \
      T d = (T)0;                                                                      \
      for (int64_t k = 0; k < j; ++k) {                                                \
        T s = (T)0;                                                                    \
        for (int64_t i = 0; i < k - MUT(1); ++i)                                         \
          s += (MUT(3) ? AT(a, i, k, rs, cs) : AT(a, k, i, rs, cs)) * AT(a, j, i, rs, cs); \
        s = (AT(a, j, k, rs, cs) - (MUT(2) ? -s : s)) / AT(a, k, k, rs, cs);           \
        AT(a, j, k, rs, cs) = s;                                                       \
        d += s * s;                                                                    \
      }                                                                                \
      d = AT(a, j, j, rs, cs) - d;                                                     \
      if (MUT(4) ? d < (T)0 : d <= (T)0) {                                             \
        set_err(e, ORC_NOT_POSITIVE_DEFINITE, 0, 0, j);                                \
        return ORC_NOT_POSITIVE_DEFINITE;                                              \
      }                                                                                \
      AT(a, j, j, rs, cs) = SQRT(d);                                                   \
    }                                                                                  \
    return ORC_OK;                                                                     \
  }                                                                                    \
                                                                                       \
  /* IntoTriangular::triangular_inplace, triangular.rs:37-53 */                        \
  int orc_triangular_inplace_##sfx(int64_t rows, int64_t cols, T* a, int64_t rs,       \
                                   int64_t cs, int uplo, orc_err* e) {                 \
    set_err(e, ORC_OK, 0, 0, -1);                                                      \
    if (check_square_##sfx(rows, cols, e)) return ORC_NOT_SQUARE;                      \
    const int64_t n = rows;                                                            \
    if (uplo == ORC_UPPER) {                                                           \
      for (int64_t i = 0; i < n; ++i)                                                  \
        for (int64_t j = 0; j < i; ++j) AT(a, i, j, rs, cs) = (T)0;                    \
    } else {                                                                           \
      for (int64_t i = 0; i < n; ++i)                                                  \
        for (int64_t j = i + 1; j < n; ++j) AT(a, i, j, rs, cs) = (T)0;                \
    }                                                                                  \
    return ORC_OK;                                                                     \
  }                                                                                    \
                                                                                       \
  /* CholeskyInplace::cholesky_inplace, cholesky.rs:78-82 */                           \
  int orc_cholesky_inplace_##sfx(int64_t rows, int64_t cols, T* a, int64_t rs,         \
                                 int64_t cs, orc_err* e) {                             \
    int st = orc_cholesky_inplace_dirty_##sfx(rows, cols, a, rs, cs, e);               \
    if (st) return st;                                                                 \
    return orc_triangular_inplace_##sfx(rows, cols, a, rs, cs, ORC_LOWER, e);          \
  }                                                                                    \
                                                                                       \
  /* Triangular::is_triangular, triangular.rs:67-90 (`is_zero` is `== 0`, so -0.0    \
   * counts as zero and NaN does not) */                                               \
  int orc_is_triangular_##sfx(int64_t rows, int64_t cols, const T* a, int64_t rs,      \
                              int64_t cs, int uplo) {                                  \
    if (rows != cols) return 0;                                                        \
    const int64_t n = rows;                                                            \
    if (uplo == ORC_UPPER) {                                                           \
      for (int64_t i = 0; i < n; ++i)                                                  \
        for (int64_t j = 0; j < i; ++j)                                                \
          if (!(AT(a, i, j, rs, cs) == (T)0)) return 0;                                \
    } else {                                                                           \
      for (int64_t i = 0; i < n; ++i)                                                  \
        for (int64_t j = i + 1; j < n; ++j)                                            \
          if (!(AT(a, i, j, rs, cs) == (T)0)) return 0;                                \
    }                                                                                  \
    return 1;                                                                          \
  }                                                                                    \
                                                                                       \
  /* solve_triangular_system_common, triangular.rs:95-128, with the dispatch of      \
   * solve_triangular_system, triangular.rs:133-144.  For each RHS column k, rows    \
   * in order (Lower: 0..n, Upper: n-1..0): coeff = b[i,k] / diag_fn(i);             \
   * b[i,k] = coeff; then ndarray's scaled_add(-coeff, x) on the rest of the column, \
   * which is y = y + (alpha * x) element by element (ndarray 0.16                   \
   * ArrayBase::scaled_add).  Lower uses a[i+1.., i] and b[i+1.., k]; Upper uses     \
   * a[..i, i] and b[..i, k], so a's diagonal is read only through diag_fn.          \
   * diag_fn is a[i,i] (triangular.rs:166) when dg is NULL, else dg[i * ds], or its  \
   * absolute value when absd (QR's |diag_i|, qr.rs:145-150, 172-177). */            \
  static int solve_tri_common_##sfx(int64_t rows, int64_t cols, const T* a,          \
                                    int64_t rs, int64_t cs, const T* dg, int64_t ds,  \
                                    int absd, int64_t brows, int64_t bcols, T* b,     \
                                    int64_t brs, int64_t bcs, int uplo, orc_err* e) { \
    set_err(e, ORC_OK, 0, 0, -1);                                                      \
    if (check_square_##sfx(rows, cols, e)) return ORC_NOT_SQUARE;                      \
    const int64_t n = rows;                                                            \
    if (brows != n) {                                                                  \
      set_err(e, ORC_WRONG_ROWS, n, brows, -1);                                        \
      return ORC_WRONG_ROWS;                                                           \
    }                                                                                  \
    for (int64_t k = 0; k < bcols; ++k) {                                              \
      for (int64_t t = 0; t < n; ++t) {                                                \
        const int64_t i = (uplo == ORC_UPPER) ? (n - 1 - t) : t;                       \
        T diag = dg ? dg[i * ds] : AT(a, i, i, rs, cs);                                \
        if (dg && absd && !MUT(9) && diag < (T)0) diag = -diag;                        \
        const T coeff = AT(b, i, k, brs, bcs) / diag;                                  \
        AT(b, i, k, brs, bcs) = coeff;                                                 \
        const T alpha = MUT(5) ? coeff : -coeff;                                       \
        if (uplo == ORC_UPPER) {                                                       \
          for (int64_t r = 0; r < i + MUT(6) * (i + 1 < n); ++r)                       \
            AT(b, r, k, brs, bcs) = AT(b, r, k, brs, bcs) + (alpha * AT(a, r, i, rs, cs)); \
        } else {                                                                       \
          for (int64_t r = i + 1; r < n; ++r)                                          \
            AT(b, r, k, brs, bcs) = AT(b, r, k, brs, bcs) + (alpha * AT(a, r, i, rs, cs)); \
        }                                                                              \
      }                                                                                \
    }                                                                                  \
    return ORC_OK;                                                                     \
  }                                                                                    \
                                                                                       \
  /* SolveTriangularInplace::solve_triangular_inplace, triangular.rs:147-169. */      \
  int orc_solve_triangular_inplace_##sfx(int64_t rows, int64_t cols, const T* a,       \
                                         int64_t rs, int64_t cs, int64_t brows,        \
                                         int64_t bcols, T* b, int64_t brs,             \
                                         int64_t bcs, int uplo, orc_err* e) {          \
    return solve_tri_common_##sfx(rows, cols, a, rs, cs, NULL, 0, 0, brows, bcols, b,  \
                                  brs, bcs, uplo, e);                                  \
  }                                                                                    \
                                                                                       \
  /* solve_triangular_system with a diag_fn that reads a separate vector             \
   * (triangular.rs:133-144; the QR callers, qr.rs:145-150 and 172-177). */           \
  int orc_solve_triangular_system_##sfx(int64_t rows, int64_t cols, const T* a,        \
                                        int64_t rs, int64_t cs, const T* dg,           \
                                        int64_t ds, int absd, int64_t brows,           \
                                        int64_t bcols, T* b, int64_t brs, int64_t bcs, \
                                        int uplo, orc_err* e) {                        \
    return solve_tri_common_##sfx(rows, cols, a, rs, cs, dg, ds, absd, brows, bcols,   \
                                  b, brs, bcs, uplo, e);                               \
  }                                                                                    \
                                                                                       \
  /* SolveCInplace::solvec_inplace, cholesky.rs:136-144: A := dirty factor, then     \
   * L y = b (Lower), then L^T x = y (Upper on the transposed view L.t(), whose      \
   * element (r, c) is L[c, r], i.e. strides swapped). */                              \
  int orc_solvec_inplace_##sfx(int64_t rows, int64_t cols, T* a, int64_t rs,           \
                               int64_t cs, int64_t brows, int64_t bcols, T* b,         \
                               int64_t brs, int64_t bcs, orc_err* e) {                 \
    int st = orc_cholesky_inplace_dirty_##sfx(rows, cols, a, rs, cs, e);               \
    if (st) return st;                                                                 \
    st = orc_solve_triangular_inplace_##sfx(rows, cols, a, rs, cs, brows, bcols, b,    \
                                            brs, bcs, ORC_LOWER, e);                   \
    if (st) return st;                                                                 \
    return orc_solve_triangular_inplace_##sfx(cols, rows, a, MUT(7) ? rs : cs,          \
                                              MUT(7) ? cs : rs, brows, bcols, b,       \
                                              brs, bcs, ORC_UPPER, e);                 \
  }                                                                                    \
                                                                                       \
  /* InverseCInplace::invc_inplace, cholesky.rs:178-182: solvec_into(eye(n)).        \
   * `out` receives the n x n result in C order (Array2::eye is C order).  Like the  \
   * reference, eye(nrows) is built before the squareness check, so a non-square A  \
   * reports NotSquare from the factor. */                                             \
  int orc_invc_inplace_##sfx(int64_t rows, int64_t cols, T* a, int64_t rs, int64_t cs, \
                             T* out, orc_err* e) {                                     \
    for (int64_t i = 0; i < rows; ++i)                                                 \
      for (int64_t j = 0; j < rows; ++j) out[i * rows + j] = (i == j) ? (T)1 : (T)0;   \
    if (MUT(8) && rows > 1) out[1] = (T)1;                                             \
    return orc_solvec_inplace_##sfx(rows, cols, a, rs, cs, rows, rows, out, rows, 1,   \
                                    e);                                                \
  }

ORC_IMPL(double, d, sqrt)
ORC_IMPL(float, s, sqrtf)

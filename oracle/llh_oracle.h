/*
 * CPU oracle for the Cholesky / triangular-solve path of linfa-linalg 0.2.1.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against (and the CPU baseline timed beside it).  It is never linked
 * into, loaded by or called from the product library.
 *
 * Each function is an operation-for-operation restatement of the reference
 * Rust code it cites, over a strided 2-D view (element strides may be negative,
 * as produced by ndarray's invert_axis / reversed_axes).  Build it with
 * -ffp-contract=off -fno-fast-math so that every `s += x*y` stays a separate,
 * correctly rounded multiply and add, as rustc emits it.
 */
#ifndef LLH_ORACLE_H
#define LLH_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; the same numbering as the product C ABI (llh.h). */
enum {
  ORC_OK = 0,
  ORC_NOT_SQUARE = 1,            /* LinalgError::NotSquare          lib.rs:37-38 */
  ORC_NOT_POSITIVE_DEFINITE = 2, /* LinalgError::NotPositiveDefinite lib.rs:43-44 */
  ORC_WRONG_ROWS = 3             /* LinalgError::WrongRows          lib.rs:55-56 */
};
enum { ORC_UPPER = 0, ORC_LOWER = 1 }; /* UPLO, triangular.rs:10-13 */

/* Error payload: NotSquare{rows, cols}, WrongRows{expected, actual}.  `column`
 * is the row j whose pivot failed the `d <= 0` test (the reference reports no
 * index; the oracle records it so the GPU path's index can be checked). */
typedef struct {
  int32_t code;
  int64_t a, b;
  int64_t column;
} orc_err;

#define ORC_DECL(T, sfx)                                                            \
  int orc_cholesky_inplace_dirty_##sfx(int64_t rows, int64_t cols, T* a,            \
                                       int64_t rs, int64_t cs, orc_err* e);         \
  int orc_cholesky_inplace_##sfx(int64_t rows, int64_t cols, T* a, int64_t rs,      \
                                 int64_t cs, orc_err* e);                           \
  int orc_triangular_inplace_##sfx(int64_t rows, int64_t cols, T* a, int64_t rs,    \
                                   int64_t cs, int uplo, orc_err* e);               \
  int orc_is_triangular_##sfx(int64_t rows, int64_t cols, const T* a, int64_t rs,   \
                              int64_t cs, int uplo);                                \
  int orc_solve_triangular_inplace_##sfx(int64_t rows, int64_t cols, const T* a,    \
                                         int64_t rs, int64_t cs, int64_t brows,     \
                                         int64_t bcols, T* b, int64_t brs,          \
                                         int64_t bcs, int uplo, orc_err* e);        \
  int orc_solve_triangular_system_##sfx(int64_t rows, int64_t cols, const T* a,     \
                                        int64_t rs, int64_t cs, const T* dg,        \
                                        int64_t ds, int absd, int64_t brows,        \
                                        int64_t bcols, T* b, int64_t brs,           \
                                        int64_t bcs, int uplo, orc_err* e);         \
  int orc_solvec_inplace_##sfx(int64_t rows, int64_t cols, T* a, int64_t rs,        \
                               int64_t cs, int64_t brows, int64_t bcols, T* b,      \
                               int64_t brs, int64_t bcs, orc_err* e);               \
  int orc_invc_inplace_##sfx(int64_t rows, int64_t cols, T* a, int64_t rs,          \
                             int64_t cs, T* out, orc_err* e);

ORC_DECL(double, d)
ORC_DECL(float, s)
#undef ORC_DECL

/* llh_oracle_par.c: the seeded workload generator and the multi-threaded,
 * bit-identical restatement of solvec for contiguous row-major f64 matrices. */
enum { ORC_GEN_NORMAL = 0, ORC_GEN_UNIFORM = 1 };
uint64_t orc_splitmix64(uint64_t state, uint64_t c); /* output c (0-based) of SplitMix64 from state */
uint64_t orc_gen_key(uint64_t seed, uint64_t stream);
double orc_log(double s);
double orc_gen_value(int dist, uint64_t key, uint64_t idx);
void orc_gen_matrix_d(int dist, uint64_t seed, uint64_t stream, int64_t rows, int64_t cols, double* x,
                      int64_t ldx);
void orc_gen_rows_d(int dist, uint64_t seed, uint64_t stream, int64_t cols, int64_t row0, int64_t nrows,
                    double* x);
void orc_gen_spd_d(int dist, uint64_t seed, int64_t n, double lam, double* a, int64_t lda, double* b_work);
int orc_par_cholesky_dirty_d(int64_t n, double* a, int64_t lda, int64_t* fail);
void orc_par_solve_lower_upper_d(int64_t n, const double* a, int64_t lda, int64_t m, double* b, int64_t ldb);
int orc_par_solvec_d(int64_t n, double* a, int64_t lda, int64_t m, double* b, int64_t ldb, int64_t* fail);

#ifdef __cplusplus
}
#endif
#endif

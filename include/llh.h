/*
 * llh.h - C ABI of linfa-linalg-hip: the Cholesky / SPD-solve / triangular-solve
 * path of linfa-linalg 0.2.1, computed by hand-written HIP kernels on one
 * MI355X (gfx950).
 *
 * The reference's interface is a set of Rust trait extension methods on
 * ndarray::ArrayBase<S, Ix2> with A: NdFloat (f32 | f64).  Each "plain" entry
 * point below replaces one of those methods for host-memory arrays of any
 * (signed) element strides, so a Rust binding can pass `arr.as_mut_ptr()`,
 * `arr.dim()` and `arr.strides()` straight through:
 *
 *   element (i, j) of a view lives at  ptr[i * rs + j * cs]
 *
 * and `ptr` addresses element (0, 0) (what ndarray's as_ptr returns, also for
 * inverted axes).  The in-place and "dirty" semantics of the reference are kept:
 * only the lower triangle of A is read, the dirty forms leave the strict upper
 * triangle untouched, and solvec/invc leave A holding the dirty factor.
 *
 * Errors mirror LinalgError (lib.rs:33-60): every entry point returns an
 * llh_status and, when `err` is non-null, fills it with the payload.  Checks run
 * in the reference's order: NotSquare, then NotPositiveDefinite, then WrongRows.
 *
 * The "dev" entry points take device-resident row-major matrices (element (i, j)
 * at ptr[i * ld + j]) and are what the throughput figures time.  The factor,
 * batch and server objects are extensions with no counterpart in the reference.
 *
 * Threading: a context owns one HIP stream; calls on one context must not
 * overlap (lock externally to share it across threads).  Calls return after the
 * work is complete unless stated otherwise.
 */
#ifndef LLH_H
#define LLH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LLH_ABI_VERSION 1

typedef enum llh_status {
  LLH_OK = 0,
  LLH_ERR_NOT_SQUARE = 1,             /* LinalgError::NotSquare{rows, cols}     lib.rs:37-38 */
  LLH_ERR_NOT_POSITIVE_DEFINITE = 2,  /* LinalgError::NotPositiveDefinite        lib.rs:43-44 */
  LLH_ERR_WRONG_ROWS = 3,             /* LinalgError::WrongRows{expected,actual} lib.rs:55-56 */
  /* Extensions (LinalgError is #[non_exhaustive], lib.rs:34): */
  LLH_ERR_INVALID_ARGUMENT = 16,      /* null pointer, bad leading dimension, bad option */
  LLH_ERR_HIP = 17,                   /* a HIP runtime call failed; see hip_error */
  LLH_ERR_OUT_OF_MEMORY = 18,         /* device allocation failed / over the HBM budget */
  LLH_ERR_NO_CONVERGENCE = 19,        /* iterative refinement stalled (result still returned) */
  LLH_ERR_REJECTED = 20               /* server admission refused the request */
} llh_status;

typedef enum llh_uplo { LLH_UPPER = 0, LLH_LOWER = 1 } llh_uplo; /* UPLO, triangular.rs:10-13 */

typedef struct llh_error {
  int32_t code;      /* llh_status */
  int32_t hip_error; /* hipError_t when code == LLH_ERR_HIP */
  uint64_t a;        /* NotSquare: rows;  WrongRows: expected */
  uint64_t b;        /* NotSquare: cols;  WrongRows: actual   */
  int64_t column;    /* NotPositiveDefinite: first row j whose pivot fails `d <= 0`; else -1 */
  char message[192];
} llh_error;

typedef struct llh_context llh_context;

/* One context per (thread, device).  device < 0 uses the current device. */
llh_status llh_context_create(int device, llh_context** out, llh_error* err);
void llh_context_destroy(llh_context* ctx);
/* Run on a caller-owned hipStream_t instead of the context's own stream. */
llh_status llh_context_set_stream(llh_context* ctx, void* hip_stream, llh_error* err);
void* llh_context_stream(llh_context* ctx);
const char* llh_status_string(int32_t code);

/* Live kernel timing (no counterpart in the reference; for benchmarks).  While
 * enabled, the device-resident drivers bracket each launch of a kernel class
 * with HIP events on the context's stream and add the launch's algorithmic
 * flops and bytes.  Enabling clears earlier records. */
typedef enum llh_kernel_class {
  LLH_KC_TRAILING_SYRK = 0, /* trailing update of the POTRF, rank = the outer panel width (MFMA) */
  LLH_KC_PANEL_UPDATE = 1,  /* updates inside the outer panel and of the look-ahead panel (MFMA) */
  LLH_KC_DIAG_FACTOR = 2,   /* 64 x 64 diagonal-block factor in LDS */
  LLH_KC_PANEL_TRSM = 3,    /* panel row solve X := X L11^-T */
  LLH_KC_SOLVE_GEMM = 4,    /* block update of the triangular solves (MFMA) */
  LLH_KC_SOLVE_DIAG = 5     /* diagonal-block solve of the triangular solves */
} llh_kernel_class;

typedef struct llh_kernel_stats {
  uint64_t launches;
  double total_ms; /* sum of event-timed launch durations */
  double flops;    /* sum of algorithmic flops */
  double bytes;    /* sum of algorithmic bytes */
} llh_kernel_stats;

llh_status llh_context_set_profiling(llh_context* ctx, int enable, llh_error* err);
/* As llh_context_set_profiling(ctx, 1, err), but records only the classes whose
 * bit (1 << llh_kernel_class) is set in class_mask (0: off): fewer events on
 * the streams when only some classes are wanted. */
llh_status llh_context_set_profiling_classes(llh_context* ctx, uint32_t class_mask, llh_error* err);
/* Synchronizes the context's stream, then sums the records of one class. */
llh_status llh_context_kernel_stats(llh_context* ctx, int kernel_class, llh_kernel_stats* out,
                                    llh_error* err);

/* Bound on each flag wait of the one-launch triangular solve (no counterpart in
 * the reference; for tests of the error path).  Past `limit` polls a waiting
 * workgroup gives up and the call reports LLH_ERR_HIP ("a block wait exceeded
 * its bound") instead of hanging; limit < 0 makes every wait give up at its
 * first poll, so a solve of more than one 64-row block fails at once while every
 * kernel still exits normally.  0 restores the default (2^22 polls). */
llh_status llh_context_set_solve_spin_limit(llh_context* ctx, int64_t limit, llh_error* err);

/* ---------------------------------------------------------------------------
 * Plain forms: host memory, any strides.  _d = f64, _s = f32.
 * ------------------------------------------------------------------------- */
#define LLH_PLAIN_DECL(T, sfx)                                                            \
  /* CholeskyInplace::cholesky_inplace_dirty, cholesky.rs:19 (impl :51-76).  On          \
     NotPositiveDefinite at row j the array is left as the reference leaves it: rows    \
     < j hold L, row j holds L[j, ..j] and its original diagonal, later rows are        \
     untouched. */                                                                       \
  llh_status llh_cholesky_inplace_dirty_##sfx(llh_context* ctx, size_t rows, size_t cols, \
                                              T* a, ptrdiff_t rs, ptrdiff_t cs,           \
                                              llh_error* err);                            \
  /* CholeskyInplace::cholesky_inplace, cholesky.rs:33 (impl :78-82): dirty, then         \
     triangular_inplace(Lower). */                                                       \
  llh_status llh_cholesky_inplace_##sfx(llh_context* ctx, size_t rows, size_t cols, T* a,  \
                                        ptrdiff_t rs, ptrdiff_t cs, llh_error* err);      \
  /* IntoTriangular::triangular_inplace, triangular.rs:19 (impl :37-53). */              \
  llh_status llh_triangular_inplace_##sfx(llh_context* ctx, size_t rows, size_t cols,     \
                                          T* a, ptrdiff_t rs, ptrdiff_t cs,               \
                                          llh_uplo uplo, llh_error* err);                 \
  /* Triangular::is_triangular, triangular.rs:59 (impl :67-90).  *out = 1 or 0. */       \
  llh_status llh_is_triangular_##sfx(llh_context* ctx, size_t rows, size_t cols,          \
                                     const T* a, ptrdiff_t rs, ptrdiff_t cs,              \
                                     llh_uplo uplo, int* out, llh_error* err);            \
  /* SolveTriangularInplace::solve_triangular_inplace, triangular.rs:149 (impl           \
     :161-168, kernel :95-128).  Reads only the `uplo` triangle and the diagonal of a;   \
     a zero pivot yields inf/NaN, not an error (triangular.rs:274-276). */               \
  llh_status llh_solve_triangular_inplace_##sfx(                                           \
      llh_context* ctx, size_t rows, size_t cols, const T* a, ptrdiff_t rs, ptrdiff_t cs, \
      size_t b_rows, size_t b_cols, T* b, ptrdiff_t brs, ptrdiff_t bcs, llh_uplo uplo,   \
      llh_error* err);                                                                    \
  /* solve_triangular_system(a, b, uplo, diag_fn), triangular.rs:133-144 (crate-      \
     private; kernel :95-128) with diag_fn(i) = |diag[i * dstride]| when abs_diag, else  \
     diag[i * dstride]: the pivots come from a separate vector, as the QR solves pass    \
     |R_ii| (qr.rs:145-150, 172-177).  Reads only the strict `uplo` triangle of a (its   \
     diagonal is never read) and the n entries of diag; errors as the solve above. */    \
  llh_status llh_solve_triangular_system_##sfx(                                           \
      llh_context* ctx, size_t rows, size_t cols, const T* a, ptrdiff_t rs, ptrdiff_t cs, \
      const T* diag, ptrdiff_t dstride, int abs_diag, size_t b_rows, size_t b_cols, T* b, \
      ptrdiff_t brs, ptrdiff_t bcs, llh_uplo uplo, llh_error* err);                       \
  /* SolveCInplace::solvec_inplace, cholesky.rs:122 (impl :136-144).  A is left         \
     holding the dirty factor, also when the later WrongRows check fails. */             \
  llh_status llh_solvec_inplace_##sfx(llh_context* ctx, size_t rows, size_t cols, T* a,    \
                                      ptrdiff_t rs, ptrdiff_t cs, size_t b_rows,          \
                                      size_t b_cols, T* b, ptrdiff_t brs, ptrdiff_t bcs,  \
                                      llh_error* err);                                    \
  /* InverseCInplace::invc_inplace, cholesky.rs:172 (impl :178-182): out (rows x rows,   \
     strides ors/ocs) receives A^-1; A is left holding the dirty factor. */              \
  llh_status llh_invc_inplace_##sfx(llh_context* ctx, size_t rows, size_t cols, T* a,      \
                                    ptrdiff_t rs, ptrdiff_t cs, T* out, ptrdiff_t ors,    \
                                    ptrdiff_t ocs, llh_error* err);

LLH_PLAIN_DECL(double, d)
LLH_PLAIN_DECL(float, s)
#undef LLH_PLAIN_DECL

/* ---------------------------------------------------------------------------
 * Device-resident forms (row-major, ld >= n; pointers are device memory).
 * ------------------------------------------------------------------------- */
#define LLH_DEV_DECL(T, sfx)                                                              \
  /* Blocked right-looking POTRF of the lower triangle, in place (the device-memory     \
     counterpart of cholesky_inplace_dirty / cholesky_inplace, cholesky.rs:51-82).      \
     zero_upper != 0 also zeroes the strict upper triangle. */                          \
  llh_status llh_dev_potrf_##sfx(llh_context* ctx, size_t n, T* a, size_t lda,            \
                                 int zero_upper, llh_error* err);                          \
  /* Solves op(A) X = B in place for m right-hand sides (triangular.rs:95-144).         \
     uplo names the triangle of op(A); trans != 0 means op(A) = A^T, so the solve       \
     reads the opposite triangle of A's storage (how solvec solves with L.t(),          \
     cholesky.rs:142).  B is n x m row-major with ldb >= m. */                          \
  llh_status llh_dev_solve_triangular_##sfx(llh_context* ctx, llh_uplo uplo, int trans,   \
                                            size_t n, const T* a, size_t lda, size_t m,    \
                                            T* b, size_t ldb, llh_error* err);             \
  /* solvec_inplace on device memory (cholesky.rs:136-144): A := dirty factor, B := X. */ \
  llh_status llh_dev_solvec_##sfx(llh_context* ctx, size_t n, T* a, size_t lda, size_t m,  \
                                  T* b, size_t ldb, llh_error* err);                      \
  /* invc_inplace on device memory (cholesky.rs:166-183): A := dirty factor, out (n x n,  \
     ldo >= n) := A^-1, both triangles.  Above the exact-order kernel's bound the inverse   \
     is L^-T L^-1 formed on MFMA (about 1.5 n^3 flops, against the 2 n^3 + n^3/3 of the     \
     reference's solve with the identity); out is untouched after NotPositiveDefinite. */  \
  llh_status llh_dev_invc_##sfx(llh_context* ctx, size_t n, T* a, size_t lda, T* out,      \
                                size_t ldo, llh_error* err);                               \
  /* LOBPCG's orthonormalize(v), lobpcg/algorithm.rs:75-97 (private to the LOBPCG        \
     module, the Cholesky path's own consumer): G = V^T V, L = cholesky_into(G), then     \
     V := (L^-1 V^T)^T.  v: n x k row-major on the device (ldv >= k), overwritten with    \
     U; r: k x k (ldr >= k) receives L with a zero upper triangle.  NotPositiveDefinite   \
     (column in err) when V^T V is not positive definite; v is then unchanged. */          \
  llh_status llh_dev_orthonormalize_##sfx(llh_context* ctx, size_t n, size_t k, T* v,      \
                                          size_t ldv, T* r, size_t ldr, llh_error* err);   \
  /* LOBPCG's apply_constraints(v, cholesky_yy, y), lobpcg/algorithm.rs:60-73:            \
     V := V - Y (L_yy^-1 (Y^T V)), with y n x ky (ldy >= ky) and l_yy the ky x ky lower    \
     Cholesky factor of Y^T Y (only its lower triangle is read). */                        \
  llh_status llh_dev_apply_constraints_##sfx(llh_context* ctx, size_t n, size_t k, T* v,   \
                                             size_t ldv, size_t ky, const T* y, size_t ldy, \
                                             const T* l_yy, size_t ldl, llh_error* err);

LLH_DEV_DECL(double, d)
LLH_DEV_DECL(float, s)
#undef LLH_DEV_DECL

/* Allocation-free llh_dev_solvec_d.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * semantics and results as llh_dev_solvec_d; the scratch that call takes from
 * the context's grow-only buffers (the inverses of the 64 x 64 diagonal blocks,
 * the right-hand sides of the forward substitution fused into the factor, and
 * the one-launch solve's ticket, error word and flags) lives instead in a
 * static device array of LLH_NOALLOC_SCRATCH_BYTES that every GPU the library
 * runs on holds until the process ends (llh_context_create locates it, once per
 * device).  The call makes no device allocation.  n > LLH_NOALLOC_MAX_N is
 * rejected with LLH_ERR_INVALID_ARGUMENT.  Every call on one GPU shares that
 * GPU's array and the array keeps the last call's contents: calls that overlap
 * on one GPU (from any context or thread) race and must be serialized by the
 * caller. */
#define LLH_NOALLOC_SCRATCH_BYTES (16u * 1024u * 1024u + 16u * 1024u)
#define LLH_NOALLOC_MAX_N 16384
llh_status llh_dev_solvec_noalloc_d(llh_context* ctx, size_t n, double* a, size_t lda, size_t m, double* b,
                                    size_t ldb, llh_error* err);

/* ---------------------------------------------------------------------------
 * Seeded synthetic workloads (no counterpart in the reference; for benchmarks
 * and full-size parity tests).  Values come from counters, so a CPU restatement
 * reproduces every one bit for bit: element (i, j) of a rows x cols matrix drawn
 * with (seed, stream) is value(key, i * cols + j), key = SplitMix64 output 0 from
 * the state seed + stream * 0x9E3779B97F4A7C15, draw c = SplitMix64 output c
 * from the state key.  LLH_GEN_UNIFORM: 2 u - 1 in [-1, 1), u = top 53 bits /
 * 2^53 of draw idx.  LLH_GEN_NORMAL: Marsaglia's polar method on draws
 * 2c, 2c + 1 (c = 64 idx + attempt), v1 sqrt(-2 ln(s) / s), ln(s) from a fixed
 * atanh series of correctly rounded operations.
 * ------------------------------------------------------------------------- */
enum { LLH_GEN_NORMAL = 0, LLH_GEN_UNIFORM = 1 };
/* x (rows x cols, row-major ldx, device f64) := the (seed, stream) draw. */
llh_status llh_dev_gen_matrix_d(llh_context* ctx, int dist, uint64_t seed, uint64_t stream, size_t rows,
                                size_t cols, double* x, size_t ldx, llh_error* err);
/* a (n x n, lda) := B B^T / n + lam I with B = the (seed, stream 0) draw, written
 * to b_work (n x n, contiguous): a_ij = (sum_k b_ik b_jk in k order, each
 * product and sum rounded) / n, plus lam on the diagonal; both triangles. */
llh_status llh_dev_gen_spd_d(llh_context* ctx, int dist, uint64_t seed, size_t n, double lam, double* a,
                             size_t lda, double* b_work, llh_error* err);

/* ---------------------------------------------------------------------------
 * Mixed precision with iterative refinement (no counterpart in the reference;
 * for callers that want an f64-accurate SPD solve at FP32-MFMA speed).
 *
 * The factor holds L in the precision `prec`: the trailing updates run in that
 * precision on MFMA, the 64 x 64 diagonal blocks are factored and inverted in
 * f64, and solves refine x in f64 against the caller's f64 A:
 *   x = solve_low(b); repeat { r = b - A x (f64); stop when every column has
 *   ||r||_inf <= ||x||_inf * ||A||_inf * eps64 * sqrt(n); x += solve_low(r) }.
 * (LLH_PREC_F16 applies the stopping test to the scaled system it factors.)
 * A is read from its lower triangle only and must stay valid and unchanged
 * while the factor is used.  A factor is bound to its context's stream.
 * ------------------------------------------------------------------------- */
typedef enum llh_precision {
  LLH_PREC_F64 = 0,          /* plain f64 factor, no refinement */
  LLH_PREC_F32 = 1,          /* f32 trailing updates on MFMA, f64 diagonal blocks, f64 refinement */
  LLH_PREC_F32_REFINED = 2,  /* batch plans only: f64 buffers, f32 factor and solves, f64 refinement */
  LLH_PREC_F16 = 3           /* factor and solve only: as LLH_PREC_F32, but of diag(s) A diag(s) with
                                s = diag(A)^-1/2, and each trailing SYRK multiplies f16 panels with f32
                                accumulation.  Refinement contracts the error by about cond(A) * 2^-11
                                per step, so it suits cond(A) up to ~1e2..1e3; a step that fails to
                                halve the residual ends the refinement as not converged. */
} llh_precision;

enum { LLH_FACTOR_FALLBACK_F64 = 1 }; /* on low-precision breakdown (pivot <= 0) or stalled refinement, redo in f64 */

typedef struct llh_refine_opts {
  int32_t max_iter; /* corrections before giving up (default 30 when <= 0) */
  int32_t flags;    /* LLH_FACTOR_FALLBACK_F64 */
} llh_refine_opts;

typedef struct llh_refine_info {
  int32_t iterations;     /* corrections applied */
  int32_t converged;      /* 1 when the stopping test passed */
  int32_t fell_back;      /* 1 when the f64 factor produced the result */
  int32_t precision_used; /* llh_precision of the factor that produced x */
  double residual;        /* max over columns of ||b - A x||_inf / (||A||_inf ||x||_inf); for
                             LLH_PREC_F16 of the scaled system S A S (x / s) = S b, S = diag(s) */
  double anorm;           /* ||A||_inf (LLH_PREC_F16: ||S A S||_inf) */
} llh_refine_info;

typedef struct llh_factor llh_factor;

/* Factors the device f64 matrix a (n x n, row-major lda, lower triangle read).
 * With LLH_PREC_F32 or LLH_PREC_F16 and LLH_FACTOR_FALLBACK_F64 in flags, a pivot
 * failure of the low-precision factor falls back to an f64 factor; otherwise it
 * reports NotPositiveDefinite with the failing row. */
llh_status llh_factor_create(llh_context* ctx, llh_precision prec, size_t n, const double* a, size_t lda,
                             int32_t flags, llh_factor** out, llh_error* err);
/* x (n x m, ldx) = A^-1 b (n x m, ldb), device f64; b and x may not alias. */
llh_status llh_factor_solve(llh_context* ctx, llh_factor* f, size_t m, const double* b, size_t ldb, double* x,
                            size_t ldx, const llh_refine_opts* opts, llh_refine_info* info, llh_error* err);
llh_precision llh_factor_precision(const llh_factor* f);
size_t llh_factor_device_bytes(const llh_factor* f);
void llh_factor_destroy(llh_factor* f);

/* One-shot form of the two calls above, with the factor's storage taken from the
 * context's reusable buffers (no allocation once they have grown). */
llh_status llh_dev_solve_mixed(llh_context* ctx, llh_precision prec, size_t n, const double* a, size_t lda,
                               size_t m, const double* b, size_t ldb, double* x, size_t ldx,
                               const llh_refine_opts* opts, llh_refine_info* info, llh_error* err);

/* ---------------------------------------------------------------------------
 * Batched solves of many independent SPD systems of uneven sizes (no
 * counterpart in the reference; for callers with thousands of small systems,
 * e.g. one per data group).  Each system i is solvec_inplace on its own pair
 * (A_i, B_i) (cholesky.rs:136-144): A_i is left holding its dirty factor and B_i
 * its solution.
 *
 * Layout: one device buffer holds every A_i, another every B_i.  A_i starts at
 * element a_off[i] and is row-major n_i x n_i with leading dimension lda[i]
 * (n_i rounded up to a multiple of 4); B_i starts at b_off[i], row-major
 * n_i x m_i with leading dimension m_i.  status[i] (device int32) receives -1
 * (the factor passed) or the first row whose pivot failed `d <= 0`.
 *
 * A plan fixes the sizes, buckets them (small systems: one workgroup each in
 * LDS, in the reference's exact operation order, from a persistent work queue;
 * larger ones: a blocked factor, fused per system in one persistent launch or
 * run as grouped launches over every system still active, one per step), and
 * captures the whole sequence in one hipGraph that
 * every llh_batch_solve replays.  llh_batch_autotune times the candidate
 * bucket bounds and launch shapes on real data and keeps the fastest.
 * ------------------------------------------------------------------------- */
typedef struct llh_batch_plan llh_batch_plan;

typedef struct llh_batch_config {
  int32_t small_limit;   /* systems with n <= small_limit take the LDS path (<= 128 f64, <= 176 f32) */
  int32_t small_waves;   /* resident workgroups per CU of the persistent small-system kernels */
  int32_t ir_iterations; /* LLH_PREC_F32_REFINED: refinement steps (0 = 2) */
  int32_t medium_mode;    /* the larger systems, when every one has n <= 512: 0 fused -- one workgroup per
                             system through its whole factor and substitutions, from a work queue; 1 the
                             factor fused so, then one substitution launch; 2 grouped launches, one per
                             blocked step over every system (always so when some n > 512) */
} llh_batch_config;

/* prec: LLH_PREC_F64 (buffers of double), LLH_PREC_F32 (buffers of float), or
 * LLH_PREC_F32_REFINED (buffers of double; every A_i is factored in f32 from a
 * copy, B_i is solved in f32 and then refined ir_iterations times with an f64
 * residual from the lower triangle of the f64 A_i, which is left unchanged; the
 * whole sequence is still one graph launch).  A system whose f32 factor fails
 * reports that row in status and is not refined. */
llh_status llh_batch_plan_create(llh_context* ctx, llh_precision prec, size_t count, const int32_t* n,
                                 const int32_t* m, const llh_batch_config* config, llh_batch_plan** out,
                                 llh_error* err);
/* Element counts of the packed A and B buffers, and the per-system layout. */
size_t llh_batch_a_elems(const llh_batch_plan* plan);
size_t llh_batch_b_elems(const llh_batch_plan* plan);
llh_status llh_batch_layout(const llh_batch_plan* plan, uint64_t* a_off, int32_t* lda, uint64_t* b_off,
                            llh_error* err);
llh_batch_config llh_batch_get_config(const llh_batch_plan* plan);
/* One graph launch: factors every A_i and solves every B_i in place (with
 * LLH_PREC_F32_REFINED, A_i is left as it was). */
llh_status llh_batch_solve(llh_context* ctx, llh_batch_plan* plan, void* a, void* b, int32_t* status,
                           llh_error* err);
/* Times the candidate configurations on copies of (a, b) (restored after each
 * trial; a and b are left unchanged) and keeps the fastest; *best_ms receives
 * its time. */
llh_status llh_batch_autotune(llh_context* ctx, llh_batch_plan* plan, const void* a, const void* b,
                              int32_t* status, double* best_ms, llh_error* err);
void llh_batch_plan_destroy(llh_batch_plan* plan);

/* ---------------------------------------------------------------------------
 * Solve server: continuous batching of SPD solve requests (no counterpart in
 * the reference; for callers that serve solve requests as they arrive).
 *
 * A worker thread admits submitted requests in arrival order while the HBM
 * budget and the in-flight limit allow, groups the small new systems admitted
 * in one step into one batched launch (the exact-order LDS kernel), runs every
 * other request on one of several HIP streams so their work overlaps, and keeps
 * each factor resident under its matrix id: a later request with the same id
 * reuses it and only runs the triangular solves.  Requests arriving while
 * others run are admitted as slots and memory free up.
 *
 * Request buffers are the caller's and must stay valid until the request is
 * done: a (n x n, f64, lower triangle read) in host memory, or already on the
 * device when a_on_device != 0; b (n x m) and x (n x m) in host memory,
 * row-major.  The server reads `a` only while that request runs: an
 * LLH_PREC_F32 entry keeps its own f64 copy of A for the refinement of later
 * requests that reuse it.  LLH_PREC_F32 requests factor in f32 (f64 diagonal
 * blocks) and apply `ir_iterations` f64 refinement steps; f64 requests solve
 * directly.
 * ------------------------------------------------------------------------- */
typedef struct llh_server llh_server;

typedef struct llh_server_config {
  uint64_t hbm_budget_bytes; /* device bytes the server may hold (0: 80% of free memory) */
  int32_t streams;           /* concurrent HIP streams (default 4) */
  int32_t max_inflight;      /* admitted, unfinished requests (default 64) */
  int32_t small_limit;       /* new systems with n <= small_limit are batched per step (default 128) */
  int32_t ir_iterations;     /* refinement steps of f32 requests (default 3) */
} llh_server_config;

typedef struct llh_request {
  uint64_t matrix_id;   /* factor cache key; reuse needs the same n */
  int32_t n;
  int32_t m;            /* right-hand sides, 1..64 */
  int32_t precision;    /* llh_precision */
  int32_t a_on_device;  /* a is a device pointer */
  const double* a;      /* may be null when the factor is resident */
  const double* b;
  double* x;
} llh_request;

typedef struct llh_result {
  int32_t status;         /* llh_status of the request */
  int32_t cache_hit;      /* 1: the resident factor was reused */
  int32_t batched;        /* 1: ran in a step's batched launch */
  int32_t precision_used;
  int64_t failed_row;     /* NotPositiveDefinite row; with LLH_ERR_HIP the hipError_t code; else -1 */
  double submit_s, admit_s, done_s; /* steady-clock seconds */
  double residual;        /* max_j ||b - A x||_inf / (||A||_inf ||x||_inf) for f32 requests, else 0 */
} llh_result;

typedef struct llh_server_stats {
  uint64_t submitted, completed, batched, cache_hits, rejected, steps;
  uint64_t bytes_in_use, bytes_peak, resident_factors;
} llh_server_stats;

llh_status llh_server_create(llh_context* ctx, const llh_server_config* config, llh_server** out,
                             llh_error* err);
/* Non-blocking.  Rejects (LLH_ERR_REJECTED) a request that cannot fit the budget
 * even alone, or one with bad sizes. */
llh_status llh_server_submit(llh_server* srv, const llh_request* req, uint64_t* ticket, llh_error* err);
/* 1 when the request is done (its result is copied to *out and forgotten), else 0. */
int llh_server_poll(llh_server* srv, uint64_t ticket, llh_result* out);
llh_status llh_server_wait(llh_server* srv, uint64_t ticket, llh_result* out, llh_error* err);
void llh_server_get_stats(llh_server* srv, llh_server_stats* out);
/* Test hooks (no counterpart in the reference; for tests of the error paths):
 * failures forced from the host, so every kernel still exits normally.
 *   LLH_SERVER_INJECT_STAGING: the next `value` batched steps fail to stage their
 *     descriptors; every request of such a step finishes with LLH_ERR_HIP.
 *   LLH_SERVER_INJECT_EVENT: the value-th event the server takes from now on
 *     fails to be created (1: the next one); the request that needed it
 *     finishes with LLH_ERR_HIP and the server goes on.
 *   LLH_SERVER_INJECT_SOLVE_SPIN: the flag-wait bound of the server's one-launch
 *     solves, as in llh_context_set_solve_spin_limit (0: the default); a request
 *     whose sweep gives up finishes with LLH_ERR_HIP. */
enum {
  LLH_SERVER_INJECT_STAGING = 1,
  LLH_SERVER_INJECT_EVENT = 2,
  LLH_SERVER_INJECT_SOLVE_SPIN = 3
};
llh_status llh_server_inject(llh_server* srv, int kind, int64_t value, llh_error* err);
/* Drops resident factors that no request is using. */
void llh_server_evict_all(llh_server* srv);
/* Finishes in-flight work, then frees everything. */
void llh_server_destroy(llh_server* srv);

#ifdef __cplusplus
}
#endif
#endif /* LLH_H */

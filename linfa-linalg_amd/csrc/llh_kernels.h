// Kernel launchers shared by the host drivers.  Every launcher enqueues on the
// given stream and returns the launch status; `gate`, when non-null, points at
// the factor's status word and makes the kernel a no-op once a pivot has failed.
#pragma once

#include "llh_common.h"

namespace llh {

// One matrix of a batched small factor/solve: A row-major n x n (lda), B
// row-major n x m (ldb), info receives kNoFail or the first failing row.
template <typename T>
struct SmallDesc {
  T* a;
  T* b;
  int64_t lda;
  int64_t ldb;
  int32_t n;
  int32_t m;
  int32_t* info;
};

// Batched solves (k_batch.hip, k_small.hip).  Matrix i of a batch: A at
// base.a + a_off (row-major n x n, ld = lda), B at base.b + b_off (row-major
// n x m, ld = m), inverses of its 64 x 64 diagonal blocks at base.linv + linv_off,
// status word base.status[id].  The base pointers live in device memory so a
// captured graph can be replayed on other buffers.
template <typename T>
struct BatchBase {
  T* a;
  T* b;
  T* linv;
  int32_t* status;
  int32_t* counters;  // work-queue counters of the persistent kernels
  T* y;               // B's layout: the forward sweep, when the grouped factor's steps carry it; else null
};
// Buffers of a batched FP32-factor / f64-refinement solve: the caller's f64 A
// and B (B becomes X), the plan's f32 copy of A (then its factor), the f32
// right-hand side / residual / correction, and the f64 copy of B.
struct MixedBase {
  const double* a64;
  double* b64;
  float* a32;
  float* x32;
  double* bk64;
  int32_t* status;
};
struct BatchItem {
  int64_t a_off;
  int64_t b_off;
  int64_t linv_off;
  int32_t n;
  int32_t m;
  int32_t lda;
  int32_t id;
};
// Work of one grouped launch: matrix item and block start k0; tiles add the
// tile origin (m0, n0) inside the block's panel or trailing matrix.
struct StepWork {
  int32_t item;
  int32_t k0;
};
struct TileWork {
  int32_t item;
  int32_t k0;
  int32_t m0;
  int32_t n0;
  int32_t kb;  // trailing tiles: width of the factored panel [k0, k0 + kb) that updates them
};

template <typename T>
constexpr int small_max() { return sizeof(T) == 8 ? kSmallMaxD : kSmallMaxS; }
template <typename T>
struct small_max_c { static constexpr int value = small_max<T>(); };

// k_small.hip: exact-order LDS Cholesky (bit-identical to the reference order).
// TS is the storage type, T the arithmetic type (T = TS except for the f64
// diagonal blocks of the mixed-precision factor).
template <typename TS, typename T = TS>
hipError_t launch_potrf_exact_single(TS* a, int64_t lda, int n, int32_t* info, int64_t info_base,
                                     const int32_t* gate, hipStream_t s);
template <typename T>
hipError_t launch_potrf_potrs_exact_batched(const SmallDesc<T>* descs, int count, int nmax, int do_solve,
                                            hipStream_t s);

// k_small.hip: persistent exact-order factor + solve over items[0..count),
// taken from a work-queue counter (base.counters[counter]); nmax <= small_max.
template <typename T>
hipError_t launch_small_batch(const BatchBase<T>* base, const BatchItem* items, int count, int nmax, int grid,
                              int counter, hipStream_t s,
                              int mode = 1);
// k_small.hip: grouped diagonal-block factor of step work (one workgroup each).
template <typename T>
hipError_t launch_grp_diag(const BatchBase<T>* base, const BatchItem* items, const StepWork* work, int count,
                           hipStream_t s);
// k_batch.hip: grouped steps of the blocked batch factor and solves.
template <typename T>
hipError_t launch_grp_trtri(const BatchBase<T>* base, const BatchItem* items, const StepWork* work, int count,
                            hipStream_t s);
template <typename T>
hipError_t launch_grp_panel(const BatchBase<T>* base, const BatchItem* items, const TileWork* work, int count,
                            hipStream_t s);
template <typename T>
hipError_t launch_grp_syrk(const BatchBase<T>* base, const BatchItem* items, const TileWork* work, int count,
                           hipStream_t s);
template <typename T>
hipError_t launch_grp_fwd(const BatchBase<T>* base, const BatchItem* items, const StepWork* work, int count,
                          hipStream_t s);
template <typename T>
hipError_t launch_grp_bwd(const BatchBase<T>* base, const BatchItem* items, const StepWork* work, int count,
                          hipStream_t s);
hipError_t launch_fill_status(int32_t* status, int64_t count, hipStream_t s);
// k_batch.hip: forward and back substitution of items[0..count), one workgroup
// per system, right-hand sides in LDS (kSolveCols columns at a time); every
// n <= kSolveLdsMaxN.
constexpr int kSolveLdsMaxN = 512;
constexpr int kSolveCols = 8;
// Panel width of the batched blocked factor (grouped launches and the fused
// per-system kernel alike): within a panel, after its j-th 64-column block the
// last lowbit(j + 1) blocks update the columns of as many next blocks; at the
// panel's end one update of that rank covers the rest (recursive halving, as
// the single-matrix factor's panels).  Measured on the 20000-system batch:
// 7.98 / 7.47 / 7.46 / 7.56 ms with panels of 64 / 128 / 256 / 512 columns.
constexpr int kBatchPanel = 256;
__host__ __device__ constexpr int lowbit(int x) { return x & -x; }
// The same substitution for one system: L in the lower triangle of a (lda),
// inverses of its 64 x 64 diagonal blocks in linv (64 x 64 each, in order), B
// row-major n x m (ldb), n <= kSolveLdsMaxN; a no-op when *gate (if given)
// records a failed factor.
template <typename T>
hipError_t launch_solve_lds(int n, int m, const T* a, int64_t lda, const T* linv, T* b, int64_t ldb,
                            const int32_t* gate, hipStream_t s);
template <typename T>
hipError_t launch_grp_solve(const BatchBase<T>* base, const BatchItem* items, int count, hipStream_t s);
// The medium bucket fused per system (k_batch.hip): factor (and with solve, the
// substitutions) of `count` items, one workgroup per system from the work
// queue bb.counters[counter] (zero before the launch), every n <= kSolveLdsMaxN.
template <typename T>
hipError_t launch_med(const BatchBase<T>* base, const BatchItem* items, int count, int counter, int grid,
                      bool solve, hipStream_t s);
// Workgroups of that kernel resident per CU (from the occupancy calculator).
template <typename T>
int med_resident_per_cu(bool solve);

// k_diag.hip: fast factor of one b x b (b <= 64) diagonal block in place, and
// its inverse into inv (64 x 64 row-major); a failure at block row f sets
// *info = min(*info, info_base + f).
template <typename TS, typename T = TS>
hipError_t launch_diag_fast(TS* a, int64_t lda, int b, TS* inv, int32_t* info, int64_t info_base,
                            const int32_t* gate, hipStream_t s);
// k_diag.hip: grouped form over step work (factor + inverse of each block).
template <typename T>
hipError_t launch_grp_diag_fast(const BatchBase<T>* base, const BatchItem* items, const StepWork* work,
                                int count, hipStream_t s);

// k_gemm.hip: C = beta*C + alpha*op(A)*op(B), optionally lower-triangle only.
template <typename T>
hipError_t launch_gemm(bool ta, bool tb, bool tri, int M, int N, int K, T alpha, const T* A, int64_t lda,
                       const T* B, int64_t ldb, T beta, T* C, int64_t ldc, const int32_t* gate,
                       hipStream_t s);

// k_batch.hip: the steps of the batched mixed-precision solve.  Before the
// factor, over the packed buffers: op 0 a32 = a64, op 1 x32 = b64 and
// bk64 = b64.  Per (system, 64-row block), systems whose factor failed left
// alone: the f64 residual r = bk - A x from the lower triangle of each A into
// x32 (as f32), and the update b64 = x32 (add false, after the first solve) or
// b64 += x32 (add true, after a correction).
hipError_t launch_mixed_flat(const MixedBase* mb, int op, int64_t elems, hipStream_t s);
hipError_t launch_grp_residual(const MixedBase* mb, const BatchItem* items, const StepWork* work, int count,
                               hipStream_t s);
hipError_t launch_grp_update(const MixedBase* mb, const BatchItem* items, const StepWork* work, int count, bool add,
                             hipStream_t s);

// k_gram.hip: C (ka x kb, ldc) = A^T B for tall row-major A (n x ka) and B
// (n x kb), split over row chunks; partial holds gram_partial_elems(...) elements.
int64_t gram_partial_elems(int n, int ka, int kb, int* chunks, int* rows_per_chunk);
template <typename T>
hipError_t launch_gram(int n, int ka, int kb, const T* A, int64_t lda, const T* B, int64_t ldb, T* C, int64_t ldc,
                       T* partial, hipStream_t s);

// k_gemm.hip: C = beta*C + alpha*A*A^T on the lower triangle of the n x n C
// (A is n x k, k-contiguous rows): the trailing update of the blocked POTRF.
template <typename T>
hipError_t launch_syrk_lower(int n, int k, T alpha, const T* A, int64_t lda, T beta, T* C, int64_t ldc,
                             const int32_t* gate, hipStream_t s);

// The same update with f16 operands and f32 accumulation and C: the trailing
// SYRK of the half-precision mixed factor.  variant selects the tile (0:
// 128 x 128, 1: 256 x 128).
hipError_t launch_syrk_lower_h(int n, int k, float alpha, const _Float16* A, int64_t lda, float beta, float* C,
                               int64_t ldc, const int32_t* gate, hipStream_t s, int variant = 0);
// C (M x N, f32, lower part of the top N x N square only) = beta C + alpha A B^T
// with f16 operands: the look-ahead panel's update of the same factor.
hipError_t launch_gemm_nt_tri_h(int M, int N, int K, float alpha, const _Float16* A, int64_t lda,
                                const _Float16* B, int64_t ldb, float beta, float* C, int64_t ldc,
                                const int32_t* gate, hipStream_t s);

// k_trsm.hip: panel solve X := X * L^-T (X is m x b, L is b x b lower, b <= kNB).
template <typename T>
hipError_t launch_trsm_rows(int m, int b, T* X, int64_t ldx, const T* L, int64_t ldl, const int32_t* gate,
                            hipStream_t s);
// k_trsm.hip: diagonal-block solve op(T) Y = Y for b <= kNB rows and m columns
// of a row-major Y; upper_op selects back substitution; trans reads T^T.
template <typename T>
hipError_t launch_trsv_block(bool upper_op, bool trans, int b, const T* Tm, int64_t ldt, int m, T* Y,
                             int64_t ldy, const int32_t* gate, hipStream_t s);

// k_trsm.hip: inverses of the 64 x 64 diagonal blocks of op(A) (n x n), written
// as consecutive row-major 64 x 64 blocks at out; upper names op(A)'s triangle.
template <typename TS, typename T = TS>
hipError_t launch_trtri_blocks(int n, const TS* A, int64_t lda, bool trans, bool upper, TS* out,
                               const int32_t* gate, hipStream_t s);
// k_trsm.hip: one-launch blocked solve op(A) X = B (B row-major n x m) using the
// diagonal-block inverses Dinv (transposed when dinv_t).  ticket/flags/err are
// device words (flags: one per 64-row block and 16-column chunk), epoch a value
// the flags have never held.  *err becomes non-zero if a wait exceeded spin_limit
// (a negative spin_limit: at the first wait).
template <typename T>
hipError_t launch_trsm_persistent(int n, int m, const T* A, int64_t lda, bool trans, bool lower_op,
                                  const T* Dinv, bool dinv_t, T* B, int64_t ldb, int32_t* ticket,
                                  int32_t* flags, int32_t epoch, int32_t* err, int64_t spin_limit,
                                  const int32_t* gate, hipStream_t s);

// k_mixed.hip: conversions, symmetric residual and norms of the refinement.
// Y = diag(rs) X diag(cs) in f32 (null scales are 1).
hipError_t launch_convert_d2s(int rows, int cols, const double* X, int64_t ldx, float* Y, int64_t ldy,
                              bool lower_only, hipStream_t s, const double* rs = nullptr,
                              const double* cs = nullptr);
// X += diag(rs) D (null rs is 1).
hipError_t launch_axpy_s2d(int rows, int cols, const float* D, int64_t ldd, double* X, int64_t ldx,
                           hipStream_t s, const double* rs = nullptr);
// Y (f16) = X (f32), round to nearest even.
hipError_t launch_convert_s2h(int rows, int cols, const float* X, int64_t ldx, _Float16* Y, int64_t ldy,
                              hipStream_t s);
// s[i] = 1 / sqrt(A[i][i]) for a positive finite diagonal entry, else 1.
hipError_t launch_diag_rsqrt(int n, const double* A, int64_t lda, double* s_out, hipStream_t s);
// R = B - A X, A symmetric from its lower triangle.  scratch (optional, in
// doubles): room for the split partial sums; symm_scratch_elems(n, m) elements
// give the most parallel split, less (or none) gives fewer splits.
hipError_t launch_symm_residual(int n, int m, const double* A, int64_t lda, const double* X, int64_t ldx,
                                const double* B, int64_t ldb, double* R, int64_t ldr, hipStream_t s,
                                double* scratch = nullptr, int64_t scratch_elems = 0);
int64_t symm_scratch_elems(int n, int m);
// out = |A| w (w null: 1), A symmetric from its lower triangle.
hipError_t launch_symm_abs_rowsum(int n, const double* A, int64_t lda, double* out, hipStream_t s,
                                  const double* w = nullptr, double* scratch = nullptr,
                                  int64_t scratch_elems = 0);
// out[c] = max_r |X[r][c]| * rs[r] (inv: / rs[r]; rs null: 1).
hipError_t launch_colmax(int rows, int cols, const double* X, int64_t ldx, double* out, hipStream_t s,
                         const double* rs = nullptr, bool inv = false);

// k_util.hip
hipError_t launch_fill_i32(int32_t* p, int32_t v, int64_t count, hipStream_t s);
template <typename T>
hipError_t launch_zero_triangle(T* a, int64_t lda, int n, bool zero_upper, const int32_t* gate,
                                hipStream_t s);
template <typename T>
hipError_t launch_is_triangular(const T* a, int64_t lda, int n, bool upper, int32_t* flag, hipStream_t s);
template <typename T>
hipError_t launch_set_identity(T* a, int64_t lda, int n, hipStream_t s);
// a's strict upper triangle := the transpose of its strict lower triangle.
template <typename T>
hipError_t launch_mirror_lower(T* a, int64_t lda, int n, hipStream_t s);
// dst (m x m) := the lower triangle of src with zeros above it.
template <typename T>
hipError_t launch_copy_lower(int m, const T* src, int64_t lds, T* dst, int64_t ldd, hipStream_t s);
// dst (rows x m, ldd) := src (lds), skipped whole once gate holds a failed row.
template <typename T>
hipError_t launch_copy_rows_gated(int rows, int m, const T* src, int64_t lds, T* dst, int64_t ldd,
                                  const int32_t* gate, hipStream_t s);

}  // namespace llh

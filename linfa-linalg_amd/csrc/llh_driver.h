// Host drivers (driver.hip): launch sequences of the factor and the solves.
#pragma once

#include "llh_common.h"
#include "llh_profile.h"

namespace llh {

// Device words of the one-launch triangular solve: ticket counter, error word
// and one flag per (64-row block, 16-column chunk).  *epoch is the owner's
// counter (a context's or a server stream's): the next value the flags have
// never held; each launch takes it and advances it at once, so no later launch
// can reuse an epoch whatever path the caller takes afterwards.  The error word
// is cleared once per call (solve_sync_begin), not per sweep, so a sweep that
// gave up stays visible after the sweeps that follow it.
struct SolveSync {
  int32_t* words = nullptr;  // [0] ticket, [1] error, [2..] flags
  int64_t nflags = 0;
  int32_t* epoch = nullptr;
  int64_t spin_limit = kSpinLimit;  // bound on each flag wait (negative: give up at once)
};

// Enqueues the clearing of the solve error word: once per call, before its
// first sweep.
hipError_t solve_sync_begin(const SolveSync& sync, hipStream_t s);

// Number of elements of the diagonal-block inverse buffer for order n.
inline int64_t dinv_elems(int n) { return (int64_t)((n + kNB - 1) / kNB) * kNB * kNB; }

// Second stream and a fork/join event pair for the factor's lookahead: while
// the trailing SYRK of outer panel k runs on the caller's stream, the next
// panel's columns are updated and factored on `side`.  All default priority.
struct Lookahead {
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  // third stream and its event: the forward substitution fused into the factor
  hipStream_t fwd = nullptr;
  hipEvent_t fwd_ev = nullptr;
};

// Forward substitution fused into the blocked factor (potrf_enqueue with la):
// y (n x m, ld ldy) first receives a copy of b (ld ldb; with b null, y already
// holds the right-hand sides, written in stream order on s); then, as each outer
// panel [k0, k0 + kb) is factored, rows [k0, k0 + kb) of y are solved with the
// panel's diagonal triangle (one-launch triangle over kb rows, with the
// diagonal-block inverses) and the rows below take -L21 Y_k (one GEMM).  This
// runs on la->fwd, ordered after each panel's factor by events, beside the
// trailing updates; when the factor is complete on s, y = L^-1 B, meaningful
// only while the status word holds no failure.  Needs m <= 64 and sync.
template <typename T>
struct FwdFuse {
  const T* b = nullptr;
  int64_t ldb = 0;
  T* y = nullptr;
  int64_t ldy = 0;
  int m = 0;
  SolveSync* sync = nullptr;
};

// Enqueues the in-place lower Cholesky of a (row-major, lda) on stream s.
// *info receives kNoFail or the first row whose pivot fails `d <= 0`.  When
// dinv is non-null it receives inv(L_ii) of every 64 x 64 diagonal block (the
// blocked path produces them on the way; they serve the panel solves and the
// later triangular solves).
// diag_f64 (f32 storage only): the diagonal blocks are factored and inverted in
// f64 arithmetic, the mixed-precision form.
// la (optional): lookahead streams; the whole factor is still complete in
// stream order on s.
// half (f32 storage only, optional): scratch of n * outer_width(n, true) f16 values; each
// trailing SYRK then rounds its panel to f16 there and multiplies on the f16
// MFMA with f32 accumulation (the half-precision mixed factor).
// ff (optional, with la and dinv, blocked sizes only): the fused forward
// substitution described at FwdFuse.
template <typename T>
hipError_t potrf_enqueue(T* a, int64_t lda, int n, int32_t* info, bool zero_upper, T* dinv, hipStream_t s,
                         Profiler* prof = nullptr, bool diag_f64 = false, const Lookahead* la = nullptr,
                         _Float16* half = nullptr, const FwdFuse<T>* ff = nullptr);

// Enqueues op(A) X = B in place (B row-major n x m); upper_op names the triangle
// of op(A) and trans != 0 means op(A) = A^T.  dinv, when non-null, holds the
// inverses of op(A)'s 64 x 64 diagonal blocks (transposed when dinv_t);
// otherwise they are computed into dinv_work.  With sync and m <= 64 the solve
// is one persistent launch; otherwise it is a sequence of blocked launches.
// gate (optional): the factor's status word; once it holds a failed row every
// kernel of the solve is a no-op and B keeps its values.
template <typename T>
hipError_t trsm_enqueue(bool upper_op, bool trans, int n, const T* a, int64_t lda, int m, T* b, int64_t ldb,
                        const T* dinv, bool dinv_t, T* dinv_work, SolveSync* sync, hipStream_t s,
                        Profiler* prof = nullptr, const int32_t* gate = nullptr);

// x := A^-1 b with the factor L of A (lower triangle of a, ld lda) and dinv,
// the inverses of L's 64 x 64 diagonal blocks: forward with L, back with L^T,
// B row-major n x m (ldb) in place.  Small systems take one single-workgroup
// launch, larger ones the two one-launch triangles of trsm_enqueue (sync).
template <typename T>
hipError_t solve_pair_enqueue(int n, const T* a, int64_t lda, int m, T* b, int64_t ldb, const T* dinv,
                              SolveSync* sync, hipStream_t s, Profiler* prof = nullptr,
                              const int32_t* gate = nullptr);

// True when factor_solve_enqueue fuses the forward substitution into the
// factor: two outer panels or more, above the single-workgroup solve's bound,
// 1 to 64 right-hand sides, and the lookahead streams with their third stream.
template <typename T>
bool forward_fusable(int n, int m, const Lookahead* la);

// The blocked factor of a (n > the exact kernel's bound, dinv required) and,
// for m > 0, x := A^-1 b (B row-major n x m, ldb, in place), on stream s, every
// solve gated on the status word info (after a failed factor B keeps its
// values).  When forward_fusable, the forward substitution runs beside the
// factor into y (n * m elements, FwdFuse), y is copied into B (gated) and only
// the back substitution follows the factor; otherwise y is unused and the two
// substitutions follow it.
template <typename T>
hipError_t factor_solve_enqueue(T* a, int64_t lda, int n, int m, T* b, int64_t ldb, T* dinv, T* y, SolveSync* sync,
                                int32_t* info, hipStream_t s, Profiler* prof, const Lookahead* la);

// Enqueues A^-1 = L^-T L^-1 from the factor L (lower triangle of a) and the
// inverses of its 64 x 64 diagonal blocks (dinv, as potrf_enqueue leaves them):
// W = L^-1 into w, then W^T W into x (both triangles).  tmp holds
// potri_tmp_elems(n) elements.
template <typename T>
hipError_t potri_enqueue(int n, const T* a, int64_t lda, const T* dinv, T* w, int64_t ldw, T* x, int64_t ldx,
                         T* tmp, hipStream_t s);
inline int64_t potri_tmp_elems(int n) { return (int64_t)(n / 2 + kNB) * (n / 2 + kNB); }

}  // namespace llh

// Host drivers: the blocked right-looking POTRF and the blocked triangular
// solves, as sequences of kernel launches on one stream.
#include <algorithm>
#include <climits>

#include "llh_driver.h"
#include "llh_kernels.h"

namespace llh {

#define LLH_TRY(expr)                  \
  do {                                 \
    hipError_t e_ = (expr);            \
    if (e_ != hipSuccess) return e_;   \
  } while (0)

// Brackets one launch on stream st with the profiler's events (no-op when
// profiling is off).
#define LLH_PROF_ON(st, cls, flops, bytes, expr)                                \
  do {                                                                          \
    const int pi_ = prof ? prof->begin((cls), (flops), (bytes), (st)) : -1;     \
    LLH_TRY(expr);                                                              \
    if (prof) prof->end(pi_, (st));                                             \
  } while (0)
#define LLH_PROF(cls, flops, bytes, expr) LLH_PROF_ON(s, cls, flops, bytes, expr)

namespace {

// Factors the outer panel of columns [k0, k0 + kb) on stream st, everything
// left of k0 having been applied to these columns.  The panel is split in two
// halves recursively down to kNB columns: factor the left half, update the
// right half's columns with it (one GEMM, rank = the left half's width, lower
// part of the top square), factor the right half.  A kNB-wide block is the
// diagonal block in LDS (with its inverse), then the rows below it as one MFMA
// product with that inverse.  Against a kNB-at-a-time right-looking sweep the
// launches are as many, but each column of the panel is read and written by
// log2(kb / kNB) updates instead of up to kb / kNB - 1, and the updates have rank
// 64 .. 256 instead of 64.
template <typename T>
hipError_t factor_panel(T* a, int64_t lda, int n, int k0, int kb, int32_t* info, T* dinv, hipStream_t st,
                        Profiler* prof, bool diag_f64) {
  const double es = (double)sizeof(T);
  if (kb > kNB) {
    const int h = (kb / kNB + 1) / 2 * kNB;  // left half: a whole number of blocks, at least one
    LLH_TRY(factor_panel<T>(a, lda, n, k0, h, info, dinv, st, prof, diag_f64));
    const int below = n - k0 - h, cols = kb - h;  // rows k0+h.., columns k0+h .. k0+kb
    T* left = a + (int64_t)(k0 + h) * lda + k0;    // L of rows k0+h.., columns k0 .. k0+h
    const double outs = (double)below * cols - (double)cols * (cols - 1) / 2.0, dh = h;
    LLH_PROF_ON(st, kPanelUpdate, 2.0 * outs * dh, es * (2.0 * outs + ((double)below + cols) * dh),
                (launch_gemm<T>(false, true, true, below, cols, h, T(-1), left, lda, left, lda, T(1), left + h, lda,
                                info, st)));
    return factor_panel<T>(a, lda, n, k0 + h, cols, info, dinv, st, prof, diag_f64);
  }
  const int k = k0, b = kb;
  T* akk = a + (int64_t)k * lda + k;
  const double db = b;
  T* dk = dinv + (int64_t)(k / kNB) * kNB * kNB;
  if (diag_f64 && sizeof(T) == 4) {
    LLH_PROF_ON(st, kDiagFactor, 2.0 * db * db * db / 3.0, es * db * db * 2.0,
                (launch_diag_fast<T, double>(akk, lda, b, dk, info, k, info, st)));
  } else {
    LLH_PROF_ON(st, kDiagFactor, 2.0 * db * db * db / 3.0, es * db * db * 2.0,
                (launch_diag_fast<T>(akk, lda, b, dk, info, k, info, st)));
  }
  const int below = n - k - b;
  if (below <= 0) return hipSuccess;
  T* panel = akk + (int64_t)b * lda;  // rows k+b.., columns k..k+b
  // L21 = A21 * L11^-T as one MFMA product with the inverse, in place (each
  // workgroup reads all of its rows' k before writing them).
  LLH_PROF_ON(st, kPanelTrsm, (double)below * db * db * 2.0, es * ((double)below * db * 2.0 + db * db),
              (launch_gemm<T>(false, true, false, below, b, b, T(1), panel, lda, dk, kNB, T(0), panel, lda, info,
                              st)));
  return hipSuccess;
}

// The trailing update C -= P P^T (lower triangle of the m x m C, P m x kb).
// With half (f32 storage only), P is first rounded to f16 into half (ld kb) and
// the SYRK multiplies in f16 with f32 accumulation.
template <typename T>
hipError_t trailing_syrk(int m, int kb, const T* p, int64_t lda, T* c, int32_t* info, _Float16* half,
                         hipStream_t s) {
  if constexpr (sizeof(T) == 4) {
    if (half != nullptr) {
      LLH_TRY(launch_convert_s2h(m, kb, p, lda, half, kb, s));
      return launch_syrk_lower_h(m, kb, -1.0f, half, kb, 1.0f, c, lda, info, s);
    }
  }
  return launch_syrk_lower<T>(m, kb, T(-1), p, lda, T(1), c, lda, info, s);
}

}  // namespace

// Right-looking blocked Cholesky of the lower triangle (the blocked equivalent
// of cholesky.rs:54-74).  Each outer panel (outer_width(n) columns) is factored
// by factor_panel; then one SYRK of that rank on MFMA updates the trailing matrix.
// With lookahead, the trailing update of panel k is split: the next panel's
// columns are updated and that panel factored on the side stream while the SYRK
// of the remaining trailing matrix runs on s.  Matrices no larger than the LDS
// kernel's limit are factored by that kernel alone, in the reference's exact
// operation order.
template <typename T>
hipError_t potrf_enqueue(T* a, int64_t lda, int n, int32_t* info, bool zero_upper, T* dinv, hipStream_t s,
                         Profiler* prof, bool diag_f64, const Lookahead* la, _Float16* half, const FwdFuse<T>* ff) {
  const double es = (double)sizeof(T);
  if (half != nullptr && sizeof(T) != 4) return hipErrorInvalidValue;
  if (ff != nullptr && (la == nullptr || la->fwd == nullptr || dinv == nullptr || n <= small_max<T>() ||
                        ff->m <= 0 || ff->m > 64 || ff->sync == nullptr))
    return hipErrorInvalidValue;
  // The fused forward step of the outer panel [k0, k0 + kb), on la->fwd once
  // the panel is factored: Y_k := L11^-1 Y_k (one-launch triangle), then
  // Y[k0 + kb:] -= L21 Y_k.
  auto fwd_panel = [&](int k0, int kb) -> hipError_t {
    const int m = ff->m;
    T* yk = ff->y + (int64_t)k0 * ff->ldy;
    LLH_TRY(trsm_enqueue<T>(false, false, kb, a + (int64_t)k0 * lda + k0, lda, m, yk, ff->ldy,
                            dinv + (int64_t)(k0 / kNB) * kNB * kNB, false, nullptr, ff->sync, la->fwd, prof, info));
    const int rest = n - k0 - kb;
    if (rest <= 0) return hipSuccess;
    LLH_PROF_ON(la->fwd, kSolveGemm, 2.0 * rest * m * kb, es * ((double)rest * kb + 2.0 * rest * m + (double)kb * m),
                (launch_gemm<T>(false, false, false, rest, m, kb, T(-1), a + (int64_t)(k0 + kb) * lda + k0, lda, yk,
                                ff->ldy, T(1), yk + (int64_t)kb * ff->ldy, ff->ldy, info, la->fwd)));
    return hipSuccess;
  };
  // bytes per panel element of the trailing update (the f16 copy is written
  // and read once more)
  const double pes = half != nullptr ? es + 4.0 : es;
  LLH_TRY(launch_fill_i32(info, kNoFail, 1, s));
  if (n <= 0) return hipSuccess;
  if (n <= small_max<T>() || dinv == nullptr) {
    if (n > small_max<T>()) return hipErrorInvalidValue;  // the blocked path needs the inverse buffer
    LLH_TRY(launch_potrf_exact_single<T>(a, lda, n, info, 0, nullptr, s));
    if (dinv != nullptr) LLH_TRY(launch_trtri_blocks<T>(n, a, lda, false, false, dinv, info, s));
  } else {
    const int nbo = outer_width(n, half != nullptr);
    LLH_TRY(factor_panel<T>(a, lda, n, 0, std::min(nbo, n), info, dinv, s, prof, diag_f64));
    if (ff != nullptr) {
      LLH_TRY(hipEventRecord(la->fwd_ev, s));
      LLH_TRY(hipStreamWaitEvent(la->fwd, la->fwd_ev, 0));
      if (ff->b != nullptr)
        LLH_TRY(hipMemcpy2DAsync(ff->y, ff->ldy * sizeof(T), ff->b, ff->ldb * sizeof(T), ff->m * sizeof(T), n,
                                 hipMemcpyDeviceToDevice, la->fwd));
      LLH_TRY(fwd_panel(0, std::min(nbo, n)));
    }
    for (int k0 = 0; k0 < n; k0 += nbo) {
      const int kb = std::min(nbo, n - k0);
      const int m2 = n - k0 - kb;
      if (m2 <= 0) break;
      const int kb2 = std::min(nbo, m2);  // width of the next panel
      T* p = a + (int64_t)(k0 + kb) * lda + k0;  // L21: rows k0+kb.., columns k0..k0+kb
      T* c = p + kb;                              // trailing matrix at (k0+kb, k0+kb)
      // The trailing SYRK's algorithmic work: outs outputs of the lower
      // triangle, each a kb-term dot product; bytes: read and write the lower
      // triangle of C once, read the panel once.
      if (la == nullptr) {
        const double outs = (double)m2 * (m2 + 1) / 2.0;
        LLH_PROF(kTrailingSyrk, 2.0 * outs * kb, es * 2.0 * outs + pes * (double)m2 * kb,
                 (trailing_syrk<T>(m2, kb, p, lda, c, info, half, s)));
        LLH_TRY(factor_panel<T>(a, lda, n, k0 + kb, kb2, info, dinv, s, prof, diag_f64));
        continue;
      }
      // With half, the whole panel (m2 x kb) is rounded to f16 once, before the
      // fork: the look-ahead update and the SYRK both read the copy.
      bool half_ready = false;
      if constexpr (sizeof(T) == 4) {
        if (half != nullptr) {
          LLH_TRY(launch_convert_s2h(m2, kb, p, lda, half, kb, s));
          half_ready = true;
        }
      }
      LLH_TRY(hipEventRecord(la->fork, s));
      LLH_TRY(hipStreamWaitEvent(la->side, la->fork, 0));
      {
        // next panel's columns (m2 rows x kb2 columns, lower part of the top
        // square), then its factor, on the side stream
        const double outs = (double)m2 * kb2 - (double)kb2 * (kb2 - 1) / 2.0;
        if (half_ready) {
          if constexpr (sizeof(T) == 4)
            LLH_PROF_ON(la->side, kPanelUpdate, 2.0 * outs * kb, 4.0 * 2.0 * outs + 2.0 * ((double)m2 + kb2) * kb,
                        (launch_gemm_nt_tri_h(m2, kb2, kb, -1.0f, half, kb, half, kb, 1.0f, c, lda, info,
                                              la->side)));
        } else {
          LLH_PROF_ON(la->side, kPanelUpdate, 2.0 * outs * kb, es * (2.0 * outs + ((double)m2 + kb2) * kb),
                      (launch_gemm<T>(false, true, true, m2, kb2, kb, T(-1), p, lda, p, lda, T(1), c, lda, info,
                                      la->side)));
        }
        LLH_TRY(factor_panel<T>(a, lda, n, k0 + kb, kb2, info, dinv, la->side, prof, diag_f64));
      }
      LLH_TRY(hipEventRecord(la->join, la->side));
      if (ff != nullptr) {
        LLH_TRY(hipStreamWaitEvent(la->fwd, la->join, 0));
        LLH_TRY(fwd_panel(k0 + kb, kb2));
      }
      const int m3 = m2 - kb2;  // the rest of the trailing matrix
      if (m3 > 0) {
        const double outs = (double)m3 * (m3 + 1) / 2.0;
        T* p3 = p + (int64_t)kb2 * lda;
        T* c3 = c + (int64_t)kb2 * lda + kb2;
        if (half_ready) {
          if constexpr (sizeof(T) == 4)
            LLH_PROF(kTrailingSyrk, 2.0 * outs * kb, es * 2.0 * outs + 2.0 * (double)m3 * kb,
                     (launch_syrk_lower_h(m3, kb, -1.0f, half + (int64_t)kb2 * kb, kb, 1.0f, c3, lda, info, s)));
        } else {
          LLH_PROF(kTrailingSyrk, 2.0 * outs * kb, es * 2.0 * outs + pes * (double)m3 * kb,
                   (trailing_syrk<T>(m3, kb, p3, lda, c3, info, half, s)));
        }
      }
      LLH_TRY(hipStreamWaitEvent(s, la->join, 0));
    }
    if (ff != nullptr) {
      LLH_TRY(hipEventRecord(la->fwd_ev, la->fwd));
      LLH_TRY(hipStreamWaitEvent(s, la->fwd_ev, 0));
    }
  }
  if (zero_upper) LLH_TRY(launch_zero_triangle<T>(a, lda, n, true, info, s));
  return hipSuccess;
}

// Blocked substitution for op(A) X = B (B row-major n x m, in place), kNB rows
// at a time: solve the diagonal block, then subtract its contribution from the
// rows still to come with one GEMM (the column sweep of triangular.rs:113-125,
// done a block of rows at a time).
template <typename T>
hipError_t trsm_enqueue(bool upper_op, bool trans, int n, const T* a, int64_t lda, int m, T* b, int64_t ldb,
                        const T* dinv, bool dinv_t, T* dinv_work, SolveSync* sync, hipStream_t s,
                        Profiler* prof, const int32_t* gate) {
  if (n <= 0 || m <= 0) return hipSuccess;
  const double es = (double)sizeof(T);
  const int nblk0 = (n + kNB - 1) / kNB;
  const int nq = (m + kSolveChunkHost - 1) / kSolveChunkHost;
  if (sync != nullptr && sync->epoch != nullptr && m <= 64 && sync->nflags >= (int64_t)nblk0 * nq &&
      (dinv != nullptr || dinv_work)) {
    if (dinv == nullptr) {
      LLH_PROF(kSolveDiag, (double)n * kNB * kNB / 3.0, es * (double)n * kNB,
               (launch_trtri_blocks<T>(n, a, lda, trans, upper_op, dinv_work, gate, s)));
      dinv = dinv_work;
      dinv_t = false;
    }
    // the launcher resets the ticket; the error word is the call's (solve_sync_begin)
    const int32_t epoch = (*sync->epoch)++;
    const double fl = (double)n * n * m;
    LLH_PROF(kSolveGemm, fl, es * ((double)n * n / 2.0 + 2.0 * n * m),
             (launch_trsm_persistent<T>(n, m, a, lda, trans, !upper_op, dinv, dinv_t, b, ldb, sync->words,
                                        sync->words + 2, epoch, sync->words + 1, sync->spin_limit, gate, s)));
    return hipSuccess;
  }
  const int nblk = (n + kNB - 1) / kNB;
  if (!upper_op) {  // forward
    for (int bi = 0; bi < nblk; ++bi) {
      const int i0 = bi * kNB, bs = std::min(kNB, n - i0);
      const T* tkk = a + (int64_t)i0 * lda + i0;
      LLH_PROF(kSolveDiag, (double)bs * bs * m, es * ((double)bs * bs / 2.0 + 2.0 * bs * m),
               (launch_trsv_block<T>(false, trans, bs, tkk, lda, m, b + (int64_t)i0 * ldb, ldb, gate, s)));
      const int rest = n - i0 - bs;
      if (rest <= 0) continue;
      // B[i0+bs:, :] -= op(A)[i0+bs:, i0:i0+bs] * X[i0:i0+bs, :]
      const T* blk = trans ? a + (int64_t)i0 * lda + (i0 + bs) : a + (int64_t)(i0 + bs) * lda + i0;
      LLH_PROF(kSolveGemm, 2.0 * rest * m * bs, es * ((double)rest * bs + 2.0 * rest * m + (double)bs * m),
               (launch_gemm<T>(trans, false, false, rest, m, bs, T(-1), blk, lda, b + (int64_t)i0 * ldb, ldb,
                               T(1), b + (int64_t)(i0 + bs) * ldb, ldb, gate, s)));
    }
  } else {  // backward
    for (int bi = nblk - 1; bi >= 0; --bi) {
      const int i0 = bi * kNB, bs = std::min(kNB, n - i0);
      const T* tkk = a + (int64_t)i0 * lda + i0;
      LLH_PROF(kSolveDiag, (double)bs * bs * m, es * ((double)bs * bs / 2.0 + 2.0 * bs * m),
               (launch_trsv_block<T>(true, trans, bs, tkk, lda, m, b + (int64_t)i0 * ldb, ldb, gate, s)));
      if (i0 == 0) continue;
      // B[:i0, :] -= op(A)[:i0, i0:i0+bs] * X[i0:i0+bs, :]
      const T* blk = trans ? a + (int64_t)i0 * lda : a + i0;
      LLH_PROF(kSolveGemm, 2.0 * i0 * m * bs, es * ((double)i0 * bs + 2.0 * i0 * m + (double)bs * m),
               (launch_gemm<T>(trans, false, false, i0, m, bs, T(-1), blk, lda, b + (int64_t)i0 * ldb, ldb,
                               T(1), b, ldb, gate, s)));
    }
  }
  return hipSuccess;
}

// The forward and back substitution with the factor L (lower triangle of a)
// and the inverses of its diagonal blocks (dinv, as the factor leaves them).
// Up to kSolvePairLdsMaxN rows and kSolveCols right-hand sides it is one
// single-workgroup launch with B in LDS and L streamed through registers
// (launch_solve_lds), otherwise the two one-launch triangles (trsm_enqueue),
// whose flag hops cost ~7 us per 64 rows.  For n = 512 and one right-hand side
// the one-workgroup launch takes 68 us against 2 x 57 us for the triangles.
constexpr int kSolvePairLdsMaxN = kSolveLdsMaxN;
template <typename T>
hipError_t solve_pair_enqueue(int n, const T* a, int64_t lda, int m, T* b, int64_t ldb, const T* dinv,
                              SolveSync* sync, hipStream_t s, Profiler* prof, const int32_t* gate) {
  const double es = (double)sizeof(T);
  if (n <= kSolvePairLdsMaxN && m <= kSolveCols && dinv != nullptr && lda <= INT32_MAX) {
    LLH_PROF(kSolveGemm, 2.0 * n * n * m, es * ((double)n * n + 2.0 * n * m),
             (launch_solve_lds<T>(n, m, a, lda, dinv, b, ldb, gate, s)));
    return hipSuccess;
  }
  LLH_TRY(trsm_enqueue<T>(false, false, n, a, lda, m, b, ldb, dinv, false, nullptr, sync, s, prof, gate));
  return trsm_enqueue<T>(true, true, n, a, lda, m, b, ldb, dinv, true, nullptr, sync, s, prof, gate);
}

template <typename T>
bool forward_fusable(int n, int m, const Lookahead* la) {
  return m > 0 && m <= 64 && n > outer_width(n, false) && n > kSolvePairLdsMaxN && n > small_max<T>() &&
         la != nullptr && la->fwd != nullptr && la->fwd_ev != nullptr;
}

// At n = 32768 with 16 right-hand sides the forward triangle after the factor
// is ~3.5 ms (512 flag hops); fused, each panel's 16-hop triangle and product
// run beside the trailing SYRK and only the last one trails the factor.
template <typename T>
hipError_t factor_solve_enqueue(T* a, int64_t lda, int n, int m, T* b, int64_t ldb, T* dinv, T* y, SolveSync* sync,
                                int32_t* info, hipStream_t s, Profiler* prof, const Lookahead* la) {
  if (!forward_fusable<T>(n, m, la)) {
    LLH_TRY(potrf_enqueue<T>(a, lda, n, info, false, dinv, s, prof, false, la));
    if (m <= 0) return hipSuccess;
    return solve_pair_enqueue<T>(n, a, lda, m, b, ldb, dinv, sync, s, prof, info);
  }
  if (y == nullptr) return hipErrorInvalidValue;
  FwdFuse<T> ff;
  ff.b = b;
  ff.ldb = ldb;
  ff.y = y;
  ff.ldy = m;
  ff.m = m;
  ff.sync = sync;
  LLH_TRY(potrf_enqueue<T>(a, lda, n, info, false, dinv, s, prof, false, la, nullptr, &ff));
  LLH_TRY(launch_copy_rows_gated<T>(n, m, y, m, b, ldb, info, s));
  return trsm_enqueue<T>(true, true, n, a, lda, m, b, ldb, dinv, true, nullptr, sync, s, prof, info);
}

namespace {

// Half of m rounded to whole kNB blocks (at least one block).
inline int split_blocks(int m) { return ((m / kNB + 1) / 2) * kNB; }

// W := L^-1 for the m x m lower-triangular block of L at (o, o) (ld lda), into
// W at (o, o) (ld ldw), by halves: [[L11, 0], [L21, L22]]^-1 =
// [[W11, 0], [-W22 L21 W11, W22]].  A kNB block copies its inverse from dinv
// (inv(L_kk) of every diagonal block, formed by the factor).  tmp holds the
// n2 x n1 product L21 W11.
template <typename T>
hipError_t trtri_rec(int o, int m, const T* L, int64_t lda, const T* dinv, T* W, int64_t ldw, T* tmp,
                     hipStream_t s) {
  if (m <= kNB)
    return launch_copy_lower<T>(m, dinv + (int64_t)(o / kNB) * kNB * kNB, kNB, W + (int64_t)o * ldw + o, ldw, s);
  const int n1 = split_blocks(m), n2 = m - n1;
  LLH_TRY(trtri_rec<T>(o, n1, L, lda, dinv, W, ldw, tmp, s));
  LLH_TRY(trtri_rec<T>(o + n1, n2, L, lda, dinv, W, ldw, tmp, s));
  const T* l21 = L + (int64_t)(o + n1) * lda + o;
  const T* w11 = W + (int64_t)o * ldw + o;
  const T* w22 = W + (int64_t)(o + n1) * ldw + o + n1;
  T* w21 = W + (int64_t)(o + n1) * ldw + o;
  LLH_TRY(launch_gemm<T>(false, false, false, n2, n1, n1, T(1), l21, lda, w11, ldw, T(0), tmp, n1, nullptr, s));
  return launch_gemm<T>(false, false, false, n2, n1, n2, T(-1), w22, ldw, tmp, n1, T(0), w21, ldw, nullptr, s);
}

// X := the lower triangle of W^T W for the m x m lower-triangular block of W at
// (o, o), by halves: with W = [[W11, 0], [W21, W22]],
// W^T W = [[W11^T W11 + W21^T W21, .], [W22^T W21, W22^T W22]].  (The blocks of
// X above the diagonal of each product are written too and then overwritten
// by the mirror.)
template <typename T>
hipError_t lauum_rec(int o, int m, const T* W, int64_t ldw, T* X, int64_t ldx, hipStream_t s) {
  const T* w = W + (int64_t)o * ldw + o;
  T* x = X + (int64_t)o * ldx + o;
  if (m <= kNB) return launch_gemm<T>(true, false, false, m, m, m, T(1), w, ldw, w, ldw, T(0), x, ldx, nullptr, s);
  const int n1 = split_blocks(m), n2 = m - n1;
  const T* w21 = W + (int64_t)(o + n1) * ldw + o;
  const T* w22 = W + (int64_t)(o + n1) * ldw + o + n1;
  T* x21 = X + (int64_t)(o + n1) * ldx + o;
  LLH_TRY(lauum_rec<T>(o, n1, W, ldw, X, ldx, s));
  LLH_TRY(launch_gemm<T>(true, false, false, n1, n1, n2, T(1), w21, ldw, w21, ldw, T(1), x, ldx, nullptr, s));
  LLH_TRY(launch_gemm<T>(true, false, false, n2, n1, n2, T(1), w22, ldw, w21, ldw, T(0), x21, ldx, nullptr, s));
  return lauum_rec<T>(o + n1, n2, W, ldw, X, ldx, s);
}

}  // namespace

// A^-1 = L^-T L^-1 from the factor L (the lower triangle of a, lda) and the
// inverses of its diagonal blocks: W = L^-1 into w (n x n, ldw; its strict
// upper triangle is zeroed), then the lower triangle of W^T W into x (ldx),
// mirrored into the upper triangle.  tmp: (n / 2 + kNB)^2 elements.  Work:
// about 2n^3/3 flops for W and n^3/2 for W^T W as MFMA products, against the
// 2n^3 of solving with the n columns of the identity.
template <typename T>
hipError_t potri_enqueue(int n, const T* a, int64_t lda, const T* dinv, T* w, int64_t ldw, T* x, int64_t ldx,
                         T* tmp, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  LLH_TRY(launch_zero_triangle<T>(w, ldw, n, true, nullptr, s));
  LLH_TRY(trtri_rec<T>(0, n, a, lda, dinv, w, ldw, tmp, s));
  LLH_TRY(lauum_rec<T>(0, n, w, ldw, x, ldx, s));
  return launch_mirror_lower<T>(x, ldx, n, s);
}

template hipError_t potri_enqueue<double>(int, const double*, int64_t, const double*, double*, int64_t, double*,
                                          int64_t, double*, hipStream_t);
template hipError_t potri_enqueue<float>(int, const float*, int64_t, const float*, float*, int64_t, float*, int64_t,
                                         float*, hipStream_t);

template hipError_t potrf_enqueue<double>(double*, int64_t, int, int32_t*, bool, double*, hipStream_t,
                                          Profiler*, bool, const Lookahead*, _Float16*, const FwdFuse<double>*);
template hipError_t potrf_enqueue<float>(float*, int64_t, int, int32_t*, bool, float*, hipStream_t, Profiler*,
                                         bool, const Lookahead*, _Float16*, const FwdFuse<float>*);
template bool forward_fusable<double>(int, int, const Lookahead*);
template bool forward_fusable<float>(int, int, const Lookahead*);
template hipError_t factor_solve_enqueue<double>(double*, int64_t, int, int, double*, int64_t, double*, double*,
                                                 SolveSync*, int32_t*, hipStream_t, Profiler*, const Lookahead*);
template hipError_t factor_solve_enqueue<float>(float*, int64_t, int, int, float*, int64_t, float*, float*,
                                                SolveSync*, int32_t*, hipStream_t, Profiler*, const Lookahead*);
template hipError_t solve_pair_enqueue<double>(int, const double*, int64_t, int, double*, int64_t, const double*,
                                               SolveSync*, hipStream_t, Profiler*, const int32_t*);
template hipError_t solve_pair_enqueue<float>(int, const float*, int64_t, int, float*, int64_t, const float*,
                                              SolveSync*, hipStream_t, Profiler*, const int32_t*);
template hipError_t trsm_enqueue<double>(bool, bool, int, const double*, int64_t, int, double*, int64_t,
                                         const double*, bool, double*, SolveSync*, hipStream_t, Profiler*,
                                         const int32_t*);
template hipError_t trsm_enqueue<float>(bool, bool, int, const float*, int64_t, int, float*, int64_t,
                                        const float*, bool, float*, SolveSync*, hipStream_t, Profiler*,
                                        const int32_t*);

hipError_t solve_sync_begin(const SolveSync& sync, hipStream_t s) {
  if (sync.words == nullptr) return hipSuccess;
  return launch_fill_i32(sync.words + 1, 0, 1, s);
}

}  // namespace llh

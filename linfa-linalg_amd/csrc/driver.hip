// Host drivers: the blocked right-looking POTRF and the blocked triangular
// solves, as sequences of kernel launches on one stream.
#include <algorithm>

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
                         Profiler* prof, bool diag_f64, const Lookahead* la, _Float16* half) {
  const double es = (double)sizeof(T);
  if (half != nullptr && sizeof(T) != 4) return hipErrorInvalidValue;
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

template hipError_t potrf_enqueue<double>(double*, int64_t, int, int32_t*, bool, double*, hipStream_t,
                                          Profiler*, bool, const Lookahead*, _Float16*);
template hipError_t potrf_enqueue<float>(float*, int64_t, int, int32_t*, bool, float*, hipStream_t, Profiler*,
                                         bool, const Lookahead*, _Float16*);
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

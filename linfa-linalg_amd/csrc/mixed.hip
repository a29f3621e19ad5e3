// Mixed-precision SPD solve with iterative refinement, and the resident factor
// object (include/llh.h: llh_factor_*, llh_dev_solve_mixed).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "llh.h"
#include "llh_context.h"
#include "llh_driver.h"
#include "llh_kernels.h"

using namespace llh;

struct llh_factor {
  static constexpr uint32_t kMagic = 0x4c4c4846;  // "LLHF"
  uint32_t magic = kMagic;
  llh_precision prec = LLH_PREC_F64;  // precision of the factor held
  int n = 0;
  const double* a = nullptr;  // caller's f64 matrix (residuals)
  int64_t lda = 0;
  bool owns = false;          // storage below is owned (else borrowed from the context)
  void* l = nullptr;          // n x n factor, row-major, ld n
  size_t l_bytes = 0;
  void* dinv = nullptr;       // inverses of the 64 x 64 diagonal blocks (f64-sized)
  size_t dinv_bytes = 0;
  void* work = nullptr;       // residual / correction / norms of the solves
  size_t work_bytes = 0;
  void* scale = nullptr;      // f16 factor: s = diag(A)^-1/2, L factors diag(s) A diag(s)
  size_t scale_bytes = 0;
  double anrm = 0.0;          // ||A||_inf
};

namespace {

#define CHECK_HIP(expr, where)                             \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return hip_fail(err, e_, where); \
  } while (0)

bool bad_ctx(llh_context* ctx) { return ctx == nullptr || ctx->magic != llh_context::kMagic; }

size_t elem_size(llh_precision p) { return (p == LLH_PREC_F32 || p == LLH_PREC_F16) ? 4 : 8; }
bool low_precision(llh_precision p) { return p == LLH_PREC_F32 || p == LLH_PREC_F16; }

// Grows an owned or context-borrowed device buffer.
hipError_t grow_buffer(llh_context* ctx, bool owns, llh_context::Buffer* borrowed, void** ptr, size_t* bytes,
                       size_t want) {
  if (*bytes >= want) return hipSuccess;
  hipError_t e = ctx_quiesce(ctx);
  if (e != hipSuccess) return e;
  if (owns) {
    if (*ptr) (void)hipFree(*ptr);
    *ptr = nullptr;
    *bytes = 0;
    e = hipMalloc(ptr, want);
    if (e == hipSuccess) *bytes = want;
    return e;
  }
  if (borrowed->bytes < want) {
    if (borrowed->ptr) (void)hipFree(borrowed->ptr);
    borrowed->ptr = nullptr;
    borrowed->bytes = 0;
    e = hipMalloc(&borrowed->ptr, want);
    if (e != hipSuccess) return e;
    borrowed->bytes = want;
  }
  *ptr = borrowed->ptr;
  *bytes = borrowed->bytes;
  return hipSuccess;
}

// The first right-hand sides of a one-shot solve, for the forward substitution
// fused into the low-precision factor: b (n x m f64, ldb) is rounded (and, for
// the f16 factor, scaled) into y (n x m f32, ld m), and y = L^-1 y is formed
// beside the factor.  done tells whether the factor took it (a fallback to f64
// or a size the fusion does not cover leaves y unused).
struct FirstRhs {
  int m = 0;
  const double* b = nullptr;
  size_t ldb = 0;
  float* y = nullptr;
  bool done = false;
};

// Factors f->a into f->l with precision f->prec; on a low-precision pivot
// failure with fallback, refactors in f64.  The f16 form factors
// diag(s) A diag(s) with s = diag(A)^-1/2 (unit diagonal, every entry of L at
// most 1 in magnitude, so nothing overflows f16), keeping the factor in f32
// and rounding each trailing SYRK's panel to f16.
llh_status factor_core(llh_context* ctx, llh_factor* f, int32_t flags, llh_error* err, FirstRhs* first = nullptr) {
  const int n = f->n;
  hipStream_t s = ctx->stream;
  // work: [rowsum n | norm 2] [partial sums of the split row-sum walk]
  const int64_t nsc = symm_scratch_elems(n, 1);
  CHECK_HIP(grow_buffer(ctx, f->owns, &ctx->mixed_work, &f->work, &f->work_bytes, (size_t)(n + 2 + nsc) * 8),
            "work buffer");
  double* rowsum = static_cast<double*>(f->work);
  for (int attempt = 0; attempt < 2; ++attempt) {
    const size_t es = elem_size(f->prec);
    CHECK_HIP(grow_buffer(ctx, f->owns, &ctx->mixed_l, &f->l, &f->l_bytes, (size_t)n * n * es), "factor buffer");
    CHECK_HIP(grow_buffer(ctx, f->owns, &ctx->dinv, &f->dinv, &f->dinv_bytes, (size_t)dinv_elems(n) * 8),
              "inverse buffer");
    const double* sc = nullptr;
    if (f->prec == LLH_PREC_F16) {
      CHECK_HIP(grow_buffer(ctx, f->owns, &ctx->mixed_scale, &f->scale, &f->scale_bytes, (size_t)n * 8),
                "scale buffer");
      sc = static_cast<const double*>(f->scale);
      CHECK_HIP(launch_diag_rsqrt(n, f->a, f->lda, static_cast<double*>(f->scale), s), "scale");
    }
    // ||A||_inf (f16: ||diag(s) A diag(s)||_inf, the system it refines) from the
    // lower triangle, in f64
    CHECK_HIP(launch_symm_abs_rowsum(n, f->a, f->lda, rowsum, s, sc, rowsum + n + 2, nsc), "row sums");
    CHECK_HIP(launch_colmax(n, 1, rowsum, 1, rowsum + n, s, sc), "norm");
    CHECK_HIP(hipMemcpyAsync(&f->anrm, rowsum + n, sizeof(double), hipMemcpyDeviceToHost, s), "norm readback");
    if (low_precision(f->prec)) {
      float* l = static_cast<float*>(f->l);
      _Float16* half = nullptr;
      if (f->prec == LLH_PREC_F16) {
        void* hp = nullptr;
        size_t hb = 0;
        const size_t want = (size_t)n * outer_width(n, true) * sizeof(_Float16);
        CHECK_HIP(grow_buffer(ctx, false, &ctx->half_panel, &hp, &hb, want), "f16 panel buffer");
        half = static_cast<_Float16*>(hp);
      }
      CHECK_HIP(launch_convert_d2s(n, n, f->a, f->lda, l, n, true, s, sc, sc), "convert");
      const Lookahead* la = ctx_lookahead(ctx);
      FwdFuse<float> ff;
      SolveSync sync;
      const bool fuse = first != nullptr && first->y != nullptr && forward_fusable<float>(n, first->m, la);
      if (fuse) {
        CHECK_HIP(ctx_sync(ctx, n, first->m, &sync), "solve sync");
        CHECK_HIP(launch_convert_d2s(n, first->m, first->b, (int64_t)first->ldb, first->y, first->m, false, s, sc),
                  "convert rhs");
        ff.y = first->y;
        ff.ldy = first->m;
        ff.m = first->m;
        ff.sync = &sync;
      }
      CHECK_HIP(potrf_enqueue<float>(l, n, n, ctx->d_info, false, static_cast<float*>(f->dinv), s,
                                     ctx->profiler(), true, la, half, fuse ? &ff : nullptr),
                "potrf f32");
      if (fuse) {
        CHECK_HIP(ctx_read_solve_error(ctx), "solve status");
        first->done = true;
      }
    } else {
      double* l = static_cast<double*>(f->l);
      CHECK_HIP(hipMemcpy2DAsync(l, (size_t)n * 8, f->a, (size_t)f->lda * 8, (size_t)n * 8, n,
                                 hipMemcpyDeviceToDevice, s),
                "copy");
      CHECK_HIP(potrf_enqueue<double>(l, n, n, ctx->d_info, false, static_cast<double*>(f->dinv), s,
                                      ctx->profiler(), false, ctx_lookahead(ctx)),
                "potrf f64");
    }
    CHECK_HIP(hipMemcpyAsync(ctx->h_info, ctx->d_info, sizeof(int32_t), hipMemcpyDeviceToHost, s), "status");
    CHECK_HIP(hipStreamSynchronize(s), "synchronize");
    const int32_t fail = ctx->h_info[0];
    if (first != nullptr && first->done && (fail != kNoFail || ctx->h_info[2] != 0))
      first->done = false;  // y is not L^-1 b: the solve starts over
    if (fail == kNoFail) return ok(err);
    if (low_precision(f->prec) && (flags & LLH_FACTOR_FALLBACK_F64)) {
      f->prec = LLH_PREC_F64;
      continue;
    }
    return not_pd(err, fail);
  }
  return not_pd(err, ctx->h_info[0]);
}

llh_status solve_stalled_mixed(llh_error* err) {
  return set_error(err, LLH_ERR_HIP, 0, 0, -1, "triangular solve: a block wait exceeded its bound", 0);
}

template <typename T>
hipError_t solve_low(llh_context* ctx, const llh_factor* f, int m, T* d, SolveSync* sync) {
  const int n = f->n;
  T* l = static_cast<T*>(f->l);
  T* dinv = static_cast<T*>(f->dinv);
  return solve_pair_enqueue<T>(n, l, n, m, d, m, dinv, sync, ctx->stream, ctx->profiler());
}

// y_fwd (optional, low precision only): L^-1 of b's rounding, left by the
// factor (FirstRhs); the first solve then takes only the back substitution.
llh_status solve_core(llh_context* ctx, llh_factor* f, size_t m_, const double* b, size_t ldb, double* x,
                      size_t ldx, const llh_refine_opts* opts, llh_refine_info* info, llh_error* err,
                      float* y_fwd = nullptr) {
  const int n = f->n, m = (int)m_;
  hipStream_t s = ctx->stream;
  llh_refine_info inf{0, 0, 0, (int32_t)f->prec, 0.0, f->anrm};
  const int max_iter = (opts && opts->max_iter > 0) ? opts->max_iter : 30;
  // work: [rowsum n | norm 2] [r: n*m f64] [colmax: 2m f64] [d: n*m f32]
  // [partial sums of the split residual walk]
  const int64_t nsc = symm_scratch_elems(n, m);
  const size_t off_r = (size_t)(n + 2) * 8, off_cm = off_r + (size_t)n * m * 8,
               off_d = off_cm + (size_t)2 * m * 8, off_p = (off_d + (size_t)n * m * 4 + 255) / 256 * 256,
               total = off_p + (size_t)nsc * 8;
  CHECK_HIP(grow_buffer(ctx, f->owns, &ctx->mixed_work, &f->work, &f->work_bytes, total), "work buffer");
  char* w = static_cast<char*>(f->work);
  double* r = reinterpret_cast<double*>(w + off_r);
  double* cm = reinterpret_cast<double*>(w + off_cm);
  float* d = reinterpret_cast<float*>(w + off_d);
  double* part = reinterpret_cast<double*>(w + off_p);
  SolveSync sync;
  CHECK_HIP(ctx_sync(ctx, n, m, &sync), "solve sync");
  const double cte = f->anrm * std::ldexp(1.0, -53) * std::sqrt((double)n);
  double hn[128];
  auto residual_check = [&](bool* conv) -> llh_status {
    CHECK_HIP(launch_symm_residual(n, m, f->a, f->lda, x, (int64_t)ldx, b, (int64_t)ldb, r, m, s, part, nsc),
              "residual");
    // the f16 factor refines diag(s) A diag(s) y = diag(s) b, y = x / s: its
    // test measures that system's residual and solution
    const double* rs = f->prec == LLH_PREC_F16 ? static_cast<const double*>(f->scale) : nullptr;
    CHECK_HIP(launch_colmax(n, m, r, m, cm, s, rs, false), "residual norm");
    CHECK_HIP(launch_colmax(n, m, x, (int64_t)ldx, cm + m, s, rs, true), "solution norm");
    CHECK_HIP(hipMemcpyAsync(hn, cm, sizeof(double) * 2 * m, hipMemcpyDeviceToHost, s), "norm readback");
    CHECK_HIP(ctx_read_solve_error(ctx), "solve status");
    CHECK_HIP(hipStreamSynchronize(s), "synchronize");
    // a sweep that gave up left x unfinished: report it before any test of x
    if (ctx->h_info[2] != 0) return solve_stalled_mixed(err);
    bool c = true;
    double worst = 0.0;
    for (int j = 0; j < m; ++j) {
      const double rn = hn[j], xn = hn[m + j];
      if (!(rn <= xn * cte)) c = false;
      const double rel = (f->anrm > 0 && xn > 0) ? rn / (f->anrm * xn) : (rn == 0 ? 0.0 : INFINITY);
      worst = std::max(worst, std::isnan(rel) ? INFINITY : rel);
    }
    inf.residual = worst;
    *conv = c;
    return LLH_OK;
  };
  if (m > 64) return invalid(err, "mixed-precision solves take at most 64 right-hand sides per call");
  if (f->prec == LLH_PREC_F64) {
    CHECK_HIP(hipMemcpy2DAsync(x, ldx * 8, b, ldb * 8, (size_t)m * 8, n, hipMemcpyDeviceToDevice, s), "copy");
    // solve in place in x (row-major n x m with ld m required by the solve)
    if (ldx != (size_t)m) return invalid(err, "ldx must equal m");
    CHECK_HIP(solve_low<double>(ctx, f, m, x, &sync), "solve f64");
    bool conv = false;
    llh_status st = residual_check(&conv);
    if (st != LLH_OK) return st;
    inf.converged = 1;
  } else {
    if (ldx != (size_t)m) return invalid(err, "ldx must equal m");
    // The f16 factor is of S A S (S = diag(scale)): A^-1 v = S (S A S)^-1 S v.
    const double* sc = f->prec == LLH_PREC_F16 ? static_cast<const double*>(f->scale) : nullptr;
    // x = solve_low(b)
    if (y_fwd != nullptr) {
      CHECK_HIP(trsm_enqueue<float>(true, true, n, static_cast<float*>(f->l), n, m, y_fwd, m,
                                    static_cast<float*>(f->dinv), true, nullptr, &sync, s, ctx->profiler()),
                "solve f32");
    } else {
      CHECK_HIP(launch_convert_d2s(n, m, b, (int64_t)ldb, d, m, false, s, sc), "convert rhs");
      CHECK_HIP(solve_low<float>(ctx, f, m, d, &sync), "solve f32");
    }
    CHECK_HIP(launch_fill_i32(reinterpret_cast<int32_t*>(x), 0, 2 * (int64_t)n * m, s), "zero x");
    CHECK_HIP(launch_axpy_s2d(n, m, y_fwd != nullptr ? y_fwd : d, m, x, (int64_t)ldx, s, sc), "update x");
    double prev = INFINITY;
    for (int it = 0;; ++it) {
      bool conv = false;
      llh_status st = residual_check(&conv);
      if (st != LLH_OK) return st;
      if (conv) {
        inf.converged = 1;
        break;
      }
      if (it >= max_iter) break;
      // The f16 factor contracts the error by about cond(A) * 2^-11 per step:
      // a step that does not halve the residual means A is too ill-conditioned
      // for it, and the caller is better served by the fallback at once.
      if (f->prec == LLH_PREC_F16 && !(inf.residual <= 0.5 * prev)) break;
      prev = inf.residual;
      CHECK_HIP(launch_convert_d2s(n, m, r, m, d, m, false, s, sc), "convert residual");
      CHECK_HIP(solve_low<float>(ctx, f, m, d, &sync), "solve f32");
      CHECK_HIP(launch_axpy_s2d(n, m, d, m, x, (int64_t)ldx, s, sc), "update x");
      inf.iterations = it + 1;
    }
  }
  CHECK_HIP(ctx_read_solve_error(ctx), "solve status");
  CHECK_HIP(hipStreamSynchronize(s), "synchronize");
  if (ctx->h_info[2] != 0) return solve_stalled_mixed(err);
  if (info) *info = inf;
  if (!inf.converged) return set_error(err, LLH_ERR_NO_CONVERGENCE, 0, 0, -1, "refinement did not converge", 0);
  return ok(err);
}

llh_status check_factor_args(llh_context* ctx, size_t n, const double* a, size_t lda, llh_error* err) {
  if (bad_ctx(ctx)) return invalid(err, "null or destroyed context");
  if (n == 0) return invalid(err, "empty matrix");
  if (a == nullptr || lda < n || n > 0x7fffffff) return invalid(err, "bad matrix arguments");
  return LLH_OK;
}

}  // namespace

extern "C" {

llh_status llh_factor_create(llh_context* ctx, llh_precision prec, size_t n, const double* a, size_t lda,
                             int32_t flags, llh_factor** out, llh_error* err) {
  if (out == nullptr) return invalid(err, "null output pointer");
  *out = nullptr;
  llh_status st = check_factor_args(ctx, n, a, lda, err);
  if (st != LLH_OK) return st;
  if (prec != LLH_PREC_F64 && !low_precision(prec)) return invalid(err, "unknown precision");
  auto* f = new llh_factor();
  f->prec = prec;
  f->n = (int)n;
  f->a = a;
  f->lda = (int64_t)lda;
  f->owns = true;
  st = factor_core(ctx, f, flags, err);
  if (st != LLH_OK) {
    llh_factor_destroy(f);
    return st;
  }
  *out = f;
  return LLH_OK;
}

}  // extern "C"

namespace {

// llh_factor_solve, its first solve optionally started by the factor (y_fwd).
llh_status factor_solve_impl(llh_context* ctx, llh_factor* f, size_t m, const double* b, size_t ldb, double* x,
                             size_t ldx, const llh_refine_opts* opts, llh_refine_info* info, llh_error* err,
                             float* y_fwd) {
  if (bad_ctx(ctx)) return invalid(err, "null or destroyed context");
  if (f == nullptr || f->magic != llh_factor::kMagic) return invalid(err, "null or destroyed factor");
  if (m == 0) return ok(err);
  if (b == nullptr || x == nullptr || ldb < m || ldx < m) return invalid(err, "bad right-hand side arguments");
  llh_status st = solve_core(ctx, f, m, b, ldb, x, ldx, opts, info, err, y_fwd);
  if (st == LLH_ERR_NO_CONVERGENCE && opts && (opts->flags & LLH_FACTOR_FALLBACK_F64) &&
      low_precision(f->prec)) {
    f->prec = LLH_PREC_F64;
    st = factor_core(ctx, f, 0, err);
    if (st != LLH_OK) return st;
    st = solve_core(ctx, f, m, b, ldb, x, ldx, opts, info, err);
    if (info) info->fell_back = 1;
  }
  return st;
}

}  // namespace

extern "C" {

llh_status llh_factor_solve(llh_context* ctx, llh_factor* f, size_t m, const double* b, size_t ldb, double* x,
                            size_t ldx, const llh_refine_opts* opts, llh_refine_info* info, llh_error* err) {
  return factor_solve_impl(ctx, f, m, b, ldb, x, ldx, opts, info, err, nullptr);
}

llh_precision llh_factor_precision(const llh_factor* f) {
  return (f && f->magic == llh_factor::kMagic) ? f->prec : LLH_PREC_F64;
}

size_t llh_factor_device_bytes(const llh_factor* f) {
  if (!f || f->magic != llh_factor::kMagic) return 0;
  return f->l_bytes + f->dinv_bytes + f->work_bytes + f->scale_bytes;
}

void llh_factor_destroy(llh_factor* f) {
  if (f == nullptr || f->magic != llh_factor::kMagic) return;
  if (f->owns) {
    if (f->l) (void)hipFree(f->l);
    if (f->dinv) (void)hipFree(f->dinv);
    if (f->work) (void)hipFree(f->work);
    if (f->scale) (void)hipFree(f->scale);
  }
  f->magic = 0;
  delete f;
}

llh_status llh_dev_solve_mixed(llh_context* ctx, llh_precision prec, size_t n, const double* a, size_t lda,
                               size_t m, const double* b, size_t ldb, double* x, size_t ldx,
                               const llh_refine_opts* opts, llh_refine_info* info, llh_error* err) {
  llh_status st = check_factor_args(ctx, n, a, lda, err);
  if (st != LLH_OK) return st;
  if (prec != LLH_PREC_F64 && !low_precision(prec)) return invalid(err, "unknown precision");
  llh_factor f;
  f.prec = prec;
  f.n = (int)n;
  f.a = a;
  f.lda = (int64_t)lda;
  f.owns = false;
  const int32_t flags = opts ? opts->flags : 0;
  // The first forward substitution runs beside the low-precision factor where
  // the fusion covers the size (one-launch triangles, m <= 64).
  FirstRhs first;
  if (low_precision(prec) && m > 0 && m <= 64 && b != nullptr && x != nullptr && ldb >= m && ldx == m &&
      forward_fusable<float>((int)n, (int)m, ctx_lookahead(ctx))) {
    void* yp = nullptr;
    size_t yb = 0;
    if (grow_buffer(ctx, false, &ctx->fwd_y, &yp, &yb, n * m * sizeof(float)) == hipSuccess) {
      first.m = (int)m;
      first.b = b;
      first.ldb = ldb;
      first.y = static_cast<float*>(yp);
    }
  }
  st = factor_core(ctx, &f, flags, err, &first);
  const bool fell = low_precision(prec) && f.prec == LLH_PREC_F64;
  if (st != LLH_OK) return st;
  if (m > 0) {
    st = factor_solve_impl(ctx, &f, m, b, ldb, x, ldx, opts, info, err,
                           first.done && low_precision(f.prec) ? first.y : nullptr);
    if (info && fell) info->fell_back = 1;
  }
  f.magic = 0;  // borrowed storage stays with the context
  return st;
}

}  // extern "C"

// C ABI (include/llh.h): contexts, the plain host-memory forms that mirror
// linfa-linalg's traits, and the device-resident forms.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "llh.h"
#include "llh_context.h"
#include "llh_driver.h"
#include "llh_kernels.h"

using namespace llh;

namespace llh {

llh_status set_error(llh_error* err, llh_status code, uint64_t a, uint64_t b, int64_t column,
                     const char* msg, int hip_error) {
  if (err != nullptr) {
    err->code = code;
    err->hip_error = hip_error;
    err->a = a;
    err->b = b;
    err->column = column;
    std::snprintf(err->message, sizeof(err->message), "%s", msg ? msg : "");
  }
  return code;
}

llh_status ok(llh_error* err) { return set_error(err, LLH_OK, 0, 0, -1, "", 0); }

llh_status hip_fail(llh_error* err, hipError_t e, const char* where) {
  char buf[160];
  std::snprintf(buf, sizeof(buf), "%s: %s", where, hipGetErrorString(e));
  return set_error(err, e == hipErrorOutOfMemory ? LLH_ERR_OUT_OF_MEMORY : LLH_ERR_HIP, 0, 0, -1, buf, (int)e);
}

llh_status invalid(llh_error* err, const char* msg) {
  return set_error(err, LLH_ERR_INVALID_ARGUMENT, 0, 0, -1, msg, 0);
}

llh_status not_square(llh_error* err, size_t rows, size_t cols) {
  char buf[96];
  std::snprintf(buf, sizeof(buf), "Matrix of (%zu, %zu) is not square", rows, cols);
  return set_error(err, LLH_ERR_NOT_SQUARE, rows, cols, -1, buf, 0);
}

llh_status not_pd(llh_error* err, int64_t col) {
  return set_error(err, LLH_ERR_NOT_POSITIVE_DEFINITE, 0, 0, col, "Matrix is not positive definite", 0);
}

llh_status wrong_rows(llh_error* err, size_t expected, size_t actual) {
  char buf[96];
  std::snprintf(buf, sizeof(buf), "Matrix must have %zu rows, not %zu", expected, actual);
  return set_error(err, LLH_ERR_WRONG_ROWS, expected, actual, -1, buf, 0);
}

// Waits for the context stream and the factor's lookahead streams: before a
// buffer any of them may still read is freed (a call that failed while
// enqueueing may have left side-stream work unjoined).
hipError_t ctx_quiesce(llh_context* ctx) {
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess && ctx->la.side != nullptr) e = hipStreamSynchronize(ctx->la.side);
  if (e == hipSuccess && ctx->la.fwd != nullptr) e = hipStreamSynchronize(ctx->la.fwd);
  return e;
}

hipError_t ctx_buffer(llh_context* ctx, int slot, size_t bytes, void** out) {
  auto& b = ctx->scratch[slot];
  if (b.bytes < bytes) {
    if (b.ptr != nullptr) {
      hipError_t e = ctx_quiesce(ctx);
      if (e != hipSuccess) return e;
      (void)hipFree(b.ptr);
      b.ptr = nullptr;
      b.bytes = 0;
    }
    const size_t want = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&b.ptr, want);
    if (e != hipSuccess) return e;
    b.bytes = want;
  }
  *out = b.ptr;
  return hipSuccess;
}

static hipError_t grow(llh_context* ctx, llh_context::Buffer& b, size_t bytes, bool* fresh) {
  *fresh = false;
  if (b.bytes >= bytes) return hipSuccess;
  if (b.ptr != nullptr) {
    hipError_t e = ctx_quiesce(ctx);
    if (e != hipSuccess) return e;
    (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
  }
  const size_t want = std::max<size_t>(bytes, 256);
  hipError_t e = hipMalloc(&b.ptr, want);
  if (e != hipSuccess) return e;
  b.bytes = want;
  *fresh = true;
  return hipSuccess;
}

hipError_t ctx_dinv(llh_context* ctx, int n, size_t es, void** out) {
  bool fresh = false;
  hipError_t e = grow(ctx, ctx->dinv, (size_t)dinv_elems(n) * es, &fresh);
  *out = ctx->dinv.ptr;
  return e;
}

const Lookahead* ctx_lookahead(llh_context* ctx) {
  if (ctx->la.side == nullptr) {
    Lookahead la;
    if (hipStreamCreateWithFlags(&la.side, hipStreamNonBlocking) != hipSuccess) return nullptr;
    if (hipEventCreateWithFlags(&la.fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&la.join, hipEventDisableTiming) != hipSuccess) {
      if (la.fork) (void)hipEventDestroy(la.fork);
      (void)hipStreamDestroy(la.side);
      return nullptr;
    }
    // the fused forward substitution's stream is optional: without it the
    // solves run after the factor
    if (hipStreamCreateWithFlags(&la.fwd, hipStreamNonBlocking) != hipSuccess) {
      la.fwd = nullptr;
    } else if (hipEventCreateWithFlags(&la.fwd_ev, hipEventDisableTiming) != hipSuccess) {
      (void)hipStreamDestroy(la.fwd);
      la.fwd = nullptr;
      la.fwd_ev = nullptr;
    }
    ctx->la = la;
  }
  return &ctx->la;
}

hipError_t ctx_sync(llh_context* ctx, int n, int m, SolveSync* sync) {
  const int64_t nb = (n + kNB - 1) / kNB, nq = (m + kSolveChunkHost - 1) / kSolveChunkHost;
  const int64_t nflags = nb * std::max<int64_t>(nq, 1);
  bool fresh = false;
  hipError_t e = grow(ctx, ctx->sync, (size_t)(2 + nflags) * sizeof(int32_t), &fresh);
  if (e != hipSuccess) return e;
  if (fresh) {  // flags start at 0, below every epoch the context hands out
    e = launch_fill_i32(static_cast<int32_t*>(ctx->sync.ptr), 0, (int64_t)(ctx->sync.bytes / sizeof(int32_t)),
                        ctx->stream);
    if (e != hipSuccess) return e;
    ctx->epoch = 1;
  }
  sync->words = static_cast<int32_t*>(ctx->sync.ptr);
  sync->nflags = (int64_t)(ctx->sync.bytes / sizeof(int32_t)) - 2;
  sync->epoch = &ctx->epoch;
  sync->spin_limit = ctx->spin_limit;
  // one error word per call: cleared here, read back once the call's sweeps are done
  return solve_sync_begin(*sync, ctx->stream);
}

hipError_t ctx_read_solve_error(llh_context* ctx) {
  if (ctx->sync.ptr == nullptr) {
    ctx->h_info[2] = 0;
    return hipSuccess;
  }
  return hipMemcpyAsync(ctx->h_info + 2, static_cast<int32_t*>(ctx->sync.ptr) + 1, sizeof(int32_t),
                        hipMemcpyDeviceToHost, ctx->stream);
}

}  // namespace llh

namespace {

llh_status solve_stalled(llh_error* err) {
  return set_error(err, LLH_ERR_HIP, 0, 0, -1, "triangular solve: a block wait exceeded its bound", 0);
}

#define CHECK_HIP(expr, where)                           \
  do {                                                   \
    hipError_t e_ = (expr);                              \
    if (e_ != hipSuccess) return hip_fail(err, e_, where); \
  } while (0)

// Gathers a strided host view into a row-major buffer and back.
template <typename T>
void gather(const T* a, ptrdiff_t rs, ptrdiff_t cs, size_t rows, size_t cols, T* out) {
  for (size_t i = 0; i < rows; ++i) {
    const T* row = a + (ptrdiff_t)i * rs;
    T* o = out + i * cols;
    if (cs == 1) {
      std::memcpy(o, row, cols * sizeof(T));
    } else {
      for (size_t j = 0; j < cols; ++j) o[j] = row[(ptrdiff_t)j * cs];
    }
  }
}

// Scatters row i's columns [0, cend(i)) of a row-major buffer into a view.
template <typename T, typename F>
void scatter(const T* in, size_t rows, size_t cols, T* a, ptrdiff_t rs, ptrdiff_t cs, F cend) {
  for (size_t i = 0; i < rows; ++i) {
    T* row = a + (ptrdiff_t)i * rs;
    const T* src = in + i * cols;
    const size_t ce = cend(i);
    if (cs == 1) {
      std::memcpy(row, src, ce * sizeof(T));
    } else {
      for (size_t j = 0; j < ce; ++j) row[(ptrdiff_t)j * cs] = src[j];
    }
  }
}

bool bad_ctx(llh_context* ctx) { return ctx == nullptr || ctx->magic != llh_context::kMagic; }

// Host staging buffer of the plain forms for count elements: a page-locked,
// grow-only context buffer (slot 0 or 1) up to kPinnedStageMax bytes, since
// copies from page-locked memory go straight over the link where pageable ones
// are bounced through the runtime's own buffers (n = 512: the plain solvec's
// two 2 MB copies); the heap vector beyond that.
constexpr size_t kPinnedStageMax = 64u << 20;
template <typename T>
hipError_t ctx_stage(llh_context* ctx, int slot, size_t count, std::vector<T>& heap, T** out) {
  const size_t bytes = count * sizeof(T);
  llh_context::Buffer& b = ctx->host_stage[slot];
  if (bytes > kPinnedStageMax) {
    heap.resize(count);
    *out = heap.data();
    return hipSuccess;
  }
  if (b.bytes < bytes) {
    if (b.ptr != nullptr) {
      hipError_t e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) return e;
      (void)hipHostFree(b.ptr);
      b.ptr = nullptr;
      b.bytes = 0;
    }
    const size_t want = std::max<size_t>(bytes, 1u << 16);
    hipError_t e = hipHostMalloc(&b.ptr, want, hipHostMallocDefault);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      heap.resize(count);
      *out = heap.data();
      return hipSuccess;
    }
    b.bytes = want;
  }
  *out = static_cast<T*>(b.ptr);
  return hipSuccess;
}

// Factors the packed host copy `h` (n x n row-major) on the device and writes
// the result back into the user's view with the reference's semantics.
// Returns the first failing row or -1.
template <typename T>
llh_status factor_plain(llh_context* ctx, size_t n, T* a, ptrdiff_t rs, ptrdiff_t cs, bool clean,
                        T* h, T** d_out, int64_t* fail_out, llh_error* err) {
  void* dv = nullptr;
  CHECK_HIP(ctx_buffer(ctx, 0, n * n * sizeof(T), &dv), "device buffer");
  T* d = static_cast<T*>(dv);
  void* dinv = nullptr;
  CHECK_HIP(ctx_dinv(ctx, (int)n, sizeof(T), &dinv), "inverse buffer");
  CHECK_HIP(hipMemcpyAsync(d, h, n * n * sizeof(T), hipMemcpyHostToDevice, ctx->stream), "upload");
  CHECK_HIP(potrf_enqueue<T>(d, (int64_t)n, (int)n, ctx->d_info, clean, static_cast<T*>(dinv), ctx->stream, nullptr,
                            false, ctx_lookahead(ctx)),
            "potrf");
  CHECK_HIP(hipMemcpyAsync(h, d, n * n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream), "download");
  CHECK_HIP(hipMemcpyAsync(ctx->h_info, ctx->d_info, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream),
            "status");
  CHECK_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
  const int32_t f = ctx->h_info[0];
  if (f == kNoFail) {
    if (clean)
      scatter(h, n, n, a, rs, cs, [n](size_t) { return n; });
    else
      scatter(h, n, n, a, rs, cs, [](size_t i) { return i + 1; });
    *fail_out = -1;
  } else {
    // cholesky.rs:54-72 returns at row f: rows < f are factored, row f holds
    // L[f, ..f], the rest of the array is as it was.
    const size_t fr = (size_t)f;
    scatter(h, fr + 1, n, a, rs, cs, [fr](size_t i) { return i < fr ? i + 1 : fr; });
    *fail_out = f;
  }
  if (d_out) *d_out = d;
  return LLH_OK;
}

template <typename T>
llh_status cholesky_plain(llh_context* ctx, size_t rows, size_t cols, T* a, ptrdiff_t rs, ptrdiff_t cs,
                          bool clean, llh_error* err) {
  if (bad_ctx(ctx)) return invalid(err, "null or destroyed context");
  if (rows != cols) return not_square(err, rows, cols);
  const size_t n = rows;
  if (n == 0) return ok(err);
  if (a == nullptr) return invalid(err, "null matrix pointer");
  if (n > 0x7fffffff) return invalid(err, "matrix too large");
  std::vector<T> hv;
  T* h = nullptr;
  CHECK_HIP(ctx_stage<T>(ctx, 0, n * n, hv, &h), "host staging");
  gather(a, rs, cs, n, n, h);
  int64_t fail = -1;
  llh_status st = factor_plain<T>(ctx, n, a, rs, cs, clean, h, nullptr, &fail, err);
  if (st != LLH_OK) return st;
  if (fail >= 0) return not_pd(err, fail);
  return ok(err);
}

template <typename T>
llh_status triangular_plain(llh_context* ctx, size_t rows, size_t cols, T* a, ptrdiff_t rs, ptrdiff_t cs,
                            llh_uplo uplo, llh_error* err) {
  if (bad_ctx(ctx)) return invalid(err, "null or destroyed context");
  if (rows != cols) return not_square(err, rows, cols);
  const size_t n = rows;
  if (n == 0) return ok(err);
  if (a == nullptr) return invalid(err, "null matrix pointer");
  std::vector<T> h(n * n);
  gather(a, rs, cs, n, n, h.data());
  void* dv = nullptr;
  CHECK_HIP(ctx_buffer(ctx, 0, n * n * sizeof(T), &dv), "device buffer");
  T* d = static_cast<T*>(dv);
  CHECK_HIP(hipMemcpyAsync(d, h.data(), n * n * sizeof(T), hipMemcpyHostToDevice, ctx->stream), "upload");
  CHECK_HIP(launch_zero_triangle<T>(d, (int64_t)n, (int)n, uplo == LLH_LOWER, nullptr, ctx->stream), "zero");
  CHECK_HIP(hipMemcpyAsync(h.data(), d, n * n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream), "download");
  CHECK_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
  scatter(h.data(), n, n, a, rs, cs, [n](size_t) { return n; });
  return ok(err);
}

template <typename T>
llh_status is_triangular_plain(llh_context* ctx, size_t rows, size_t cols, const T* a, ptrdiff_t rs,
                               ptrdiff_t cs, llh_uplo uplo, int* out, llh_error* err) {
  if (bad_ctx(ctx)) return invalid(err, "null or destroyed context");
  if (out == nullptr) return invalid(err, "null output pointer");
  if (rows != cols) {  // triangular.rs:88-89: non-square is not triangular
    *out = 0;
    return ok(err);
  }
  const size_t n = rows;
  if (n <= 1) {
    *out = 1;
    return ok(err);
  }
  if (a == nullptr) return invalid(err, "null matrix pointer");
  std::vector<T> h(n * n);
  gather(a, rs, cs, n, n, h.data());
  void* dv = nullptr;
  CHECK_HIP(ctx_buffer(ctx, 0, n * n * sizeof(T), &dv), "device buffer");
  T* d = static_cast<T*>(dv);
  CHECK_HIP(hipMemcpyAsync(d, h.data(), n * n * sizeof(T), hipMemcpyHostToDevice, ctx->stream), "upload");
  CHECK_HIP(launch_fill_i32(ctx->d_info + 1, 1, 1, ctx->stream), "flag");
  CHECK_HIP(launch_is_triangular<T>(d, (int64_t)n, (int)n, uplo == LLH_UPPER, ctx->d_info + 1, ctx->stream),
            "is_triangular");
  CHECK_HIP(hipMemcpyAsync(ctx->h_info + 1, ctx->d_info + 1, sizeof(int32_t), hipMemcpyDeviceToHost,
                           ctx->stream),
            "flag download");
  CHECK_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
  *out = ctx->h_info[1] ? 1 : 0;
  return ok(err);
}

template <typename T>
llh_status solve_triangular_plain(llh_context* ctx, size_t rows, size_t cols, const T* a, ptrdiff_t rs,
                                  ptrdiff_t cs, size_t b_rows, size_t b_cols, T* b, ptrdiff_t brs,
                                  ptrdiff_t bcs, llh_uplo uplo, llh_error* err, const T* diag = nullptr,
                                  ptrdiff_t dstride = 0, int abs_diag = 0) {
  if (bad_ctx(ctx)) return invalid(err, "null or destroyed context");
  if (rows != cols) return not_square(err, rows, cols);  // triangular.rs:102
  const size_t n = rows;
  if (b_rows != n) return wrong_rows(err, n, b_rows);  // triangular.rs:103-108
  if (n == 0 || b_cols == 0) return ok(err);
  if (a == nullptr || b == nullptr) return invalid(err, "null matrix pointer");
  std::vector<T> hav, hbv;
  T *ha = nullptr, *hb = nullptr;
  CHECK_HIP(ctx_stage<T>(ctx, 0, n * n, hav, &ha), "host staging");
  CHECK_HIP(ctx_stage<T>(ctx, 1, n * b_cols, hbv, &hb), "host staging");
  gather(a, rs, cs, n, n, ha);
  gather(b, brs, bcs, n, b_cols, hb);
  if (diag != nullptr) {
    // diag_fn: the kernel (triangular.rs:113-124) reads a's diagonal only as the
    // pivot, so the packed copy carries the caller's pivots in its place
    for (size_t i = 0; i < n; ++i) {
      const T d = diag[(ptrdiff_t)i * dstride];
      ha[i * n + i] = (abs_diag && d < T(0)) ? -d : d;
    }
  }
  void *dav = nullptr, *dbv = nullptr;
  CHECK_HIP(ctx_buffer(ctx, 0, n * n * sizeof(T), &dav), "device buffer");
  CHECK_HIP(ctx_buffer(ctx, 1, n * b_cols * sizeof(T), &dbv), "device buffer");
  T* da = static_cast<T*>(dav);
  T* db = static_cast<T*>(dbv);
  CHECK_HIP(hipMemcpyAsync(da, ha, n * n * sizeof(T), hipMemcpyHostToDevice, ctx->stream), "upload");
  CHECK_HIP(hipMemcpyAsync(db, hb, n * b_cols * sizeof(T), hipMemcpyHostToDevice, ctx->stream),
            "upload");
  void* dinv = nullptr;
  CHECK_HIP(ctx_dinv(ctx, (int)n, sizeof(T), &dinv), "inverse buffer");
  SolveSync sync;
  CHECK_HIP(ctx_sync(ctx, (int)n, (int)b_cols, &sync), "solve sync");
  CHECK_HIP(trsm_enqueue<T>(uplo == LLH_UPPER, false, (int)n, da, (int64_t)n, (int)b_cols, db,
                            (int64_t)b_cols, nullptr, false, static_cast<T*>(dinv), &sync, ctx->stream,
                            ctx->profiler()),
            "trsm");
  CHECK_HIP(ctx_read_solve_error(ctx), "solve status");
  CHECK_HIP(hipMemcpyAsync(hb, db, n * b_cols * sizeof(T), hipMemcpyDeviceToHost, ctx->stream),
            "download");
  CHECK_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
  if (ctx->h_info[2] != 0) return solve_stalled(err);
  scatter(hb, n, b_cols, b, brs, bcs, [b_cols](size_t) { return b_cols; });
  return ok(err);
}

// factor_solve_enqueue on the context stream, the fused forward substitution's
// workspace taken from the grow-only buffer fwd_y.
template <typename T>
hipError_t ctx_factor_solve(llh_context* ctx, int n, T* a, int64_t lda, int m, T* b, int64_t ldb, T* dinv,
                            SolveSync* sync) {
  const Lookahead* la = ctx_lookahead(ctx);
  T* y = nullptr;
  if (forward_fusable<T>(n, m, la)) {
    bool fresh = false;
    hipError_t e = grow(ctx, ctx->fwd_y, (size_t)n * m * sizeof(T), &fresh);
    if (e != hipSuccess) return e;
    y = static_cast<T*>(ctx->fwd_y.ptr);
  }
  return factor_solve_enqueue<T>(a, lda, n, m, b, ldb, dinv, y, sync, ctx->d_info, ctx->stream, ctx->profiler(), la);
}

// Small systems go through the exact-order fused kernel (bit-identical to the
// reference's factor and both sweeps); larger ones through the blocked path.
template <typename T>
llh_status solve_on_device(llh_context* ctx, size_t n, T* da, size_t m, T* db, llh_error* err) {
  if ((int)n <= small_max<T>()) {
    SmallDesc<T> desc{da, db, (int64_t)n, (int64_t)m, (int32_t)n, (int32_t)m, ctx->d_info};
    void* ddv = nullptr;
    CHECK_HIP(ctx_buffer(ctx, 2, sizeof(desc), &ddv), "descriptor buffer");
    CHECK_HIP(hipMemcpyAsync(ddv, &desc, sizeof(desc), hipMemcpyHostToDevice, ctx->stream), "descriptor");
    CHECK_HIP(launch_potrf_potrs_exact_batched<T>(static_cast<SmallDesc<T>*>(ddv), 1, (int)n, 1, ctx->stream),
              "small factor+solve");
  } else {
    void* dv = nullptr;
    CHECK_HIP(ctx_dinv(ctx, (int)n, sizeof(T), &dv), "inverse buffer");
    T* dinv = static_cast<T*>(dv);
    SolveSync sync;
    CHECK_HIP(ctx_sync(ctx, (int)n, (int)m, &sync), "solve sync");
    CHECK_HIP(ctx_factor_solve<T>(ctx, (int)n, da, (int64_t)n, (int)m, db, (int64_t)m, dinv, &sync),
              "factor and solve");
  }
  CHECK_HIP(ctx_read_solve_error(ctx), "solve status");
  return LLH_OK;
}

// A^-1 of the device matrix da (n x n, lda) into out (ldo): the factor in place
// (da := dirty factor), then, when it succeeded, W = L^-1 and W^T W on MFMA
// (potri_enqueue) -- about 1.5 n^3 flops where the reference's solvec on the
// identity (cholesky.rs:178-182) takes 2 n^3 + n^3 / 3.  The status is left in
// ctx->h_info[0]; out is untouched after NotPositiveDefinite.  n must exceed
// the exact-order kernel's bound (below it the caller takes that kernel).
template <typename T>
llh_status invert_on_device(llh_context* ctx, size_t n, T* da, size_t lda, T* out, size_t ldo, llh_error* err) {
  void* dv = nullptr;
  CHECK_HIP(ctx_dinv(ctx, (int)n, sizeof(T), &dv), "inverse buffer");
  T* dinv = static_cast<T*>(dv);
  void* wv = nullptr;
  const size_t wel = n * n, tel = (size_t)potri_tmp_elems((int)n);
  CHECK_HIP(ctx_buffer(ctx, 3, (wel + tel) * sizeof(T), &wv), "inverse work buffer");
  T* w = static_cast<T*>(wv);
  CHECK_HIP(potrf_enqueue<T>(da, (int64_t)lda, (int)n, ctx->d_info, false, dinv, ctx->stream, ctx->profiler(), false,
                             ctx_lookahead(ctx)),
            "potrf");
  CHECK_HIP(hipMemcpyAsync(ctx->h_info, ctx->d_info, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream),
            "status");
  CHECK_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
  ctx->h_info[2] = 0;  // no one-launch solve runs here: no wait to report
  if (ctx->h_info[0] != kNoFail) return LLH_OK;
  CHECK_HIP(potri_enqueue<T>((int)n, da, (int64_t)lda, dinv, w, (int64_t)n, out, (int64_t)ldo, w + wel, ctx->stream),
            "inverse");
  return LLH_OK;
}

template <typename T>
llh_status solvec_plain(llh_context* ctx, size_t rows, size_t cols, T* a, ptrdiff_t rs, ptrdiff_t cs,
                        size_t b_rows, size_t b_cols, T* b, ptrdiff_t brs, ptrdiff_t bcs, T* inv_out,
                        ptrdiff_t ors, ptrdiff_t ocs, llh_error* err) {
  if (bad_ctx(ctx)) return invalid(err, "null or destroyed context");
  if (rows != cols) return not_square(err, rows, cols);  // cholesky.rs:52 via :140
  const size_t n = rows;
  const bool invert = inv_out != nullptr;
  const size_t m = invert ? n : b_cols;
  if (n == 0) {
    if (!invert && b_rows != 0) return wrong_rows(err, 0, b_rows);
    return ok(err);
  }
  if (a == nullptr || (!invert && b == nullptr && b_cols > 0)) return invalid(err, "null matrix pointer");
  if (n > 0x7fffffff || m > 0x7fffffff) return invalid(err, "matrix too large");
  std::vector<T> hav, hbv;
  T* ha = nullptr;
  CHECK_HIP(ctx_stage<T>(ctx, 0, n * n, hav, &ha), "host staging");
  gather(a, rs, cs, n, n, ha);
  const bool rows_ok = invert || b_rows == n;
  if (!rows_ok || m == 0) {
    // The factor runs (and A keeps the dirty factor) before WrongRows is
    // reported, cholesky.rs:140-141 with triangular.rs:103-108.
    int64_t fail = -1;
    llh_status st = factor_plain<T>(ctx, n, a, rs, cs, false, ha, nullptr, &fail, err);
    if (st != LLH_OK) return st;
    if (fail >= 0) return not_pd(err, fail);
    if (!rows_ok) return wrong_rows(err, n, b_rows);
    return ok(err);
  }
  T* hb = nullptr;
  CHECK_HIP(ctx_stage<T>(ctx, 1, n * m, hbv, &hb), "host staging");
  void *dav = nullptr, *dbv = nullptr;
  CHECK_HIP(ctx_buffer(ctx, 0, n * n * sizeof(T), &dav), "device buffer");
  CHECK_HIP(ctx_buffer(ctx, 1, n * m * sizeof(T), &dbv), "device buffer");
  T* da = static_cast<T*>(dav);
  T* db = static_cast<T*>(dbv);
  CHECK_HIP(hipMemcpyAsync(da, ha, n * n * sizeof(T), hipMemcpyHostToDevice, ctx->stream), "upload");
  llh_status st = LLH_OK;
  if (invert && (int)n > small_max<T>()) {
    st = invert_on_device<T>(ctx, n, da, n, db, n, err);
  } else {
    if (invert) {
      CHECK_HIP(launch_set_identity<T>(db, (int64_t)n, (int)n, ctx->stream), "identity");  // cholesky.rs:179
    } else {
      gather(b, brs, bcs, n, m, hb);
      CHECK_HIP(hipMemcpyAsync(db, hb, n * m * sizeof(T), hipMemcpyHostToDevice, ctx->stream), "upload");
    }
    st = solve_on_device<T>(ctx, n, da, m, db, err);
  }
  if (st != LLH_OK) return st;
  CHECK_HIP(hipMemcpyAsync(ha, da, n * n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream), "download");
  CHECK_HIP(hipMemcpyAsync(hb, db, n * m * sizeof(T), hipMemcpyDeviceToHost, ctx->stream), "download");
  CHECK_HIP(hipMemcpyAsync(ctx->h_info, ctx->d_info, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream),
            "status");
  CHECK_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
  const int32_t f = ctx->h_info[0];
  if (f != kNoFail) {
    const size_t fr = (size_t)f;
    scatter(ha, fr + 1, n, a, rs, cs, [fr](size_t i) { return i < fr ? i + 1 : fr; });
    return not_pd(err, f);
  }
  if (ctx->h_info[2] != 0) return solve_stalled(err);
  scatter(ha, n, n, a, rs, cs, [](size_t i) { return i + 1; });  // dirty factor
  if (invert)
    scatter(hb, n, n, inv_out, ors, ocs, [n](size_t) { return n; });
  else
    scatter(hb, n, m, b, brs, bcs, [m](size_t) { return m; });
  return ok(err);
}

// ---- device-resident forms ------------------------------------------------

template <typename T>
llh_status dev_potrf(llh_context* ctx, size_t n, T* a, size_t lda, int zero_upper, llh_error* err) {
  if (bad_ctx(ctx)) return invalid(err, "null or destroyed context");
  if (n == 0) return ok(err);
  if (a == nullptr || lda < n) return invalid(err, "null matrix or lda < n");
  if (n > 0x7fffffff) return invalid(err, "matrix too large");
  void* dinv = nullptr;
  CHECK_HIP(ctx_dinv(ctx, (int)n, sizeof(T), &dinv), "inverse buffer");
  CHECK_HIP(potrf_enqueue<T>(a, (int64_t)lda, (int)n, ctx->d_info, zero_upper != 0, static_cast<T*>(dinv),
                            ctx->stream, ctx->profiler(), false, ctx_lookahead(ctx)),
            "potrf");
  CHECK_HIP(hipMemcpyAsync(ctx->h_info, ctx->d_info, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream),
            "status");
  CHECK_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
  if (ctx->h_info[0] != kNoFail) return not_pd(err, ctx->h_info[0]);
  return ok(err);
}

template <typename T>
llh_status dev_solve_triangular(llh_context* ctx, llh_uplo uplo, int trans, size_t n, const T* a, size_t lda,
                                size_t m, T* b, size_t ldb, llh_error* err) {
  if (bad_ctx(ctx)) return invalid(err, "null or destroyed context");
  if (n == 0 || m == 0) return ok(err);
  if (a == nullptr || b == nullptr || lda < n || ldb < m) return invalid(err, "bad matrix arguments");
  void* dinv = nullptr;
  CHECK_HIP(ctx_dinv(ctx, (int)n, sizeof(T), &dinv), "inverse buffer");
  SolveSync sync;
  CHECK_HIP(ctx_sync(ctx, (int)n, (int)m, &sync), "solve sync");
  CHECK_HIP(trsm_enqueue<T>(uplo == LLH_UPPER, trans != 0, (int)n, a, (int64_t)lda, (int)m, b, (int64_t)ldb,
                            nullptr, false, static_cast<T*>(dinv), &sync, ctx->stream, ctx->profiler()),
            "trsm");
  CHECK_HIP(ctx_read_solve_error(ctx), "solve status");
  CHECK_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
  if (ctx->h_info[2] != 0) return solve_stalled(err);
  return ok(err);
}

template <typename T>
llh_status dev_solvec(llh_context* ctx, size_t n, T* a, size_t lda, size_t m, T* b, size_t ldb,
                      llh_error* err) {
  if (bad_ctx(ctx)) return invalid(err, "null or destroyed context");
  if (n == 0) return ok(err);
  if (a == nullptr || lda < n || (m > 0 && (b == nullptr || ldb < m))) return invalid(err, "bad arguments");
  void* dv = nullptr;
  CHECK_HIP(ctx_dinv(ctx, (int)n, sizeof(T), &dv), "inverse buffer");
  T* dinv = static_cast<T*>(dv);
  SolveSync sync;
  CHECK_HIP(ctx_sync(ctx, (int)n, (int)std::max<size_t>(m, 1), &sync), "solve sync");
  // The solves are enqueued at once (one synchronisation per call) but gated on
  // the factor's status word: after NotPositiveDefinite every solve kernel is a
  // no-op and B keeps the caller's values, as the reference returns before
  // touching b (cholesky.rs:140).  The status is read back after them, so its
  // copy does not sit between the factor and the solves.
  CHECK_HIP(ctx_factor_solve<T>(ctx, (int)n, a, (int64_t)lda, (int)m, b, (int64_t)ldb, dinv, &sync),
            "factor and solve");
  CHECK_HIP(hipMemcpyAsync(ctx->h_info, ctx->d_info, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream),
            "status");
  CHECK_HIP(ctx_read_solve_error(ctx), "solve status");
  CHECK_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
  if (ctx->h_info[0] != kNoFail) return not_pd(err, ctx->h_info[0]);
  if (ctx->h_info[2] != 0) return solve_stalled(err);
  return ok(err);
}

// invc_inplace on device memory (cholesky.rs:166-183): a := dirty factor,
// out := A^-1 (both triangles).  Systems within the exact-order kernel's bound
// solve the identity in the reference's operation order; larger ones take the
// MFMA inverse (invert_on_device).
template <typename T>
llh_status dev_invc(llh_context* ctx, size_t n, T* a, size_t lda, T* out, size_t ldo, llh_error* err) {
  if (bad_ctx(ctx)) return invalid(err, "null or destroyed context");
  if (n == 0) return ok(err);
  if (a == nullptr || out == nullptr || lda < n || ldo < n) return invalid(err, "bad matrix arguments");
  if (n > 0x7fffffff) return invalid(err, "matrix too large");
  if ((int)n <= small_max<T>()) {
    CHECK_HIP(launch_set_identity<T>(out, (int64_t)ldo, (int)n, ctx->stream), "identity");
    SmallDesc<T> desc{a, out, (int64_t)lda, (int64_t)ldo, (int32_t)n, (int32_t)n, ctx->d_info};
    void* ddv = nullptr;
    CHECK_HIP(ctx_buffer(ctx, 2, sizeof(desc), &ddv), "descriptor buffer");
    CHECK_HIP(hipMemcpyAsync(ddv, &desc, sizeof(desc), hipMemcpyHostToDevice, ctx->stream), "descriptor");
    CHECK_HIP(launch_potrf_potrs_exact_batched<T>(static_cast<SmallDesc<T>*>(ddv), 1, (int)n, 1, ctx->stream),
              "small factor+solve");
    CHECK_HIP(hipMemcpyAsync(ctx->h_info, ctx->d_info, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream),
              "status");
  } else {
    llh_status st = invert_on_device<T>(ctx, n, a, lda, out, ldo, err);
    if (st != LLH_OK) return st;
  }
  CHECK_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
  if (ctx->h_info[0] != kNoFail) return not_pd(err, ctx->h_info[0]);
  return ok(err);
}

// ---- LOBPCG consumers (lobpcg/algorithm.rs:60-97) --------------------------

// Scratch of the LOBPCG forms in context slot 3: the Gram partial tiles, then a
// k x k (or ky x k) matrix, then (k > 64 only) an n x k product.
template <typename T>
hipError_t lobpcg_scratch(llh_context* ctx, int n, int ka, int kb, size_t extra, T** partial, T** small,
                          T** big) {
  int chunks = 0, rpc = 0;
  const size_t np = (size_t)gram_partial_elems(n, ka, kb, &chunks, &rpc);
  const size_t ns = (size_t)ka * kb;
  void* p = nullptr;
  hipError_t e = ctx_buffer(ctx, 3, (np + ns + extra) * sizeof(T), &p);
  if (e != hipSuccess) return e;
  *partial = static_cast<T*>(p);
  *small = *partial + np;
  *big = *small + ns;
  return hipSuccess;
}

// orthonormalize(v), lobpcg/algorithm.rs:75-97: G = V^T V; L = cholesky_into(G)
// (the clean lower factor, into r); U = (L^-1 V^T)^T = V L^-T, into v.  The
// triangular solve on V^T is done as V W^T with W = L^-1 (a k x k solve on the
// identity), one MFMA product over the n rows.
template <typename T>
llh_status dev_orthonormalize(llh_context* ctx, size_t n, size_t k, T* v, size_t ldv, T* r, size_t ldr,
                              llh_error* err) {
  if (bad_ctx(ctx)) return invalid(err, "null or destroyed context");
  if (k == 0) return ok(err);
  if (v == nullptr || r == nullptr || ldv < k || ldr < k) return invalid(err, "bad matrix arguments");
  if (n > 0x7fffffff || k > 0x7fffffff) return invalid(err, "matrix too large");
  const int N = (int)n, K = (int)k;
  T *part = nullptr, *w = nullptr, *tmp = nullptr;
  CHECK_HIP(lobpcg_scratch<T>(ctx, N, K, K, K > 64 ? n * k : 0, &part, &w, &tmp), "scratch");
  hipStream_t s = ctx->stream;
  CHECK_HIP(launch_gram<T>(N, K, K, v, (int64_t)ldv, v, (int64_t)ldv, r, (int64_t)ldr, part, s), "gram");
  void* dv = nullptr;
  CHECK_HIP(ctx_dinv(ctx, K, sizeof(T), &dv), "inverse buffer");
  T* dinv = static_cast<T*>(dv);
  CHECK_HIP(potrf_enqueue<T>(r, (int64_t)ldr, K, ctx->d_info, true, dinv, s, ctx->profiler(), false,
                            ctx_lookahead(ctx)),
            "potrf");
  CHECK_HIP(hipMemcpyAsync(ctx->h_info, ctx->d_info, sizeof(int32_t), hipMemcpyDeviceToHost, s), "status");
  CHECK_HIP(hipStreamSynchronize(s), "synchronize");
  if (ctx->h_info[0] != kNoFail) return not_pd(err, ctx->h_info[0]);  // cholesky_into()? (:81)
  SolveSync sync;
  CHECK_HIP(ctx_sync(ctx, K, K, &sync), "solve sync");
  CHECK_HIP(launch_set_identity<T>(w, (int64_t)K, K, s), "identity");
  CHECK_HIP(trsm_enqueue<T>(false, false, K, r, (int64_t)ldr, K, w, (int64_t)K, dinv, false, nullptr, &sync, s,
                            ctx->profiler()),
            "inverse");
  if (n > 0) {
    if (K <= 64) {
      // one 64-column tile holds all k columns of its rows: in place
      CHECK_HIP(launch_gemm<T>(false, true, false, N, K, K, T(1), v, (int64_t)ldv, w, (int64_t)K, T(0), v,
                               (int64_t)ldv, nullptr, s),
                "product");
    } else {
      CHECK_HIP(launch_gemm<T>(false, true, false, N, K, K, T(1), v, (int64_t)ldv, w, (int64_t)K, T(0), tmp,
                               (int64_t)K, nullptr, s),
                "product");
      CHECK_HIP(hipMemcpy2DAsync(v, ldv * sizeof(T), tmp, k * sizeof(T), k * sizeof(T), n, hipMemcpyDeviceToDevice,
                                 s),
                "copy");
    }
  }
  CHECK_HIP(ctx_read_solve_error(ctx), "solve status");
  CHECK_HIP(hipStreamSynchronize(s), "synchronize");
  if (ctx->h_info[2] != 0) return solve_stalled(err);
  return ok(err);
}

// apply_constraints(v, cholesky_yy, y), lobpcg/algorithm.rs:60-73:
// gram_yv = Y^T V; u = cholesky_yy.solve_triangular_into(gram_yv, Lower);
// V = V - Y u (general_mat_mul(-1, y, u, 1, v)).
template <typename T>
llh_status dev_apply_constraints(llh_context* ctx, size_t n, size_t k, T* v, size_t ldv, size_t ky, const T* y,
                                 size_t ldy, const T* l_yy, size_t ldl, llh_error* err) {
  if (bad_ctx(ctx)) return invalid(err, "null or destroyed context");
  if (k == 0 || ky == 0 || n == 0) return ok(err);
  if (v == nullptr || y == nullptr || l_yy == nullptr || ldv < k || ldy < ky || ldl < ky)
    return invalid(err, "bad matrix arguments");
  if (n > 0x7fffffff || k > 0x7fffffff || ky > 0x7fffffff) return invalid(err, "matrix too large");
  const int N = (int)n, K = (int)k, KY = (int)ky;
  T *part = nullptr, *g = nullptr, *unused = nullptr;
  CHECK_HIP(lobpcg_scratch<T>(ctx, N, KY, K, 0, &part, &g, &unused), "scratch");
  hipStream_t s = ctx->stream;
  CHECK_HIP(launch_gram<T>(N, KY, K, y, (int64_t)ldy, v, (int64_t)ldv, g, (int64_t)K, part, s), "gram");
  void* dv = nullptr;
  CHECK_HIP(ctx_dinv(ctx, KY, sizeof(T), &dv), "inverse buffer");
  SolveSync sync;
  CHECK_HIP(ctx_sync(ctx, KY, K, &sync), "solve sync");
  CHECK_HIP(trsm_enqueue<T>(false, false, KY, l_yy, (int64_t)ldl, K, g, (int64_t)K, nullptr, false,
                            static_cast<T*>(dv), &sync, s, ctx->profiler()),
            "solve");
  CHECK_HIP(launch_gemm<T>(false, false, false, N, K, KY, T(-1), y, (int64_t)ldy, g, (int64_t)K, T(1), v,
                           (int64_t)ldv, nullptr, s),
            "update");
  CHECK_HIP(ctx_read_solve_error(ctx), "solve status");
  CHECK_HIP(hipStreamSynchronize(s), "synchronize");
  if (ctx->h_info[2] != 0) return solve_stalled(err);
  return ok(err);
}

}  // namespace

extern "C" {

const char* llh_status_string(int32_t code) {
  switch (code) {
    case LLH_OK: return "ok";
    case LLH_ERR_NOT_SQUARE: return "NotSquare";
    case LLH_ERR_NOT_POSITIVE_DEFINITE: return "NotPositiveDefinite";
    case LLH_ERR_WRONG_ROWS: return "WrongRows";
    case LLH_ERR_INVALID_ARGUMENT: return "InvalidArgument";
    case LLH_ERR_HIP: return "HipError";
    case LLH_ERR_OUT_OF_MEMORY: return "OutOfMemory";
    case LLH_ERR_NO_CONVERGENCE: return "NoConvergence";
    case LLH_ERR_REJECTED: return "Rejected";
    default: return "unknown";
  }
}

llh_status llh_context_create(int device, llh_context** out, llh_error* err) {
  if (out == nullptr) return invalid(err, "null output pointer");
  *out = nullptr;
  int dev = device;
  if (dev < 0) CHECK_HIP(hipGetDevice(&dev), "hipGetDevice");
  CHECK_HIP(hipSetDevice(dev), "hipSetDevice");
  auto* ctx = new llh_context();
  ctx->device = dev;
  hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&ctx->d_info, 64 * sizeof(int32_t));
  if (e == hipSuccess) e = hipHostMalloc(&ctx->h_info, 64 * sizeof(int32_t), hipHostMallocDefault);
  if (e != hipSuccess) {
    llh_context_destroy(ctx);
    return hip_fail(err, e, "context create");
  }
  ctx->own_stream = true;
  // Set-up that later calls must not repeat: the factor's lookahead streams
  // (without them the factor runs on one stream), and this device's copy of the
  // allocation-free entry point's static scratch (that entry point reports a
  // failure here when it is called).
  (void)ctx_lookahead(ctx);
  (void)noalloc_init(dev, ctx->stream);
  *out = ctx;
  return ok(err);
}

void llh_context_destroy(llh_context* ctx) {
  if (ctx == nullptr || ctx->magic != llh_context::kMagic) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto& b : ctx->scratch)
    if (b.ptr) (void)hipFree(b.ptr);
  if (ctx->dinv.ptr) (void)hipFree(ctx->dinv.ptr);
  if (ctx->mixed_l.ptr) (void)hipFree(ctx->mixed_l.ptr);
  if (ctx->mixed_work.ptr) (void)hipFree(ctx->mixed_work.ptr);
  if (ctx->mixed_scale.ptr) (void)hipFree(ctx->mixed_scale.ptr);
  if (ctx->half_panel.ptr) (void)hipFree(ctx->half_panel.ptr);
  if (ctx->sync.ptr) (void)hipFree(ctx->sync.ptr);
  if (ctx->fwd_y.ptr) (void)hipFree(ctx->fwd_y.ptr);
  for (auto& b : ctx->host_stage)
    if (b.ptr) (void)hipHostFree(b.ptr);
  if (ctx->d_info) (void)hipFree(ctx->d_info);
  if (ctx->h_info) (void)hipHostFree(ctx->h_info);
  if (ctx->la.side) {
    (void)hipStreamSynchronize(ctx->la.side);
    (void)hipEventDestroy(ctx->la.fork);
    (void)hipEventDestroy(ctx->la.join);
    (void)hipStreamDestroy(ctx->la.side);
    if (ctx->la.fwd) {
      (void)hipStreamSynchronize(ctx->la.fwd);
      (void)hipEventDestroy(ctx->la.fwd_ev);
      (void)hipStreamDestroy(ctx->la.fwd);
    }
  }
  if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
  ctx->magic = 0;
  delete ctx;
}

llh_status llh_context_set_stream(llh_context* ctx, void* hip_stream, llh_error* err) {
  if (bad_ctx(ctx)) return invalid(err, "null or destroyed context");
  CHECK_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
  ctx->stream = static_cast<hipStream_t>(hip_stream);
  ctx->own_stream = false;
  return ok(err);
}

void* llh_context_stream(llh_context* ctx) { return bad_ctx(ctx) ? nullptr : (void*)ctx->stream; }

llh_status llh_context_set_solve_spin_limit(llh_context* ctx, int64_t limit, llh_error* err) {
  if (bad_ctx(ctx)) return invalid(err, "null or destroyed context");
  ctx->spin_limit = limit == 0 ? kSpinLimit : limit;
  return ok(err);
}

llh_status llh_context_set_profiling(llh_context* ctx, int enable, llh_error* err) {
  if (bad_ctx(ctx)) return invalid(err, "null or destroyed context");
  CHECK_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
  for (auto& r : ctx->prof.recs) {
    ctx->prof.pool.push_back(r.start);
    ctx->prof.pool.push_back(r.stop);
  }
  ctx->prof.recs.clear();
  ctx->prof.on = enable != 0;
  ctx->prof.mask = ~0u;
  return ok(err);
}

llh_status llh_context_set_profiling_classes(llh_context* ctx, uint32_t class_mask, llh_error* err) {
  llh_status st = llh_context_set_profiling(ctx, class_mask != 0, err);
  if (st != LLH_OK) return st;
  ctx->prof.mask = class_mask;
  return ok(err);
}

llh_status llh_context_kernel_stats(llh_context* ctx, int kernel_class, llh_kernel_stats* out,
                                    llh_error* err) {
  if (bad_ctx(ctx)) return invalid(err, "null or destroyed context");
  if (out == nullptr || kernel_class < 0 || kernel_class >= kNumClasses) return invalid(err, "bad arguments");
  CHECK_HIP(hipStreamSynchronize(ctx->stream), "synchronize");
  llh_kernel_stats st{0, 0.0, 0.0, 0.0};
  for (auto& r : ctx->prof.recs) {
    if (r.cls != kernel_class) continue;
    float ms = 0.f;
    CHECK_HIP(hipEventElapsedTime(&ms, r.start, r.stop), "hipEventElapsedTime");
    st.launches += 1;
    st.total_ms += ms;
    st.flops += r.flops;
    st.bytes += r.bytes;
  }
  *out = st;
  return ok(err);
}

#define LLH_PLAIN_IMPL(T, sfx)                                                                          \
  llh_status llh_cholesky_inplace_dirty_##sfx(llh_context* ctx, size_t rows, size_t cols, T* a,         \
                                              ptrdiff_t rs, ptrdiff_t cs, llh_error* err) {             \
    return cholesky_plain<T>(ctx, rows, cols, a, rs, cs, false, err);                                   \
  }                                                                                                     \
  llh_status llh_cholesky_inplace_##sfx(llh_context* ctx, size_t rows, size_t cols, T* a, ptrdiff_t rs,  \
                                        ptrdiff_t cs, llh_error* err) {                                 \
    return cholesky_plain<T>(ctx, rows, cols, a, rs, cs, true, err);                                    \
  }                                                                                                     \
  llh_status llh_triangular_inplace_##sfx(llh_context* ctx, size_t rows, size_t cols, T* a, ptrdiff_t rs,\
                                          ptrdiff_t cs, llh_uplo uplo, llh_error* err) {                \
    return triangular_plain<T>(ctx, rows, cols, a, rs, cs, uplo, err);                                  \
  }                                                                                                     \
  llh_status llh_is_triangular_##sfx(llh_context* ctx, size_t rows, size_t cols, const T* a,            \
                                     ptrdiff_t rs, ptrdiff_t cs, llh_uplo uplo, int* out,               \
                                     llh_error* err) {                                                  \
    return is_triangular_plain<T>(ctx, rows, cols, a, rs, cs, uplo, out, err);                          \
  }                                                                                                     \
  llh_status llh_solve_triangular_inplace_##sfx(llh_context* ctx, size_t rows, size_t cols, const T* a, \
                                                ptrdiff_t rs, ptrdiff_t cs, size_t b_rows,              \
                                                size_t b_cols, T* b, ptrdiff_t brs, ptrdiff_t bcs,      \
                                                llh_uplo uplo, llh_error* err) {                        \
    return solve_triangular_plain<T>(ctx, rows, cols, a, rs, cs, b_rows, b_cols, b, brs, bcs, uplo,     \
                                     err);                                                              \
  }                                                                                                     \
  llh_status llh_solve_triangular_system_##sfx(                                                         \
      llh_context* ctx, size_t rows, size_t cols, const T* a, ptrdiff_t rs, ptrdiff_t cs, const T* diag,  \
      ptrdiff_t dstride, int abs_diag, size_t b_rows, size_t b_cols, T* b, ptrdiff_t brs, ptrdiff_t bcs,  \
      llh_uplo uplo, llh_error* err) {                                                                  \
    if (diag == nullptr && rows > 0 && rows == cols && b_rows == rows && b_cols > 0)                    \
      return invalid(err, "null diagonal pointer");                                                     \
    return solve_triangular_plain<T>(ctx, rows, cols, a, rs, cs, b_rows, b_cols, b, brs, bcs, uplo, err, \
                                     diag, dstride, abs_diag);                                          \
  }                                                                                                     \
  llh_status llh_solvec_inplace_##sfx(llh_context* ctx, size_t rows, size_t cols, T* a, ptrdiff_t rs,    \
                                      ptrdiff_t cs, size_t b_rows, size_t b_cols, T* b, ptrdiff_t brs,  \
                                      ptrdiff_t bcs, llh_error* err) {                                  \
    return solvec_plain<T>(ctx, rows, cols, a, rs, cs, b_rows, b_cols, b, brs, bcs, nullptr, 0, 0, err); \
  }                                                                                                     \
  llh_status llh_invc_inplace_##sfx(llh_context* ctx, size_t rows, size_t cols, T* a, ptrdiff_t rs,      \
                                    ptrdiff_t cs, T* out, ptrdiff_t ors, ptrdiff_t ocs, llh_error* err) { \
    if (out == nullptr && rows > 0) return invalid(err, "null output pointer");                         \
    if (rows == 0) {                                                                                    \
      if (rows != cols) return not_square(err, rows, cols);                                             \
      return ok(err);                                                                                   \
    }                                                                                                   \
    return solvec_plain<T>(ctx, rows, cols, a, rs, cs, rows, rows, nullptr, 0, 0, out, ors, ocs, err);   \
  }                                                                                                     \
  llh_status llh_dev_potrf_##sfx(llh_context* ctx, size_t n, T* a, size_t lda, int zero_upper,          \
                                 llh_error* err) {                                                      \
    return dev_potrf<T>(ctx, n, a, lda, zero_upper, err);                                               \
  }                                                                                                     \
  llh_status llh_dev_solve_triangular_##sfx(llh_context* ctx, llh_uplo uplo, int trans, size_t n,       \
                                            const T* a, size_t lda, size_t m, T* b, size_t ldb,         \
                                            llh_error* err) {                                           \
    return dev_solve_triangular<T>(ctx, uplo, trans, n, a, lda, m, b, ldb, err);                        \
  }                                                                                                     \
  llh_status llh_dev_solvec_##sfx(llh_context* ctx, size_t n, T* a, size_t lda, size_t m, T* b,         \
                                  size_t ldb, llh_error* err) {                                         \
    return dev_solvec<T>(ctx, n, a, lda, m, b, ldb, err);                                               \
  }                                                                                                     \
  llh_status llh_dev_invc_##sfx(llh_context* ctx, size_t n, T* a, size_t lda, T* out, size_t ldo,        \
                                llh_error* err) {                                                       \
    return dev_invc<T>(ctx, n, a, lda, out, ldo, err);                                                  \
  }                                                                                                     \
  llh_status llh_dev_orthonormalize_##sfx(llh_context* ctx, size_t n, size_t k, T* v, size_t ldv, T* r,  \
                                          size_t ldr, llh_error* err) {                                 \
    return dev_orthonormalize<T>(ctx, n, k, v, ldv, r, ldr, err);                                       \
  }                                                                                                     \
  llh_status llh_dev_apply_constraints_##sfx(llh_context* ctx, size_t n, size_t k, T* v, size_t ldv,     \
                                             size_t ky, const T* y, size_t ldy, const T* l_yy,          \
                                             size_t ldl, llh_error* err) {                              \
    return dev_apply_constraints<T>(ctx, n, k, v, ldv, ky, y, ldy, l_yy, ldl, err);                     \
  }

LLH_PLAIN_IMPL(double, d)
LLH_PLAIN_IMPL(float, s)

}  // extern "C"

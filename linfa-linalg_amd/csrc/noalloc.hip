// Allocation-free device solvec (llh_dev_solvec_noalloc_d): the same factor
// and solves as llh_dev_solvec_d, with the scratch those take from the
// context's grow-only buffers (the inverses of the 64 x 64 diagonal blocks, the
// right-hand sides of the forward substitution fused into the factor, and the
// ticket, error word and flags of the one-launch solve) placed in a static
// device array instead.  Each GPU the library runs on holds its own copy of the
// array until the process ends; llh_context_create locates it once per device
// (hipGetSymbolAddress with that device current) and every later use goes
// through that per-device address.
#include <cstdio>
#include <mutex>

#include "llh.h"
#include "llh_context.h"
#include "llh_driver.h"
#include "llh_kernels.h"

namespace {

constexpr size_t kSyncBytes = 16u * 1024u;          // [0] ticket, [1] error word, [2..] flags
constexpr size_t kInvBytes = 8u * 1024u * 1024u;    // inverses of the diagonal blocks
constexpr size_t kFwdBytes = 8u * 1024u * 1024u;    // fused forward substitution: n x m, m <= 64
constexpr size_t kScratchBytes = kSyncBytes + kInvBytes + kFwdBytes;
static_assert(kScratchBytes == LLH_NOALLOC_SCRATCH_BYTES, "the header states the array's size");
static_assert(kInvBytes / (llh::kNB * sizeof(double)) == LLH_NOALLOC_MAX_N, "the header states the bound");
static_assert(kFwdBytes == (size_t)LLH_NOALLOC_MAX_N * 64 * sizeof(double),
              "the fused forward substitution's right-hand sides fit (it takes m <= 64)");
// flags of the largest solve: one per (64-row block, 16-column chunk) at m <= 64
static_assert(2 + (LLH_NOALLOC_MAX_N / llh::kNB) * (64 / llh::kSolveChunkHost) <= kSyncBytes / 4,
              "the flags of the largest one-launch solve fit");

// The array itself.  In host code its name is only a stand-in for the symbol:
// kernels and copies take the per-device address below, never the name.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
struct DeviceScratch {
  bool located = false;
  hipError_t status = hipErrorNotInitialized;
  unsigned char* ptr = nullptr;  // this device's address of `scratch`
  int32_t epoch = 1;             // flag epochs of the solves that use this device's array
};
std::mutex g_mu;
DeviceScratch g_scratch[kMaxDevices];

}  // namespace

namespace llh {

// Locates this device's copy of the array and clears its flags (they start at
// 0, below every epoch handed out).  Runs with `device` current, once per
// device; later calls return the first outcome.
hipError_t noalloc_init(int device, hipStream_t s) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_mu);
  DeviceScratch& d = g_scratch[device];
  if (d.located) return d.status;
  d.located = true;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e == hipSuccess && p == nullptr) e = hipErrorInvalidSymbol;
  if (e == hipSuccess) e = launch_fill_i32(static_cast<int32_t*>(p), 0, (int64_t)(kSyncBytes / 4), s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) (void)hipGetLastError();
  d.status = e;
  d.ptr = e == hipSuccess ? static_cast<unsigned char*>(p) : nullptr;
  return e;
}

}  // namespace llh

using namespace llh;

extern "C" llh_status llh_dev_solvec_noalloc_d(llh_context* ctx, size_t n, double* a, size_t lda, size_t m,
                                               double* b, size_t ldb, llh_error* err) {
  if (ctx == nullptr || ctx->magic != llh_context::kMagic) return invalid(err, "null or destroyed context");
  if (n == 0) return ok(err);
  if (a == nullptr || lda < n || (m > 0 && (b == nullptr || ldb < m))) return invalid(err, "bad arguments");
  if (n > (size_t)LLH_NOALLOC_MAX_N) {
    char msg[128];
    std::snprintf(msg, sizeof(msg), "n = %zu exceeds the allocation-free scratch (n <= %d)", n, LLH_NOALLOC_MAX_N);
    return invalid(err, msg);
  }
  if (m > 0x7fffffff) return invalid(err, "too many right-hand sides");
  if (ctx->device < 0 || ctx->device >= kMaxDevices) return invalid(err, "device out of range");
  unsigned char* base = nullptr;
  int32_t* epoch = nullptr;
  {
    std::lock_guard<std::mutex> g(g_mu);
    DeviceScratch& d = g_scratch[ctx->device];
    if (d.ptr == nullptr) return hip_fail(err, d.status, "allocation-free scratch");
    base = d.ptr;
    epoch = &d.epoch;
  }
  hipStream_t s = ctx->stream;
  double* dinv = reinterpret_cast<double*>(base + kSyncBytes);
  double* fwd_y = reinterpret_cast<double*>(base + kSyncBytes + kInvBytes);
  SolveSync sync;
  sync.words = reinterpret_cast<int32_t*>(base);
  sync.nflags = (int64_t)(kSyncBytes / 4) - 2;
  sync.epoch = epoch;
  sync.spin_limit = ctx->spin_limit;
  // the lookahead streams exist from llh_context_create on; nothing is created here
  const Lookahead* la = ctx->la.side != nullptr ? &ctx->la : nullptr;
#define NA_TRY(expr, where)                                 \
  do {                                                      \
    hipError_t e_ = (expr);                                 \
    if (e_ != hipSuccess) return hip_fail(err, e_, where);  \
  } while (0)
  NA_TRY(solve_sync_begin(sync, s), "solve sync");
  // the same sequence as llh_dev_solvec_d (solves gated on the factor's status
  // word); the fused forward substitution, where it applies, takes m <= 64
  NA_TRY(factor_solve_enqueue<double>(a, (int64_t)lda, (int)n, (int)m, b, (int64_t)ldb, dinv, fwd_y, &sync,
                                      ctx->d_info, s, nullptr, la),
         "factor and solve");
  NA_TRY(hipMemcpyAsync(ctx->h_info, ctx->d_info, sizeof(int32_t), hipMemcpyDeviceToHost, s), "status");
  NA_TRY(hipMemcpyAsync(ctx->h_info + 2, sync.words + 1, sizeof(int32_t), hipMemcpyDeviceToHost, s),
         "solve status");
  NA_TRY(hipStreamSynchronize(s), "synchronize");
#undef NA_TRY
  if (ctx->h_info[0] != kNoFail) return not_pd(err, ctx->h_info[0]);
  if (ctx->h_info[2] != 0)
    return set_error(err, LLH_ERR_HIP, 0, 0, -1, "triangular solve: a block wait exceeded its bound", 0);
  return ok(err);
}

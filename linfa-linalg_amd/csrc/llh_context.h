// The context behind the opaque llh_context handle of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "llh.h"
#include "llh_driver.h"
#include "llh_profile.h"

struct llh_context {
  static constexpr uint32_t kMagic = 0x4c4c4843;  // "LLHC"
  uint32_t magic = kMagic;
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int32_t* d_info = nullptr;  // device status words: [0] factor status, [1] flag
  int32_t* h_info = nullptr;  // pinned host mirror
  struct Buffer {
    void* ptr = nullptr;
    size_t bytes = 0;
  };
  Buffer scratch[4];  // grow-only device buffers of the plain forms
  Buffer dinv;        // inverses of the 64 x 64 diagonal blocks
  Buffer mixed_l;     // factor storage of the one-shot mixed-precision solve
  Buffer mixed_work;  // its residual / correction / norm buffers
  Buffer mixed_scale; // its diagonal scaling (f16 factor)
  Buffer half_panel;  // f16 copy of the panel of the f16 factor's trailing SYRK
  Buffer sync;        // ticket / error / flags of the one-launch solve
  Buffer fwd_y;       // right-hand sides of the forward substitution fused into the factor
  Buffer host_stage[2];  // page-locked host staging of the plain forms (hipHostMalloc)
  int32_t epoch = 1;
  int64_t spin_limit = llh::kSpinLimit;  // bound on the one-launch solve's flag waits
  llh::Lookahead la;  // side stream and events of the factor's lookahead (created on first use)
  llh::Profiler prof;  // live kernel timing (off unless enabled)
  llh::Profiler* profiler() { return prof.on ? &prof : nullptr; }
};

namespace llh {
llh_status set_error(llh_error* err, llh_status code, uint64_t a, uint64_t b, int64_t column, const char* msg,
                     int hip_error);
llh_status ok(llh_error* err);
llh_status hip_fail(llh_error* err, hipError_t e, const char* where);
llh_status invalid(llh_error* err, const char* msg);
llh_status not_square(llh_error* err, size_t rows, size_t cols);
llh_status not_pd(llh_error* err, int64_t col);
llh_status wrong_rows(llh_error* err, size_t expected, size_t actual);
hipError_t ctx_buffer(llh_context* ctx, int slot, size_t bytes, void** out);
// Synchronizes the context stream and the lookahead streams.
hipError_t ctx_quiesce(llh_context* ctx);
// Grow-only diagonal-inverse buffer for order n (elements of type size es).
hipError_t ctx_dinv(llh_context* ctx, int n, size_t es, void** out);
// Solve-sync words sized for order n and m right-hand sides; fills *sync.
hipError_t ctx_sync(llh_context* ctx, int n, int m, SolveSync* sync);
// The lookahead streams of the blocked factor, created on first use (nullptr if
// they cannot be created: the factor then runs on the context stream alone).
const Lookahead* ctx_lookahead(llh_context* ctx);
// Enqueues the read-back of the solve error word into h_info[2].
hipError_t ctx_read_solve_error(llh_context* ctx);
// Locates `device`'s copy of the allocation-free entry point's static scratch
// (noalloc.hip); runs once per device, with that device current.
hipError_t noalloc_init(int device, hipStream_t s);
}  // namespace llh

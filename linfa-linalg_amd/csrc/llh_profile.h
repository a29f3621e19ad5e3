// Live per-kernel-class timing with HIP events recorded on the launch stream.
// A profiler is attached to a context; the drivers bracket the launches of each
// kernel class with an event pair and add the launch's algorithmic work.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace llh {

enum KernelClass : int {
  kTrailingSyrk = 0,  // trailing update of the POTRF, rank = the outer panel width (MFMA)
  kPanelUpdate = 1,   // inner-panel update of the POTRF (MFMA)
  kDiagFactor = 2,    // diagonal-block factor in LDS
  kPanelTrsm = 3,     // panel row solve
  kSolveGemm = 4,     // block update of the triangular solves (MFMA)
  kSolveDiag = 5,     // diagonal-block solve of the triangular solves
  kNumClasses = 6
};

struct Profiler {
  bool on = false;
  uint32_t mask = ~0u;  // classes recorded while on
  struct Rec {
    hipEvent_t start, stop;
    int cls;
    double flops;
    double bytes;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;

  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  // Records the start event of one launch; returns its index or -1 when off.
  int begin(int cls, double flops, double bytes, hipStream_t s) {
    if (!on || !(mask & (1u << cls))) return -1;
    Rec r{get(), get(), cls, flops, bytes};
    (void)hipEventRecord(r.start, s);
    recs.push_back(r);
    return (int)recs.size() - 1;
  }
  void end(int idx, hipStream_t s) {
    if (idx >= 0) (void)hipEventRecord(recs[idx].stop, s);
  }
  ~Profiler() {
    for (auto& r : recs) {
      (void)hipEventDestroy(r.start);
      (void)hipEventDestroy(r.stop);
    }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

}  // namespace llh

// Fast factor + inverse of one 64 x 64 diagonal block for the blocked paths
// (the device code is in k_diag_impl.h): one workgroup per block, or per
// (system, block) of a grouped batch step.
#include "k_diag_impl.h"
#include "k_subst_impl.h"
#include "llh_kernels.h"

namespace llh {
namespace {

template <typename TS, typename T>
__global__ __launch_bounds__(kDiagT) void diag_fast_kernel(TS* __restrict__ a, int64_t lda, int b,
                                                           TS* __restrict__ inv, int32_t* __restrict__ info,
                                                           int64_t info_base, const int32_t* __restrict__ gate) {
  __shared__ __attribute__((aligned(16))) DiagLds<T> sm;
  // the gate is read once and shared, so all four waves leave or stay together
  if (threadIdx.x == 0) sm.skip = gate != nullptr && *gate != kNoFail;
  __syncthreads();
  if (sm.skip) return;
  diag_quad_block<TS, T>(a, lda, b, inv, info, info_base, sm);
}

// Grouped form for the batch planner: work w = (item, k0) factors and inverts
// the block at (k0, k0) of its system; a failure sets the system's status.
// With a forward workspace (bb.y) it also takes the block's forward step.
template <typename T>
__global__ __launch_bounds__(kDiagT, 4) void grp_diag_fast_kernel(const BatchBase<T>* __restrict__ base,
                                                           const BatchItem* __restrict__ items,
                                                           const StepWork* __restrict__ work, int count) {
  if ((int)blockIdx.x >= count) return;
  const StepWork w = work[blockIdx.x];
  const BatchItem it = items[w.item];
  const BatchBase<T> bb = *base;
  int32_t* st = bb.status + it.id;
  __shared__ __attribute__((aligned(16))) DiagLds<T> sm;
  if (threadIdx.x == 0) sm.skip = *st != kNoFail;
  __syncthreads();
  if (sm.skip) return;
  const bool fwd = bb.y != nullptr;
  FwdDiagPrefetch<T> f{};
  if (fwd) f = fwd_diag_prefetch<T>(bb, it, w.k0);
  if (diag_quad_block<T, T>(bb.a + it.a_off + (int64_t)w.k0 * it.lda + w.k0, it.lda, min(kNB, it.n - w.k0),
                            bb.linv + it.linv_off + (int64_t)(w.k0 / kNB) * kNB * kNB, st, w.k0, sm))
    return;
  if (fwd) {
    __syncthreads();  // the inverse is stored; the block's LDS is free
    static_assert(sizeof(sm.l) >= kNB * (kNB + 1) * sizeof(T) && sizeof(sm.xd) >= kNB * kSolveCols * sizeof(T),
                  "the forward step's tiles fit the block's LDS");
    grp_forward_diag<T>(bb, it, w.k0, sm.l, &sm.xd[0][0], f);
  }
}

}  // namespace

template <typename T>
hipError_t launch_grp_diag_fast(const BatchBase<T>* base, const BatchItem* items, const StepWork* work,
                                int count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL((grp_diag_fast_kernel<T>), dim3(count), dim3(kDiagT), 0, s, base, items, work, count);
  return hipGetLastError();
}
template hipError_t launch_grp_diag_fast<double>(const BatchBase<double>*, const BatchItem*, const StepWork*,
                                                 int, hipStream_t);
template hipError_t launch_grp_diag_fast<float>(const BatchBase<float>*, const BatchItem*, const StepWork*, int,
                                                hipStream_t);

template <typename TS, typename T>
hipError_t launch_diag_fast(TS* a, int64_t lda, int b, TS* inv, int32_t* info, int64_t info_base,
                            const int32_t* gate, hipStream_t s) {
  if (b <= 0) return hipSuccess;
  if (b > kNB) return hipErrorInvalidValue;
  hipLaunchKernelGGL((diag_fast_kernel<TS, T>), dim3(1), dim3(kDiagT), 0, s, a, lda, b, inv, info, info_base,
                     gate);
  return hipGetLastError();
}

template hipError_t launch_diag_fast<double, double>(double*, int64_t, int, double*, int32_t*, int64_t,
                                                     const int32_t*, hipStream_t);
template hipError_t launch_diag_fast<float, float>(float*, int64_t, int, float*, int32_t*, int64_t,
                                                   const int32_t*, hipStream_t);
template hipError_t launch_diag_fast<float, double>(float*, int64_t, int, float*, int32_t*, int64_t,
                                                    const int32_t*, hipStream_t);

}  // namespace llh

// Exact-order Cholesky and SPD solve for small matrices, one workgroup per
// matrix, the matrix held in LDS.
//
// The reference factors row by row (cholesky.rs:54-74): L[j,k] = (A[j,k] - s) /
// L[k,k] with s = sum_{i<k} L[k,i]*L[j,i] accumulated in i order, and
// d = A[j,j] - sum_{k<j} L[j,k]^2 accumulated in k order.  Here the same sums are
// accumulated right-looking: after column k is final, every pending sum
// S[j,k'] (k < k' <= j) receives L[k',k]*L[j,k].  Each S[j,k'] therefore sees
// its terms in the same order i = 0, 1, ..., k'-1, starting from +0, with a
// separately rounded multiply and add, so every entry of L is bit-identical to
// the reference's while all entries of a column are computed in parallel.
//
// The solves follow triangular.rs:113-125 the same way: for each pivot i (in the
// reference's row order), coeff = b[i] / diag and every later entry gets
// b[r] = b[r] + ((-coeff) * a[r, i]) (ndarray's scaled_add), in parallel over r
// and over right-hand-side columns, one pivot after another.
#pragma clang fp contract(off)

#include <algorithm>

#include "llh_common.h"
#include "llh_kernels.h"
#include "k_small_impl.h"

namespace llh {
namespace {

// TS is the storage type of `a`, T the arithmetic type (f32 storage with f64
// arithmetic factors the diagonal blocks of the mixed-precision path).
template <typename TS, typename T, int NMAX, int NT>
__global__ __launch_bounds__(NT) void potrf_exact_single(TS* a, int64_t lda, int n, int32_t* info,
                                                         int64_t info_base, const int32_t* gate) {
  if (gate != nullptr && *gate != kNoFail) return;
  __shared__ T W[(NMAX * (NMAX + 1)) / 2];
  __shared__ T Ap[(NMAX * (NMAX + 1)) / 2];
  __shared__ T Dg[NMAX];
  const int fail = factor_exact<TS, T, NT, NMAX>(a, lda, n, W, Ap, Dg);
  __syncthreads();
  write_back<TS, T, NT>(a, lda, n, fail, W, Dg);
  if (fail >= 0 && threadIdx.x == 0) atomicMin(info, (int32_t)(info_base + fail));
}

// Batched factor + solve: matrix i of the batch is descs[i].
template <typename T, int NMAX, int NT, int MCH>
__global__ __launch_bounds__(NT) void potrf_potrs_exact_batched(const SmallDesc<T>* __restrict__ descs,
                                                                int count, int do_solve) {
  const int bi = blockIdx.x;
  if (bi >= count) return;
  __shared__ SmallLds<T, NMAX, MCH> lds;
  small_solve_one<T, NMAX, NT, MCH>(descs[bi], do_solve, lds);
}

// Persistent batch kernel: each workgroup takes matrices from a work queue
// until it is empty (items are ordered largest first by the planner).
template <typename T, int NMAX, int NT, int MCH>
__global__ __launch_bounds__(NT) void small_batch_kernel(const BatchBase<T>* __restrict__ base,
                                                         const BatchItem* __restrict__ items, int count,
                                                         int counter, int mode) {
  __shared__ SmallLds<T, NMAX, MCH> lds;
  __shared__ int s_next;
  const BatchBase<T> bb = *base;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_next = atomicAdd(bb.counters + counter, 1);
    __syncthreads();
    const int i = s_next;
    if (i >= count) return;
    const BatchItem it = items[i];
    SmallDesc<T> ds{bb.a + it.a_off, bb.b + it.b_off, it.lda, it.m, it.n, it.m, bb.status + it.id};
    small_solve_one<T, NMAX, NT, MCH>(ds, mode, lds);
  }
}

// Grouped diagonal-block factor: work w = (item, k0) factors the block of order
// min(64, n - k0) at (k0, k0) in the exact order; a failure sets the matrix's
// status to its global row and makes the matrix's later launches no-ops.
template <typename T>
__global__ __launch_bounds__(256) void grp_diag_kernel(const BatchBase<T>* __restrict__ base,
                                                       const BatchItem* __restrict__ items,
                                                       const StepWork* __restrict__ work, int count) {
  if ((int)blockIdx.x >= count) return;
  const StepWork w = work[blockIdx.x];
  const BatchItem it = items[w.item];
  const BatchBase<T> bb = *base;
  int32_t* st = bb.status + it.id;
  if (*st != kNoFail) return;
  __shared__ T W[(kNB * (kNB + 1)) / 2];
  __shared__ T Ap[(kNB * (kNB + 1)) / 2];
  __shared__ T Dg[kNB];
  T* a = bb.a + it.a_off + (int64_t)w.k0 * it.lda + w.k0;
  const int b = min(kNB, it.n - w.k0);
  const int fail = factor_exact<T, T, 256, kNB>(a, it.lda, b, W, Ap, Dg);
  __syncthreads();
  write_back<T, T, 256>(a, it.lda, b, fail, W, Dg);
  if (fail >= 0 && threadIdx.x == 0) atomicMin(st, w.k0 + fail);
}

}  // namespace

template <typename T>
hipError_t launch_small_batch(const BatchBase<T>* base, const BatchItem* items, int count, int nmax, int grid,
                              int counter, hipStream_t s, int mode) {
  if (count <= 0) return hipSuccess;
  grid = std::max(1, std::min(grid, count));
  if (nmax <= 32) {
    hipLaunchKernelGGL((small_batch_kernel<T, 32, 64, 8>), dim3(grid), dim3(64), 0, s, base, items, count, counter,
                       mode);
  } else if (nmax <= 64) {
    hipLaunchKernelGGL((small_batch_kernel<T, 64, 128, 8>), dim3(grid), dim3(128), 0, s, base, items, count,
                       counter, mode);
  } else if (nmax <= small_max<T>()) {
    hipLaunchKernelGGL((small_batch_kernel<T, small_max_c<T>::value, 256, 8>), dim3(grid), dim3(256), 0, s, base,
                       items, count, counter, mode);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_grp_diag(const BatchBase<T>* base, const BatchItem* items, const StepWork* work, int count,
                           hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL((grp_diag_kernel<T>), dim3(count), dim3(256), 0, s, base, items, work, count);
  return hipGetLastError();
}

template hipError_t launch_small_batch<double>(const BatchBase<double>*, const BatchItem*, int, int, int, int,
                                               hipStream_t, int);
template hipError_t launch_small_batch<float>(const BatchBase<float>*, const BatchItem*, int, int, int, int,
                                              hipStream_t, int);
template hipError_t launch_grp_diag<double>(const BatchBase<double>*, const BatchItem*, const StepWork*, int,
                                            hipStream_t);
template hipError_t launch_grp_diag<float>(const BatchBase<float>*, const BatchItem*, const StepWork*, int,
                                           hipStream_t);

template <typename TS, typename T>
hipError_t launch_potrf_exact_single(TS* a, int64_t lda, int n, int32_t* info, int64_t info_base,
                                     const int32_t* gate, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (n <= 64) {
    hipLaunchKernelGGL((potrf_exact_single<TS, T, 64, 256>), dim3(1), dim3(256), 0, s, a, lda, n, info,
                       info_base, gate);
  } else if (n <= small_max<T>()) {
    hipLaunchKernelGGL((potrf_exact_single<TS, T, small_max_c<T>::value, 256>), dim3(1), dim3(256), 0, s,
                       a, lda, n, info, info_base, gate);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_potrf_potrs_exact_batched(const SmallDesc<T>* descs, int count, int nmax, int do_solve,
                                            hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (nmax <= 32) {
    hipLaunchKernelGGL((potrf_potrs_exact_batched<T, 32, 64, 8>), dim3(count), dim3(64), 0, s, descs,
                       count, do_solve);
  } else if (nmax <= 64) {
    hipLaunchKernelGGL((potrf_potrs_exact_batched<T, 64, 128, 8>), dim3(count), dim3(128), 0, s, descs,
                       count, do_solve);
  } else if (nmax <= small_max<T>()) {
    hipLaunchKernelGGL((potrf_potrs_exact_batched<T, small_max_c<T>::value, 256, 8>), dim3(count),
                       dim3(256), 0, s, descs, count, do_solve);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template hipError_t launch_potrf_exact_single<double, double>(double*, int64_t, int, int32_t*, int64_t,
                                                              const int32_t*, hipStream_t);
template hipError_t launch_potrf_exact_single<float, float>(float*, int64_t, int, int32_t*, int64_t,
                                                            const int32_t*, hipStream_t);
template hipError_t launch_potrf_exact_single<float, double>(float*, int64_t, int, int32_t*, int64_t,
                                                             const int32_t*, hipStream_t);
template hipError_t launch_potrf_potrs_exact_batched<double>(const SmallDesc<double>*, int, int, int,
                                                             hipStream_t);
template hipError_t launch_potrf_potrs_exact_batched<float>(const SmallDesc<float>*, int, int, int,
                                                            hipStream_t);

}  // namespace llh

// Triangular solves on small diagonal blocks.
//
// trsm_rows: the panel step of the blocked POTRF, X := X * L11^-T, i.e. for
// every row x of the panel, x[c] = (a[c] - sum_{p<c} x[p] * L11[c,p]) / L11[c,c]
// (the division of cholesky.rs:62 applied to a whole block of rows).  One
// thread owns one row, held in registers; L11 sits in LDS transposed and every
// read of it is a wave-wide broadcast.  The sums are accumulated column-sweep
// style (after x[c] is final, every later x[q] -= x[c] * L11[q,c]), so each
// step issues up to 63 independent FMAs instead of one dependent chain.
//
// trsv_block: the diagonal-block step of the blocked triangular solve: one
// thread owns one right-hand-side column, held in registers, swept exactly as
// triangular.rs:113-125 sweeps it (divide by the pivot, then update the rest of
// the column), restricted to b <= 64 rows.
#include "llh_common.h"
#include "llh_kernels.h"

namespace llh {
namespace {

template <typename T, bool VEC>
__global__ __launch_bounds__(64) void trsm_rows_kernel(int m, int b, T* __restrict__ X, int64_t ldx,
                                                       const T* __restrict__ L, int64_t ldl,
                                                       const int32_t* __restrict__ gate) {
  if (gate != nullptr && *gate != kNoFail) return;
  __shared__ __attribute__((aligned(16))) T Lt[kNB * kNB];  // Lt[c * kNB + q] = L11[q, c]
  const int tid = threadIdx.x;
  for (int e = tid; e < kNB * kNB; e += 64) {
    const int q = e / kNB, c = e % kNB;  // coalesced read of row q
    Lt[c * kNB + q] = (q < b && c <= q) ? L[(int64_t)q * ldl + c] : T(0);
  }
  __syncthreads();
  const int row = blockIdx.x * 64 + tid;
  if (row >= m) return;
  T* xr = X + (int64_t)row * ldx;
  T x[kNB];
  if constexpr (VEC && sizeof(T) == 8) {
#pragma unroll
    for (int c = 0; c < kNB; c += 2) {
      if (c < b) {
        const double2 v = *reinterpret_cast<const double2*>(xr + c);
        x[c] = v.x;
        x[c + 1] = v.y;
      } else {
        x[c] = T(0);
        x[c + 1] = T(0);
      }
    }
  } else {
#pragma unroll
    for (int c = 0; c < kNB; ++c) x[c] = (c < b) ? xr[c] : T(0);
  }
#pragma unroll
  for (int c = 0; c < kNB; ++c) {
    if (c < b) {
      x[c] = x[c] / Lt[c * kNB + c];
#pragma unroll
      for (int q = c + 1; q < kNB; ++q) x[q] -= x[c] * Lt[c * kNB + q];
    }
  }
#pragma unroll
  for (int c = 0; c < kNB; ++c)
    if (c < b) xr[c] = x[c];
}

template <typename T, bool UPPER_OP>
__global__ __launch_bounds__(64) void trsv_block_kernel(int b, bool trans, const T* __restrict__ Tm,
                                                        int64_t ldt, int m, T* __restrict__ Y, int64_t ldy) {
  // St[c * kNB + r] = op(T)[r, c], only op(T)'s triangle: column c of op(T) is
  // contiguous, which is what the column sweep reads.
  __shared__ __attribute__((aligned(16))) T St[kNB * kNB];
  const int tid = threadIdx.x;
  for (int e = tid; e < kNB * kNB; e += 64) {
    const int i = e / kNB, j = e % kNB;               // reads T[i, j], coalesced over j
    const int r = trans ? j : i, c = trans ? i : j;   // T[i, j] is op(T)[r, c]
    const bool in = r < b && c < b && (UPPER_OP ? c >= r : c <= r);
    St[c * kNB + r] = in ? Tm[(int64_t)i * ldt + j] : T(0);
  }
  __syncthreads();
  const int col = blockIdx.x * 64 + tid;
  if (col >= m) return;
  T y[kNB];
#pragma unroll
  for (int r = 0; r < kNB; ++r) y[r] = (r < b) ? Y[(int64_t)r * ldy + col] : T(0);
  if constexpr (!UPPER_OP) {
#pragma unroll
    for (int r = 0; r < kNB; ++r) {
      if (r < b) {
        y[r] = y[r] / St[r * kNB + r];
#pragma unroll
        for (int p = r + 1; p < kNB; ++p) y[p] -= y[r] * St[r * kNB + p];
      }
    }
  } else {
#pragma unroll
    for (int r = kNB - 1; r >= 0; --r) {
      if (r < b) {
        y[r] = y[r] / St[r * kNB + r];
#pragma unroll
        for (int p = 0; p < r; ++p) y[p] -= y[r] * St[r * kNB + p];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kNB; ++r)
    if (r < b) Y[(int64_t)r * ldy + col] = y[r];
}

}  // namespace

template <typename T>
hipError_t launch_trsm_rows(int m, int b, T* X, int64_t ldx, const T* L, int64_t ldl, const int32_t* gate,
                            hipStream_t s) {
  if (m <= 0 || b <= 0) return hipSuccess;
  if (b > kNB) return hipErrorInvalidValue;
  const dim3 grid((m + 63) / 64);
  const bool vec = (reinterpret_cast<uintptr_t>(X) % 16 == 0) && (ldx % 2 == 0) && (b % 2 == 0);
  if (vec)
    hipLaunchKernelGGL((trsm_rows_kernel<T, true>), grid, dim3(64), 0, s, m, b, X, ldx, L, ldl, gate);
  else
    hipLaunchKernelGGL((trsm_rows_kernel<T, false>), grid, dim3(64), 0, s, m, b, X, ldx, L, ldl, gate);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_trsv_block(bool upper_op, bool trans, int b, const T* Tm, int64_t ldt, int m, T* Y,
                             int64_t ldy, hipStream_t s) {
  if (m <= 0 || b <= 0) return hipSuccess;
  if (b > kNB) return hipErrorInvalidValue;
  const dim3 grid((m + 63) / 64);
  if (upper_op)
    hipLaunchKernelGGL((trsv_block_kernel<T, true>), grid, dim3(64), 0, s, b, trans, Tm, ldt, m, Y, ldy);
  else
    hipLaunchKernelGGL((trsv_block_kernel<T, false>), grid, dim3(64), 0, s, b, trans, Tm, ldt, m, Y, ldy);
  return hipGetLastError();
}

template hipError_t launch_trsm_rows<double>(int, int, double*, int64_t, const double*, int64_t,
                                             const int32_t*, hipStream_t);
template hipError_t launch_trsm_rows<float>(int, int, float*, int64_t, const float*, int64_t, const int32_t*,
                                            hipStream_t);
template hipError_t launch_trsv_block<double>(bool, bool, int, const double*, int64_t, int, double*, int64_t,
                                              hipStream_t);
template hipError_t launch_trsv_block<float>(bool, bool, int, const float*, int64_t, int, float*, int64_t,
                                             hipStream_t);

}  // namespace llh

// Elementwise helpers of the path: status initialisation, triangular_inplace
// (triangular.rs:37-53), is_triangular (triangular.rs:67-90) and the identity
// right-hand side of invc (cholesky.rs:179).
#include "llh_common.h"
#include "llh_kernels.h"

namespace llh {
namespace {

__global__ void fill_i32_kernel(int32_t* p, int32_t v, int64_t count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}

// Zeroes the strict upper (zero_upper) or strict lower triangle.  2-D grid:
// x over columns, y over rows.
template <typename T>
__global__ void zero_triangle_kernel(T* a, int64_t lda, int n, bool zero_upper, const int32_t* gate) {
  if (gate != nullptr && *gate != kNoFail) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (c >= n || r >= n) return;
  if (zero_upper ? (c > r) : (c < r)) a[(int64_t)r * lda + c] = T(0);
}

// Clears *flag if any entry of the other triangle is non-zero (`is_zero` is
// `== 0`, so -0.0 counts as zero and NaN does not).
template <typename T>
__global__ void is_triangular_kernel(const T* a, int64_t lda, int n, bool upper, int32_t* flag) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (c >= n || r >= n) return;
  const bool other = upper ? (c < r) : (c > r);
  if (other && !(a[(int64_t)r * lda + c] == T(0))) *flag = 0;
}

template <typename T>
__global__ void set_identity_kernel(T* a, int64_t lda, int n) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (c >= n || r >= n) return;
  a[(int64_t)r * lda + c] = (r == c) ? T(1) : T(0);
}

dim3 grid2(int n) { return dim3((n + 255) / 256, n); }

}  // namespace

hipError_t launch_fill_i32(int32_t* p, int32_t v, int64_t count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_i32_kernel, dim3((count + 255) / 256), dim3(256), 0, s, p, v, count);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_zero_triangle(T* a, int64_t lda, int n, bool zero_upper, const int32_t* gate,
                                hipStream_t s) {
  if (n <= 1) return hipSuccess;
  hipLaunchKernelGGL((zero_triangle_kernel<T>), grid2(n), dim3(256), 0, s, a, lda, n, zero_upper, gate);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_is_triangular(const T* a, int64_t lda, int n, bool upper, int32_t* flag, hipStream_t s) {
  if (n <= 1) return hipSuccess;
  hipLaunchKernelGGL((is_triangular_kernel<T>), grid2(n), dim3(256), 0, s, a, lda, n, upper, flag);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_set_identity(T* a, int64_t lda, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL((set_identity_kernel<T>), grid2(n), dim3(256), 0, s, a, lda, n);
  return hipGetLastError();
}

template hipError_t launch_zero_triangle<double>(double*, int64_t, int, bool, const int32_t*, hipStream_t);
template hipError_t launch_zero_triangle<float>(float*, int64_t, int, bool, const int32_t*, hipStream_t);
template hipError_t launch_is_triangular<double>(const double*, int64_t, int, bool, int32_t*, hipStream_t);
template hipError_t launch_is_triangular<float>(const float*, int64_t, int, bool, int32_t*, hipStream_t);
template hipError_t launch_set_identity<double>(double*, int64_t, int, hipStream_t);
template hipError_t launch_set_identity<float>(float*, int64_t, int, hipStream_t);

}  // namespace llh

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

// a[r][c] = a[c][r] for c > r: the upper triangle mirrors the lower.  One
// 32 x 32 tile pair per workgroup (tile (ti, tj), tj >= ti, receives the
// transpose of tile (tj, ti)), staged through LDS so reads and writes are both
// along rows.
template <typename T>
__global__ __launch_bounds__(256) void mirror_lower_kernel(T* a, int64_t lda, int n) {
  __shared__ T tile[32][33];
  const int64_t t = blockIdx.x;
  int64_t tj = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while (tj * (tj + 1) / 2 > t) --tj;
  while ((tj + 1) * (tj + 2) / 2 <= t) ++tj;
  const int64_t ti = t - tj * (tj + 1) / 2;  // ti <= tj
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  // read tile (tj, ti): rows tj*32 + y, columns ti*32 + x (lower part)
  for (int y = ty; y < 32; y += 8) {
    const int64_t r = tj * 32 + y, c = ti * 32 + tx;
    tile[y][tx] = (r < n && c < n) ? a[r * lda + c] : T(0);
  }
  __syncthreads();
  // write tile (ti, tj): a[ti*32 + y][tj*32 + x] = a[tj*32 + x][ti*32 + y]
  for (int y = ty; y < 32; y += 8) {
    const int64_t r = ti * 32 + y, c = tj * 32 + tx;
    if (r < n && c < n && c > r) a[r * lda + c] = tile[tx][y];
  }
}

// dst (m x m, ldd) = the lower triangle of src (ldsrc), zeros above it.
template <typename T>
__global__ void copy_lower_kernel(int m, const T* src, int64_t lds, T* dst, int64_t ldd) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (c >= m || r >= m) return;
  dst[(int64_t)r * ldd + c] = c <= r ? src[(int64_t)r * lds + c] : T(0);
}

// dst (rows x m, ldd) = src (lds), unless the gate holds a failed row.  One
// thread per element; 2-D grid: x over columns in 64-wide strips, y over rows.
template <typename T>
__global__ void copy_rows_gated_kernel(int rows, int m, const T* __restrict__ src, int64_t lds, T* __restrict__ dst,
                                       int64_t ldd, const int32_t* __restrict__ gate) {
  if (gate != nullptr && *gate != kNoFail) return;
  const int c = blockIdx.x * 64 + (threadIdx.x & 63);
  const int r = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (c >= m || r >= rows) return;
  dst[(int64_t)r * ldd + c] = src[(int64_t)r * lds + c];
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

template <typename T>
hipError_t launch_mirror_lower(T* a, int64_t lda, int n, hipStream_t s) {
  if (n <= 1) return hipSuccess;
  const int64_t nt = (n + 31) / 32;
  hipLaunchKernelGGL((mirror_lower_kernel<T>), dim3((unsigned)(nt * (nt + 1) / 2)), dim3(256), 0, s, a, lda, n);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_copy_lower(int m, const T* src, int64_t lds, T* dst, int64_t ldd, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL((copy_lower_kernel<T>), dim3((m + 63) / 64, m), dim3(64), 0, s, m, src, lds, dst, ldd);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_copy_rows_gated(int rows, int m, const T* src, int64_t lds, T* dst, int64_t ldd,
                                  const int32_t* gate, hipStream_t s) {
  if (rows <= 0 || m <= 0) return hipSuccess;
  if ((rows + 3) / 4 > 65535) return hipErrorInvalidValue;  // grid.y bound
  hipLaunchKernelGGL((copy_rows_gated_kernel<T>), dim3((m + 63) / 64, (rows + 3) / 4), dim3(256), 0, s, rows, m,
                     src, lds, dst, ldd, gate);
  return hipGetLastError();
}

template hipError_t launch_copy_rows_gated<double>(int, int, const double*, int64_t, double*, int64_t,
                                                   const int32_t*, hipStream_t);
template hipError_t launch_copy_rows_gated<float>(int, int, const float*, int64_t, float*, int64_t, const int32_t*,
                                                  hipStream_t);
template hipError_t launch_mirror_lower<double>(double*, int64_t, int, hipStream_t);
template hipError_t launch_mirror_lower<float>(float*, int64_t, int, hipStream_t);
template hipError_t launch_copy_lower<double>(int, const double*, int64_t, double*, int64_t, hipStream_t);
template hipError_t launch_copy_lower<float>(int, const float*, int64_t, float*, int64_t, hipStream_t);
template hipError_t launch_zero_triangle<double>(double*, int64_t, int, bool, const int32_t*, hipStream_t);
template hipError_t launch_zero_triangle<float>(float*, int64_t, int, bool, const int32_t*, hipStream_t);
template hipError_t launch_is_triangular<double>(const double*, int64_t, int, bool, int32_t*, hipStream_t);
template hipError_t launch_is_triangular<float>(const float*, int64_t, int, bool, int32_t*, hipStream_t);
template hipError_t launch_set_identity<double>(double*, int64_t, int, hipStream_t);
template hipError_t launch_set_identity<float>(float*, int64_t, int, hipStream_t);

}  // namespace llh

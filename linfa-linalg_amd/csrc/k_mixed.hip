// Kernels of the mixed-precision solve: precision conversions, the f64 residual
// r = b - A x of a symmetric A held in its lower triangle (the part the
// reference reads, cholesky.rs:58-66), and the column norms of the refinement
// test.
#include "llh_common.h"
#include "llh_kernels.h"

namespace llh {
namespace {

// Y (rows x cols, f32) = rs[r] * X (f64) * cs[c]; lower_only restricts to
// j <= i; a null scale is 1.
__global__ void convert_d2s_kernel(int rows, int cols, const double* __restrict__ X, int64_t ldx,
                                   float* __restrict__ Y, int64_t ldy, bool lower_only,
                                   const double* __restrict__ rs, const double* __restrict__ cs) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y + blockIdx.z * 65535;
  if (r >= rows || c >= cols || (lower_only && c > r)) return;
  double x = X[(int64_t)r * ldx + c];
  if (rs != nullptr) x *= rs[r];
  if (cs != nullptr) x *= cs[c];
  Y[(int64_t)r * ldy + c] = (float)x;
}

// X (f64) += rs[r] * D (f32); a null scale is 1.
__global__ void axpy_s2d_kernel(int rows, int cols, const float* __restrict__ D, int64_t ldd,
                                double* __restrict__ X, int64_t ldx, const double* __restrict__ rs) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y + blockIdx.z * 65535;
  if (r >= rows || c >= cols) return;
  double d = (double)D[(int64_t)r * ldd + c];
  if (rs != nullptr) d *= rs[r];
  X[(int64_t)r * ldx + c] += d;
}

// Y (f16) = X (f32), rounded to nearest even, 8 columns per thread when the
// rows are 16-byte aligned.
template <bool VEC>
__global__ void convert_s2h_kernel(int rows, int cols, const float* __restrict__ X, int64_t ldx,
                                   _Float16* __restrict__ Y, int64_t ldy) {
  const int c = (blockIdx.x * blockDim.x + threadIdx.x) * 8;
  const int r = blockIdx.y + blockIdx.z * 65535;
  if (r >= rows || c >= cols) return;
  const float* x = X + (int64_t)r * ldx + c;
  _Float16* y = Y + (int64_t)r * ldy + c;
  if (VEC && c + 8 <= cols) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    typedef _Float16 h8 __attribute__((ext_vector_type(8)));
    const v4 a = *reinterpret_cast<const v4*>(x), b = *reinterpret_cast<const v4*>(x + 4);
    h8 h;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      h[e] = (_Float16)a[e];
      h[e + 4] = (_Float16)b[e];
    }
    *reinterpret_cast<h8*>(y) = h;
  } else {
    for (int e = 0; e < 8 && c + e < cols; ++e) y[e] = (_Float16)x[e];
  }
}

// s[i] = 1 / sqrt(A[i][i]) where the diagonal is positive and finite, else 1
// (the factor then meets the same non-positive pivot unscaled).
__global__ void diag_rsqrt_kernel(int n, const double* __restrict__ A, int64_t lda, double* __restrict__ s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = A[(int64_t)i * lda + i];
  s[i] = (d > 0.0 && d < INFINITY) ? 1.0 / sqrt(d) : 1.0;
}

// R = B - A X (ABS: R = |A| w, w = X as a vector or 1 when X is null: the row
// sums of |A|, weighted), A symmetric from its lower
// triangle.  Workgroup (w, q): 64 rows of R, 16 columns.  It walks the 64 x 64
// blocks of row block w of the full symmetric A: blocks left of the diagonal are
// read from the lower triangle row-wise, blocks right of it as transposed lower
// blocks, the diagonal block from both halves.  The next block is prefetched into
// registers while the current one is applied.
template <bool ABS>
__global__ __launch_bounds__(256) void symm_kernel(int n, int m, const double* __restrict__ A, int64_t lda,
                                                   const double* __restrict__ X, int64_t ldx,
                                                   const double* __restrict__ B, int64_t ldb,
                                                   double* __restrict__ R, int64_t ldr) {
  constexpr int S = kNB + 1, MC = 16, SC = MC + 1;
  __shared__ double As[kNB * S];
  __shared__ double Xs[kNB * SC];
  const int tid = threadIdx.x;
  const int w = blockIdx.x, q = blockIdx.y;
  const int nb = (n + kNB - 1) / kNB;
  const int w0 = w * kNB, bw = min(kNB, n - w0);
  const int c0 = q * MC, mc = ABS ? 1 : min(MC, m - c0);
  const int r = tid >> 2, cc = (tid & 3) * 4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  double pre[16];
  // Element e = tid + 256 p of block j: (i, jj) = (e / 64, e % 64) of the stored
  // 64 x 64 tile that is read coalesced; it lands at As[rr][kk].
  auto fetch = [&](int j) {
    const int j0 = j * kNB, bj = min(kNB, n - j0);
#pragma unroll
    for (int p = 0; p < 16; ++p) {
      const int e = tid + 256 * p, i = e / kNB, jj = e % kNB;
      double v = 0.0;
      if (j < w) {  // A[w0+i][j0+jj], lower
        if (i < bw && jj < bj) v = A[(int64_t)(w0 + i) * lda + j0 + jj];
      } else if (j > w) {  // A[j0+i][w0+jj] = As[jj][i]
        if (i < bj && jj < bw) v = A[(int64_t)(j0 + i) * lda + w0 + jj];
      } else {  // diagonal block: lower part read once, mirrored on store
        if (i < bw && jj <= i) v = A[(int64_t)(w0 + i) * lda + w0 + jj];
      }
      pre[p] = v;
    }
  };
  auto place = [&](int j) {
#pragma unroll
    for (int p = 0; p < 16; ++p) {
      const int e = tid + 256 * p, i = e / kNB, jj = e % kNB;
      const double v = ABS ? fabs(pre[p]) : pre[p];
      if (j < w) {
        As[i * S + jj] = v;
      } else if (j > w) {
        As[jj * S + i] = v;
      } else {
        if (jj <= i) {
          As[i * S + jj] = v;
          As[jj * S + i] = v;
        }
      }
    }
  };
  fetch(0);
  for (int j = 0; j < nb; ++j) {
    const int j0 = j * kNB, bj = min(kNB, n - j0);
    place(j);
    for (int e = tid; e < kNB * MC; e += 256) {
      const int k = e / MC, c = e % MC;
      Xs[k * SC + c] = ABS ? (k < bj && c == 0 ? (X != nullptr ? X[j0 + k] : 1.0) : 0.0)
                           : ((k < bj && c < mc) ? X[(int64_t)(j0 + k) * ldx + c0 + c] : 0.0);
    }
    __syncthreads();
    if (j + 1 < nb) fetch(j + 1);
#pragma unroll 8
    for (int k = 0; k < kNB; ++k) {
      const double a = As[r * S + k];
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[e] += a * Xs[k * SC + cc + e];
    }
    __syncthreads();
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (r < bw && cc + e < mc) {
      const int64_t row = w0 + r;
      R[row * ldr + c0 + cc + e] = ABS ? acc[e] : B[row * ldb + c0 + cc + e] - acc[e];
    }
  }
}

// out[c] = max(out[c], max_r |X[r][c] w_r|) for c < cols, w_r = rs[r] (or
// 1 / rs[r] with inv; 1 when rs is null).  out must start at 0; the maximum of
// non-negative doubles is the maximum of their bit patterns.
__global__ void colmax_kernel(int rows, int cols, const double* __restrict__ X, int64_t ldx,
                              unsigned long long* __restrict__ out, const double* __restrict__ rs, bool inv) {
  const int c = blockIdx.y;
  double mx = 0.0;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) {
    double v = fabs(X[(int64_t)r * ldx + c]);
    if (rs != nullptr) v = inv ? v / rs[r] : v * rs[r];
    mx = (v > mx || v != v) ? v : mx;  // a NaN wins, so it cannot look converged
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_down(mx, off, 64);
    mx = (o > mx || o != o) ? o : mx;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(out + c, (unsigned long long)__double_as_longlong(mx));
}

dim3 grid_rows(int rows, int cols) {
  return dim3((cols + 255) / 256, rows < 65535 ? rows : 65535, (rows + 65534) / 65535);
}

}  // namespace

hipError_t launch_convert_d2s(int rows, int cols, const double* X, int64_t ldx, float* Y, int64_t ldy,
                              bool lower_only, hipStream_t s, const double* rs, const double* cs) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  hipLaunchKernelGGL(convert_d2s_kernel, grid_rows(rows, cols), dim3(256), 0, s, rows, cols, X, ldx, Y, ldy,
                     lower_only, rs, cs);
  return hipGetLastError();
}

hipError_t launch_axpy_s2d(int rows, int cols, const float* D, int64_t ldd, double* X, int64_t ldx,
                           hipStream_t s, const double* rs) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  hipLaunchKernelGGL(axpy_s2d_kernel, grid_rows(rows, cols), dim3(256), 0, s, rows, cols, D, ldd, X, ldx, rs);
  return hipGetLastError();
}

hipError_t launch_convert_s2h(int rows, int cols, const float* X, int64_t ldx, _Float16* Y, int64_t ldy,
                              hipStream_t s) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  const int groups = (cols + 7) / 8;
  const dim3 grid((groups + 255) / 256, rows < 65535 ? rows : 65535, (rows + 65534) / 65535);
  const bool vec = reinterpret_cast<uintptr_t>(X) % 16 == 0 && reinterpret_cast<uintptr_t>(Y) % 16 == 0 &&
                   ldx % 4 == 0 && ldy % 8 == 0;
  if (vec)
    hipLaunchKernelGGL((convert_s2h_kernel<true>), grid, dim3(256), 0, s, rows, cols, X, ldx, Y, ldy);
  else
    hipLaunchKernelGGL((convert_s2h_kernel<false>), grid, dim3(256), 0, s, rows, cols, X, ldx, Y, ldy);
  return hipGetLastError();
}

hipError_t launch_diag_rsqrt(int n, const double* A, int64_t lda, double* s_out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(diag_rsqrt_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, A, lda, s_out);
  return hipGetLastError();
}

hipError_t launch_symm_residual(int n, int m, const double* A, int64_t lda, const double* X, int64_t ldx,
                                const double* B, int64_t ldb, double* R, int64_t ldr, hipStream_t s) {
  if (n <= 0 || m <= 0) return hipSuccess;
  const dim3 grid((n + kNB - 1) / kNB, (m + 15) / 16);
  hipLaunchKernelGGL((symm_kernel<false>), grid, dim3(256), 0, s, n, m, A, lda, X, ldx, B, ldb, R, ldr);
  return hipGetLastError();
}

hipError_t launch_symm_abs_rowsum(int n, const double* A, int64_t lda, double* out, hipStream_t s,
                                  const double* w) {
  if (n <= 0) return hipSuccess;
  const dim3 grid((n + kNB - 1) / kNB, 1);
  hipLaunchKernelGGL((symm_kernel<true>), grid, dim3(256), 0, s, n, 1, A, lda, w, 1, nullptr, 0, out, 1);
  return hipGetLastError();
}

hipError_t launch_colmax(int rows, int cols, const double* X, int64_t ldx, double* out, hipStream_t s,
                         const double* rs, bool inv) {
  if (cols <= 0) return hipSuccess;
  hipError_t e = launch_fill_i32(reinterpret_cast<int32_t*>(out), 0, 2 * (int64_t)cols, s);
  if (e != hipSuccess || rows <= 0) return e;
  const int bx = std::min(64, (rows + 255) / 256);
  hipLaunchKernelGGL(colmax_kernel, dim3(bx, cols), dim3(256), 0, s, rows, cols, X, ldx,
                     reinterpret_cast<unsigned long long*>(out), rs, inv);
  return hipGetLastError();
}

}  // namespace llh

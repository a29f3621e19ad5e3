// Kernels of the mixed-precision solve: precision conversions, the f64 residual
// r = b - A x of a symmetric A held in its lower triangle (the part the
// reference reads, cholesky.rs:58-66), and the column norms of the refinement
// test.
#include <cstdint>

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

// R = B - A X (ABS: R = |A| w, w = X as a vector or 1 when X is null: the
// weighted row sums of |A|), A symmetric from its lower triangle.
//
// Workgroup (i, q, sp): 64 rows of R (row block i), 16 columns (chunk q), and
// the blocks j of split sp of row block i of the full symmetric A.  Lane l of
// wave v holds rows p = l mod 32 and p + 32 of block j at the 8 columns
// k = 16v + 8h .. 16v + 8h + 7, h = l / 32, in registers, loaded one block
// ahead: blocks left of the diagonal row-wise from the lower triangle, blocks
// right of it as transposed lower blocks (for each k the 32 lanes of a half
// read 32 consecutive entries of one row), the diagonal block from both halves.
// X_j goes through a double-buffered LDS slab, one barrier per block; every
// value read from it (one address per half-wave) feeds two FMAs, one per row.
// With one row per lane and 16 k those broadcast reads bounded the walk (128
// ds_read_b128 per wave and block; ~4.2 TB/s of A at n = 32768).  The halves'
// partial sums are added by one lane exchange and the four waves' once through
// LDS at the end.  With one split the workgroup writes R; with several it
// writes its partial sum to P[sp] and symm_reduce_kernel adds the splits in
// order, so the result does not depend on scheduling either way.
constexpr int kSymmMC = 16;

template <bool ABS, bool VEC>
__global__ __launch_bounds__(256) void symm_walk_kernel(int n, int m, const double* __restrict__ A, int64_t lda,
                                                        const double* __restrict__ X, int64_t ldx,
                                                        const double* __restrict__ B, int64_t ldb,
                                                        double* __restrict__ R, int64_t ldr,
                                                        double* __restrict__ P, int nsplit, int mp) {
  constexpr int MC = kSymmMC, KH = kNB / 8, SL = kNB * MC;  // KH: k per lane
  __shared__ __attribute__((aligned(16))) double Ys[2][SL];
  __shared__ double Part[3][SL];
  const int i = blockIdx.x, q = blockIdx.y, sp = blockIdx.z;
  const int nb = (n + kNB - 1) / kNB;
  const int per = (nb + nsplit - 1) / nsplit;
  const int jb = sp * per, je = min(nb, jb + per);
  const int i0 = i * kNB, bi = min(kNB, n - i0);
  const int c0 = q * MC, mc = ABS ? 1 : min(MC, m - c0);
  const int tid = threadIdx.x, l = tid & 63, wv = tid >> 6;
  const int p = l & 31, h = l >> 5, k0 = 16 * wv + KH * h;
  // clamped rows p and p + 32: rows past the matrix compute and are not written
  const int rr0 = min(p, bi - 1), rr1 = min(p + 32, bi - 1);
  typedef double v2d __attribute__((ext_vector_type(2)));
  // out[0 .. KH) row p, out[KH .. 2 KH) row p + 32
  auto fetch = [&](int j, double* out) {
    const int j0 = j * kNB, bj = min(kNB, n - j0);
    if (j < i) {  // A[i0 + r][j0 + k]
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const double* row = A + (int64_t)(i0 + (s ? rr1 : rr0)) * lda + j0 + k0;
        if (VEC && k0 + KH <= bj) {
#pragma unroll
          for (int v = 0; v < KH / 2; ++v) {
            const v2d x = *reinterpret_cast<const v2d*>(row + 2 * v);
            out[s * KH + 2 * v] = x[0];
            out[s * KH + 2 * v + 1] = x[1];
          }
        } else {
#pragma unroll
          for (int kk = 0; kk < KH; ++kk) out[s * KH + kk] = (k0 + kk < bj) ? row[kk] : 0.0;
        }
      }
    } else if (j > i) {  // A[j0 + k][i0 + r]
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const double* col = A + (int64_t)(j0 + k0) * lda + i0 + (s ? rr1 : rr0);
#pragma unroll
        for (int kk = 0; kk < KH; ++kk) out[s * KH + kk] = (k0 + kk < bj) ? col[(int64_t)kk * lda] : 0.0;
      }
    } else {  // diagonal block, from its lower half
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int rr = s ? rr1 : rr0;
#pragma unroll
        for (int kk = 0; kk < KH; ++kk) {
          const int k = k0 + kk;
          const int64_t o =
              k <= rr ? (int64_t)(i0 + rr) * lda + i0 + k : (int64_t)(i0 + min(k, bi - 1)) * lda + i0 + rr;
          const double v = A[o];
          out[s * KH + kk] = k < bi ? v : 0.0;
        }
      }
    }
    if (ABS) {
#pragma unroll
      for (int kk = 0; kk < 2 * KH; ++kk) out[kk] = fabs(out[kk]);
    }
  };
  // X_j rows [0, 64) x chunk columns: 4 values per thread (row e / 16, column e % 16)
  auto loady = [&](int j, double* out) {
    const int j0 = j * kNB, bj = min(kNB, n - j0);
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int e = tid + 256 * g, k = e / MC, c = e % MC;
      if (ABS)
        out[g] = (k < bj && c == 0) ? (X != nullptr ? X[j0 + k] : 1.0) : 0.0;
      else
        out[g] = (k < bj && c < mc) ? X[(int64_t)(j0 + k) * ldx + c0 + c] : 0.0;
    }
  };

  double acc0[MC], acc1[MC];
#pragma unroll
  for (int c = 0; c < MC; ++c) acc0[c] = acc1[c] = 0.0;
  double a0[2 * KH], a1[2 * KH], y0[4];
  if (jb < je) {
    fetch(jb, a0);
    loady(jb, y0);
  }
  for (int j = jb; j < je; ++j) {
    double* ys = Ys[(j - jb) & 1];
#pragma unroll
    for (int e = 0; e < 4; ++e) ys[tid + 256 * e] = y0[e];
    __syncthreads();
    if (j + 1 < je) {
      fetch(j + 1, a1);
      loady(j + 1, y0);
    }
    const double* yk = ys + k0 * MC;
#pragma unroll
    for (int kk = 0; kk < KH; ++kk) {
#pragma unroll
      for (int c = 0; c < MC; ++c) {
        const double y = yk[kk * MC + c];
        acc0[c] += a0[kk] * y;
        acc1[c] += a0[KH + kk] * y;
      }
    }
#pragma unroll
    for (int kk = 0; kk < 2 * KH; ++kk) a0[kk] = a1[kk];
  }
  // the two halves of a wave hold the same rows at different k
#pragma unroll
  for (int c = 0; c < MC; ++c) {
    acc0[c] += __shfl_xor(acc0[c], 32, 64);
    acc1[c] += __shfl_xor(acc1[c], 32, 64);
  }
  __syncthreads();
  if (wv > 0 && h == 0) {
#pragma unroll
    for (int c = 0; c < MC; ++c) {
      Part[wv - 1][p * MC + c] = acc0[c];
      Part[wv - 1][(p + 32) * MC + c] = acc1[c];
    }
  }
  __syncthreads();
  if (wv == 0) {
    // lane l < 32 finishes row p, lane l >= 32 row p + 32 (both halves hold both sums)
    const int row_in = h ? p + 32 : p;
    if (row_in < bi) {
#pragma unroll
      for (int c = 0; c < MC; ++c) {
        if (c < mc) {
          const double own = h ? acc1[c] : acc0[c];
          const double sum = own + Part[0][row_in * MC + c] + Part[1][row_in * MC + c] + Part[2][row_in * MC + c];
          const int64_t row = i0 + row_in;
          if (nsplit == 1)
            R[row * ldr + c0 + c] = ABS ? sum : B[row * ldb + c0 + c] - sum;
          else
            P[((int64_t)sp * n + row) * mp + c0 + c] = sum;
        }
      }
    }
  }
}

// R[row][c] = B[row][c] - sum_sp P[sp][row][c] (ABS: the sum), splits in order.
template <bool ABS>
__global__ void symm_reduce_kernel(int n, int m, int nsplit, int mp, const double* __restrict__ P,
                                   const double* __restrict__ B, int64_t ldb, double* __restrict__ R,
                                   int64_t ldr) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * m) return;
  const int row = (int)(e / m), c = (int)(e % m);
  double sum = 0.0;
  for (int sp = 0; sp < nsplit; ++sp) sum += P[((int64_t)sp * n + row) * mp + c];
  R[(int64_t)row * ldr + c] = ABS ? sum : B[(int64_t)row * ldb + c] - sum;
}

// Splits of the symmetric walk: about 2048 workgroups, each with at least 4
// blocks, within the scratch given.
int symm_splits(int n, int m, int64_t scratch_elems) {
  const int nb = (n + kNB - 1) / kNB, nq = (m + kSymmMC - 1) / kSymmMC, mp = nq * kSymmMC;
  int sp = std::max(1, std::min(2048 / std::max(1, nb * nq), nb / 4));
  while (sp > 1 && (int64_t)sp * n * mp > scratch_elems) --sp;
  return sp;
}

template <bool ABS>
hipError_t launch_symm(int n, int m, const double* A, int64_t lda, const double* X, int64_t ldx, const double* B,
                       int64_t ldb, double* R, int64_t ldr, double* scratch, int64_t scratch_elems,
                       hipStream_t s) {
  const int nb = (n + kNB - 1) / kNB, nq = (m + kSymmMC - 1) / kSymmMC, mp = nq * kSymmMC;
  const int sp = symm_splits(n, m, scratch != nullptr ? scratch_elems : 0);
  const bool vec = reinterpret_cast<uintptr_t>(A) % 16 == 0 && lda % 2 == 0;
  const dim3 grid(nb, nq, sp);
  if (vec)
    hipLaunchKernelGGL((symm_walk_kernel<ABS, true>), grid, dim3(256), 0, s, n, m, A, lda, X, ldx, B, ldb, R, ldr,
                       scratch, sp, mp);
  else
    hipLaunchKernelGGL((symm_walk_kernel<ABS, false>), grid, dim3(256), 0, s, n, m, A, lda, X, ldx, B, ldb, R, ldr,
                       scratch, sp, mp);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || sp == 1) return e;
  const int64_t total = (int64_t)n * m;
  hipLaunchKernelGGL((symm_reduce_kernel<ABS>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, n, m, sp, mp,
                     scratch, B, ldb, R, ldr);
  return hipGetLastError();
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

int64_t symm_scratch_elems(int n, int m) {
  const int nq = (m + kSymmMC - 1) / kSymmMC;
  return (int64_t)symm_splits(n, m, INT64_MAX) * n * nq * kSymmMC;
}

hipError_t launch_symm_residual(int n, int m, const double* A, int64_t lda, const double* X, int64_t ldx,
                                const double* B, int64_t ldb, double* R, int64_t ldr, hipStream_t s,
                                double* scratch, int64_t scratch_elems) {
  if (n <= 0 || m <= 0) return hipSuccess;
  return launch_symm<false>(n, m, A, lda, X, ldx, B, ldb, R, ldr, scratch, scratch_elems, s);
}

hipError_t launch_symm_abs_rowsum(int n, const double* A, int64_t lda, double* out, hipStream_t s,
                                  const double* w, double* scratch, int64_t scratch_elems) {
  if (n <= 0) return hipSuccess;
  return launch_symm<true>(n, 1, A, lda, w, 1, nullptr, 0, out, 1, scratch, scratch_elems, s);
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

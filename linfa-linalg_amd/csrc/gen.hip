// Seeded synthetic workloads on the device (llh_dev_gen_*): the matrices the
// benchmarks and the full-size parity tests solve, drawn from counters so that
// the oracle's CPU generator reproduces every value bit for bit.
//
// Element (i, j) of a rows x cols matrix drawn with (seed, stream) is
// value(key, i * cols + j), key = SplitMix64 output 0 from the state
// seed + stream * 0x9E3779B97F4A7C15; draw c is SplitMix64 output c from the
// state key (Steele, Lea and Flood's SplitMix64: state += 0x9E3779B97F4A7C15,
// then two multiply-xorshift rounds).  Uniform values are 2 u - 1 with u the
// top 53 bits over 2^53; normal values follow Marsaglia's polar method
// (v1 = draw 2c, v2 = draw 2c + 1 with c = 64 idx + attempt; accept
// 0 < s = v1^2 + v2^2 < 1; return v1 sqrt(-2 ln(s) / s)), with ln(s) from an
// atanh series of correctly rounded operations only.
//
// The SPD matrix is A = B B^T / n + lam I: a_ij = (sum_k b_ik b_jk, k in order
// from +0, each product and sum rounded on its own) / n, plus lam on the
// diagonal; the upper triangle mirrors the lower.  This file is compiled with
// -ffp-contract=off: every multiply and add above stays a separate IEEE
// operation, as in the CPU restatement.
#pragma clang fp contract(off)

#include <algorithm>
#include <cstdio>

#include "llh.h"
#include "llh_context.h"
#include "llh_kernels.h"

namespace llh {
namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ __forceinline__ uint64_t key_of(uint64_t seed, uint64_t stream) {
  uint64_t z = seed + kGolden * stream + kGolden;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t u) { return (double)(u >> 11) * 0x1.0p-53; }

// ln(s), s positive and normal: s = m 2^e, m in [sqrt(1/2), sqrt(2)),
// ln(m) = 2 atanh(f), f = (m - 1) / (m + 1), series in w = f^2 to w^12.
__device__ double ln_series(double s) {
  uint64_t bits = (uint64_t)__double_as_longlong(s);
  int64_t e = (int64_t)((bits >> 52) & 0x7ff) - 1023;
  bits = (bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
  double m = __longlong_as_double((long long)bits);
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e += 1;
  }
  const double f = (m - 1.0) / (m + 1.0);
  const double w = f * f;
  double p = 1.0 / 25.0;
  p = p * w + 1.0 / 23.0;
  p = p * w + 1.0 / 21.0;
  p = p * w + 1.0 / 19.0;
  p = p * w + 1.0 / 17.0;
  p = p * w + 1.0 / 15.0;
  p = p * w + 1.0 / 13.0;
  p = p * w + 1.0 / 11.0;
  p = p * w + 1.0 / 9.0;
  p = p * w + 1.0 / 7.0;
  p = p * w + 1.0 / 5.0;
  p = p * w + 1.0 / 3.0;
  p = p * w + 1.0;
  const double lm = (2.0 * f) * p;
  return (double)e * 0.6931471805599453 + lm;
}

__device__ double gen_value(int dist, uint64_t key, uint64_t idx) {
  if (dist == LLH_GEN_UNIFORM) return 2.0 * u01(mix64(key + kGolden * (idx + 1))) - 1.0;
  for (uint64_t t = 0; t < 64; ++t) {
    const uint64_t c = 2 * (idx * 64 + t);
    const double v1 = 2.0 * u01(mix64(key + kGolden * (c + 1))) - 1.0;
    const double v2 = 2.0 * u01(mix64(key + kGolden * (c + 2))) - 1.0;
    const double s = v1 * v1 + v2 * v2;
    if (s < 1.0 && s != 0.0) {
      const double l = ln_series(s);
      const double q = (-2.0 * l) / s;
      return v1 * __builtin_sqrt(q);
    }
  }
  return 0.0;  // 64 rejections in a row: probability below 1e-42
}

__global__ __launch_bounds__(256) void gen_matrix_kernel(int dist, uint64_t key, int64_t rows, int64_t cols,
                                                         double* __restrict__ x, int64_t ldx) {
  const int64_t total = rows * cols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / cols, j = e % cols;
    x[i * ldx + j] = gen_value(dist, key, (uint64_t)e);
  }
}

// One 64 x 64 lower tile of A = B B^T / n + lam I per workgroup; thread (ty, tx)
// owns rows 4 ty .. 4 ty + 3 and columns 4 tx .. 4 tx + 3 of the tile.  Each
// entry's sum runs over k in order.
constexpr int kGT = 64, kGK = 32;
__global__ __launch_bounds__(256) void gen_spd_kernel(int64_t n, const double* __restrict__ B, double lam,
                                                      double* __restrict__ A, int64_t lda) {
  __shared__ double Bi[kGK][kGT + 1];
  __shared__ double Bj[kGK][kGT + 1];
  // lower tiles enumerated row by row: tile t -> (ti, tj), tj <= ti
  const int64_t t = blockIdx.x;
  int64_t ti = (int64_t)((__builtin_sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while (ti * (ti + 1) / 2 > t) --ti;
  while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
  const int64_t tj = t - ti * (ti + 1) / 2;
  const int64_t i0 = ti * kGT, j0 = tj * kGT;
  const int tid = threadIdx.x, ty = tid / 16, tx = tid % 16;
  double acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = 0.0;
  for (int64_t k0 = 0; k0 < n; k0 += kGK) {
    __syncthreads();
#pragma unroll
    for (int g = 0; g < kGT * kGK / 256; ++g) {
      const int e = tid + 256 * g, row = e / kGK, k = e % kGK;
      const int64_t gi = i0 + row, gj = j0 + row, gk = k0 + k;
      Bi[k][row] = (gi < n && gk < n) ? B[gi * n + gk] : 0.0;
      Bj[k][row] = (gj < n && gk < n) ? B[gj * n + gk] : 0.0;
    }
    __syncthreads();
    const int kc = (int)min<int64_t>(kGK, n - k0);
    for (int k = 0; k < kc; ++k) {
      double bi[4], bj[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) bi[r] = Bi[k][4 * ty + r];
#pragma unroll
      for (int c = 0; c < 4; ++c) bj[c] = Bj[k][4 * tx + c];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = acc[r][c] + bi[r] * bj[c];
    }
  }
  const double dn = (double)n;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int64_t i = i0 + 4 * ty + r, j = j0 + 4 * tx + c;
      if (i < n && j <= i) {
        double v = acc[r][c] / dn;
        if (i == j) v = v + lam;
        A[i * lda + j] = v;
        A[j * lda + i] = v;
      }
    }
}

}  // namespace
}  // namespace llh

using namespace llh;

extern "C" llh_status llh_dev_gen_matrix_d(llh_context* ctx, int dist, uint64_t seed, uint64_t stream, size_t rows,
                                           size_t cols, double* x, size_t ldx, llh_error* err) {
  if (ctx == nullptr || ctx->magic != llh_context::kMagic) return invalid(err, "null or destroyed context");
  if (dist != LLH_GEN_NORMAL && dist != LLH_GEN_UNIFORM) return invalid(err, "unknown distribution");
  if (rows == 0 || cols == 0) return ok(err);
  if (x == nullptr || ldx < cols) return invalid(err, "bad matrix arguments");
  const int64_t total = (int64_t)(rows * cols);
  const int grid = (int)std::min<int64_t>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(gen_matrix_kernel, dim3(grid), dim3(256), 0, ctx->stream, dist, key_of(seed, stream),
                     (int64_t)rows, (int64_t)cols, x, (int64_t)ldx);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(err, e, "generator");
  return ok(err);
}

extern "C" llh_status llh_dev_gen_spd_d(llh_context* ctx, int dist, uint64_t seed, size_t n, double lam, double* a,
                                        size_t lda, double* b_work, llh_error* err) {
  if (ctx == nullptr || ctx->magic != llh_context::kMagic) return invalid(err, "null or destroyed context");
  if (n == 0) return ok(err);
  if (a == nullptr || b_work == nullptr || lda < n || n > 0x7fffffff) return invalid(err, "bad matrix arguments");
  llh_status st = llh_dev_gen_matrix_d(ctx, dist, seed, 0, n, n, b_work, n, err);
  if (st != LLH_OK) return st;
  const int64_t nt = ((int64_t)n + kGT - 1) / kGT;
  hipLaunchKernelGGL(gen_spd_kernel, dim3((unsigned)(nt * (nt + 1) / 2)), dim3(256), 0, ctx->stream, (int64_t)n,
                     b_work, lam, a, (int64_t)lda);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(err, e, "generator");
  return ok(err);
}

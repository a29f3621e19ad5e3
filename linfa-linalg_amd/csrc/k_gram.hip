// C (ka x kb) = A^T B for tall row-major A (n x ka) and B (n x kb): the Gram
// products of LOBPCG's orthonormalize and apply_constraints (V^T V, Y^T V).
//
// n is large and ka, kb are a block size, so the reduction runs over rows: the
// rows are cut into chunks, and each workgroup computes one 64 x 64 output tile
// over one chunk with MFMA, its four waves a 32 x 32 quadrant each.  MFMA step
// s takes rows r .. r+3: the A operand of lane l is A[r + l/16][a0 + l%16] and
// the B operand B[r + l/16][b0 + l%16], so both are read straight from global
// memory as 16 consecutive columns of 4 rows (no LDS).  Each workgroup writes its
// tile to its own partial slot; a second kernel sums the slots in chunk order,
// so the result does not depend on scheduling.
#include "llh_common.h"
#include "k_gemm_impl.h"
#include "llh_kernels.h"

namespace llh {
namespace {

constexpr int kGramTile = 64;

template <typename T>
__global__ __launch_bounds__(256) void gram_partial_kernel(int n, int ka, int kb, const T* __restrict__ A,
                                                           int64_t lda, const T* __restrict__ B, int64_t ldb,
                                                           int rows_per_chunk, T* __restrict__ partial) {
  using MF = Mfma<T>;
  using acc_t = typename MF::acc_t;
  const int chunk = blockIdx.x, ta = blockIdx.y, tb = blockIdx.z;
  const int r0 = chunk * rows_per_chunk, r1 = min(n, r0 + rows_per_chunk);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int qa = ta * kGramTile + (wave >> 1) * 32, qb = tb * kGramTile + (wave & 1) * 32;
  const int m = lane & 15, kk = lane >> 4;
  // clamped columns: products of out-of-range columns land in outputs that the
  // reduction never reads
  int ca[2], cb[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    ca[i] = min(qa + 16 * i + m, ka - 1);
    cb[i] = min(qb + 16 * i + m, kb - 1);
  }
  acc_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
  constexpr int U = 4;  // MFMA steps per iteration (16 rows), loads issued together
  for (int r = r0; r < r1; r += 4 * U) {
    T av[U][2], bv[U][2];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int row = r + 4 * u + kk;
      const bool ok = row < r1;
      const int64_t rr = ok ? row : r0;
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const T x = A[rr * lda + ca[i]];
        const T y = B[rr * ldb + cb[i]];
        av[u][i] = ok ? x : T(0);
        bv[u][i] = ok ? y : T(0);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = MF::mma(av[u][i], bv[u][j], acc[i][j]);
  }
  T* P = partial + ((int64_t)(chunk * gridDim.y + ta) * gridDim.z + tb) * kGramTile * kGramTile;
  const int la = (wave >> 1) * 32, lb = (wave & 1) * 32;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        P[(la + 16 * i + MF::row(lane, e)) * kGramTile + lb + 16 * j + m] = acc[i][j][e];
}

// C[a][b] = sum over chunks, in chunk order, of the partial tiles.
template <typename T>
__global__ void gram_reduce_kernel(int chunks, int tiles_a, int tiles_b, int ka, int kb,
                                   const T* __restrict__ partial, T* __restrict__ C, int64_t ldc) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)ka * kb) return;
  const int a = (int)(e / kb), b = (int)(e % kb);
  const int ta = a / kGramTile, tb = b / kGramTile, off = (a % kGramTile) * kGramTile + b % kGramTile;
  const int64_t stride = (int64_t)tiles_a * tiles_b * kGramTile * kGramTile;
  const T* p = partial + ((int64_t)ta * tiles_b + tb) * kGramTile * kGramTile + off;
  T s = T(0);
  for (int c = 0; c < chunks; ++c) s += p[c * stride];
  C[(int64_t)a * ldc + b] = s;
}

}  // namespace

int64_t gram_partial_elems(int n, int ka, int kb, int* chunks, int* rows_per_chunk) {
  const int tiles_a = (ka + kGramTile - 1) / kGramTile, tiles_b = (kb + kGramTile - 1) / kGramTile;
  // about 2048 workgroups, at least 256 rows (16 MFMA iterations) each
  const int want = std::max(1, 2048 / (tiles_a * tiles_b));
  int rpc = (n + want - 1) / want;
  rpc = std::max(256, (rpc + 15) / 16 * 16);
  *rows_per_chunk = rpc;
  *chunks = std::max(1, (n + rpc - 1) / rpc);
  return (int64_t)*chunks * tiles_a * tiles_b * kGramTile * kGramTile;
}

template <typename T>
hipError_t launch_gram(int n, int ka, int kb, const T* A, int64_t lda, const T* B, int64_t ldb, T* C, int64_t ldc,
                       T* partial, hipStream_t s) {
  if (ka <= 0 || kb <= 0) return hipSuccess;
  int chunks = 0, rpc = 0;
  gram_partial_elems(n, ka, kb, &chunks, &rpc);
  const int tiles_a = (ka + kGramTile - 1) / kGramTile, tiles_b = (kb + kGramTile - 1) / kGramTile;
  if (n > 0)
    hipLaunchKernelGGL((gram_partial_kernel<T>), dim3(chunks, tiles_a, tiles_b), dim3(256), 0, s, n, ka, kb, A, lda,
                       B, ldb, rpc, partial);
  else
    chunks = 0;  // an empty product: the reduction writes zeros
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int64_t total = (int64_t)ka * kb;
  hipLaunchKernelGGL((gram_reduce_kernel<T>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, chunks, tiles_a,
                     tiles_b, ka, kb, partial, C, ldc);
  return hipGetLastError();
}

template hipError_t launch_gram<double>(int, int, int, const double*, int64_t, const double*, int64_t, double*,
                                        int64_t, double*, hipStream_t);
template hipError_t launch_gram<float>(int, int, int, const float*, int64_t, const float*, int64_t, float*, int64_t,
                                       float*, hipStream_t);

}  // namespace llh

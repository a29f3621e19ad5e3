// Fast factor + inverse of one 64 x 64 diagonal block for the blocked paths
// (the trailing updates there round differently from the reference anyway, so
// the block may too; small whole matrices keep the exact-order kernel).
//
// The block is staged in LDS.  It is factored in eight panels of 8 columns:
// wave 0 factors a panel with each lane holding its row's 8 panel values in
// registers, the pivot and the column values broadcast by v_readlane (no
// barrier inside a panel); then all four waves apply the panel's rank-8 update
// to the rest of the block.  The inverse of the factor is then formed in LDS by
// the recursive doubling of trtri64_block.  Pivot test, failure row and
// write-back follow cholesky.rs:54-74: on a failure at row f the block holds
// rows < f and L[f, ..f].
#include "llh_common.h"
#include "k_trsm_impl.h"
#include "llh_kernels.h"

namespace llh {
namespace {

__device__ __forceinline__ double bcast(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float bcast(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// TS: storage type of the block and of the inverse; T: arithmetic type.  Sm,
// Xs, Ts: LDS arrays of 64 x 65 elements.
template <typename TS, typename T>
__device__ void diag_fast_block(TS* __restrict__ a, int64_t lda, int b, TS* __restrict__ inv,
                                int32_t* __restrict__ info, int64_t info_base, T* Sm, T* Xs, T* Ts,
                                int* s_fail_p) {
  constexpr int N = kNB, S = kNB + 1, P = 8;
  int& s_fail = *s_fail_p;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // stage the lower triangle (identity beyond b)
  {
    TS v[16];
#pragma unroll
    for (int g = 0; g < 16; ++g) {
      const int e = tid + 256 * g, i = e / N, j = e % N;
      v[g] = (i < b && j <= i) ? a[(int64_t)i * lda + j] : TS(0);
    }
#pragma unroll
    for (int g = 0; g < 16; ++g) {
      const int e = tid + 256 * g, i = e / N, j = e % N;
      Sm[i * S + j] = (i < b) ? T(v[g]) : (i == j ? T(1) : T(0));
    }
  }
  if (tid == 0) s_fail = -1;
  __syncthreads();
  int fail = -1;
  for (int c0 = 0; c0 < b; c0 += P) {
    if (wave == 0) {
      T v[P];
#pragma unroll
      for (int c = 0; c < P; ++c) v[c] = Sm[lane * S + c0 + c];
      int f = -1;
#pragma unroll
      for (int c = 0; c < P; ++c) {
        const int col = c0 + c;
        if (col < b && f < 0) {
          const T d = bcast(v[c], col);
          if (d <= T(0)) {
            f = col;  // cholesky.rs:69 (NaN passes)
          } else {
            const T r = sqrt_rn(d);
            v[c] = lane > col ? v[c] / r : (lane == col ? r : v[c]);
#pragma unroll
            for (int q = c + 1; q < P; ++q) {
              const T lq = bcast(v[c], c0 + q);
              if (lane >= c0 + q) v[q] -= v[c] * lq;
            }
          }
        }
      }
      // write the panel: every finished column, and on failure the columns
      // before the failing one (which hold row f's L[f, ..f])
#pragma unroll
      for (int c = 0; c < P; ++c)
        if (lane >= c0 && (f < 0 || c0 + c < f)) Sm[lane * S + c0 + c] = v[c];
      if (lane == 0) s_fail = f;
    }
    __syncthreads();
    fail = s_fail;
    if (fail >= 0) break;
    // rank-8 update of the rest: thread (row j = tid / 4, quarter q) updates
    // S[j][k] for k = c0 + P + q + 4t <= j.
    const int j = tid >> 2, q = tid & 3, k1 = c0 + P;
    if (j >= k1 && j < b) {
      T lj[P];
#pragma unroll
      for (int p = 0; p < P; ++p) lj[p] = Sm[j * S + c0 + p];
      for (int k = k1 + q; k <= j; k += 4) {
        T s = Sm[j * S + k];
#pragma unroll
        for (int p = 0; p < P; ++p) s -= lj[p] * Sm[k * S + c0 + p];
        Sm[j * S + k] = s;
      }
    }
    __syncthreads();
  }
  // write back: rows < f (or all rows) with their lower part, row f up to f-1
  {
    const int rows = fail < 0 ? b : fail + 1;
    for (int e = tid; e < N * N; e += 256) {
      const int i = e / N, j = e % N;
      if (i < rows && (i == fail ? j < i : j <= i)) a[(int64_t)i * lda + j] = TS(Sm[i * S + j]);
    }
  }
  if (fail >= 0) {
    if (tid == 0) atomicMin(info, (int32_t)(info_base + fail));
    return;
  }
  // inverse of the factor held in Sm (read through a generic pointer, ld S;
  // Sm doubles as the inverse's copy of the block)
  __syncthreads();
  trtri64_block<T, TS, T>(b, Sm, S, false, false, inv, Sm, Xs, Ts);
}

template <typename TS, typename T>
__global__ __launch_bounds__(256) void diag_fast_kernel(TS* __restrict__ a, int64_t lda, int b,
                                                        TS* __restrict__ inv, int32_t* __restrict__ info,
                                                        int64_t info_base, const int32_t* __restrict__ gate) {
  if (gate != nullptr && *gate != kNoFail) return;
  __shared__ T Sm[kNB * (kNB + 1)];  // the block, then (aliased) the inverse's working copy
  __shared__ T Xs[kNB * (kNB + 1)];
  __shared__ T Ts[kNB * (kNB + 1)];
  __shared__ int s_fail;
  diag_fast_block<TS, T>(a, lda, b, inv, info, info_base, Sm, Xs, Ts, &s_fail);
}

// Grouped form for the batch planner: work w = (item, k0) factors and inverts
// the block at (k0, k0) of its system; a failure sets the system's status.
template <typename T>
__global__ __launch_bounds__(256) void grp_diag_fast_kernel(const BatchBase<T>* __restrict__ base,
                                                            const BatchItem* __restrict__ items,
                                                            const StepWork* __restrict__ work, int count) {
  if ((int)blockIdx.x >= count) return;
  const StepWork w = work[blockIdx.x];
  const BatchItem it = items[w.item];
  const BatchBase<T> bb = *base;
  int32_t* st = bb.status + it.id;
  if (*st != kNoFail) return;
  __shared__ T Sm[kNB * (kNB + 1)];
  __shared__ T Xs[kNB * (kNB + 1)];
  __shared__ T Ts[kNB * (kNB + 1)];
  __shared__ int s_fail;
  diag_fast_block<T, T>(bb.a + it.a_off + (int64_t)w.k0 * it.lda + w.k0, it.lda, min(kNB, it.n - w.k0),
                        bb.linv + it.linv_off + (int64_t)(w.k0 / kNB) * kNB * kNB, st, w.k0, Sm, Xs, Ts,
                        &s_fail);
}

}  // namespace

template <typename T>
hipError_t launch_grp_diag_fast(const BatchBase<T>* base, const BatchItem* items, const StepWork* work,
                                int count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL((grp_diag_fast_kernel<T>), dim3(count), dim3(256), 0, s, base, items, work, count);
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
  hipLaunchKernelGGL((diag_fast_kernel<TS, T>), dim3(1), dim3(256), 0, s, a, lda, b, inv, info, info_base, gate);
  return hipGetLastError();
}

template hipError_t launch_diag_fast<double, double>(double*, int64_t, int, double*, int32_t*, int64_t,
                                                     const int32_t*, hipStream_t);
template hipError_t launch_diag_fast<float, float>(float*, int64_t, int, float*, int32_t*, int64_t,
                                                   const int32_t*, hipStream_t);
template hipError_t launch_diag_fast<float, double>(float*, int64_t, int, float*, int32_t*, int64_t,
                                                    const int32_t*, hipStream_t);

}  // namespace llh

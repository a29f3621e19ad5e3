// Fast factor + inverse of one 64 x 64 diagonal block for the blocked paths
// (the trailing updates there round differently from the reference anyway, so
// the block may too; small whole matrices keep the exact-order kernel).
//
// One wave does it all, in panels of 8 columns.  Lane i holds row i of the
// block in registers.  Inside a panel the column chain uses registers only:
// the pivot and each L[k, j] come from their lanes by v_readlane.  The panel is
// then written to a row-major copy of L in LDS, and the rank-8 update of the
// columns to its right reads L[k, j0 .. j0+8) as 128-bit uniform-address
// (broadcast) loads, software-pipelined behind the FMAs (rank2_rows).  The inverse runs the same way with lane c holding column c
// of inv(L): inside a panel x_k is scaled by 1/L[k, k] and eliminated from the
// panel's later rows; then the rows below take the panel's rank-8 update.
// Each lane loads and stores its own row, all its loads issued together.  Rows
// beyond b are identity rows, so the block factors to diag(L, I) and its
// inverse likewise.  Pivot test, failure row and write-back follow
// cholesky.rs:54-74: on a failure at row f the block holds rows < f and
// L[f, ..f].
#include <type_traits>
#include <utility>

#include "llh_common.h"
#include "llh_kernels.h"

namespace llh {
namespace {

// Calls f(std::integral_constant<int, I>) for I = 0 .. n-1: the panel loops are
// expanded at compile time, so every index into the register arrays is a
// constant (a rolled loop would move them to scratch memory).
template <typename F, int... I>
__device__ __forceinline__ void static_for(std::integer_sequence<int, I...>, F&& f) {
  (f(std::integral_constant<int, I>{}), ...);
}

__device__ __forceinline__ double bcast(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float bcast(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

constexpr int kDiagS = kNB + 2;  // LDS row stride (keeps rows 16-byte aligned)
constexpr int kDiagP = 8;        // panel width

// Applies f(y[k], (col[k*S], col[k*S+1])) to the rows k = K0 .. 63 of the
// register array y, the column pair read from LDS as one 128-bit uniform load
// per row.  The rows go in chunks of 8 with the loads of the next chunk issued
// before the FMAs of the current one (and a scheduling barrier per chunk to
// keep that order): left alone, the compiler keeps about 3 loads in flight and
// each row pays most of an LDS round trip.
template <typename T, int K0, typename F>
__device__ __forceinline__ void rank2_rows(T* y, const T* col, F&& f) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  constexpr int NC = (kNB - K0) / 8;
  static_assert((kNB - K0) % 8 == 0, "rows in chunks of 8");
  if constexpr (NC > 0) {
    v2 buf[2][8];
#pragma unroll
    for (int u = 0; u < 8; ++u) buf[0][u] = *reinterpret_cast<const v2*>(&col[(K0 + u) * kDiagS]);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (c + 1 < NC) {
#pragma unroll
        for (int u = 0; u < 8; ++u)
          buf[(c + 1) & 1][u] = *reinterpret_cast<const v2*>(&col[(K0 + 8 * (c + 1) + u) * kDiagS]);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) f(y[K0 + 8 * c + u], buf[c & 1][u]);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

template <typename T>
struct DiagLds {
  T l[kNB * kDiagS];  // L (row-major), written a panel at a time
  T rinv[kNB];        // 1 / L[j, j]
};

// TS: storage type of the block and of the inverse; T: arithmetic type.
// Launched with exactly one wave.
template <typename TS, typename T>
__device__ __forceinline__ void diag_wave_block(TS* __restrict__ a, int64_t lda, int b, TS* __restrict__ inv,
                                                int32_t* __restrict__ info, int64_t info_base, DiagLds<T>& sm) {
  constexpr int N = kNB, S = kDiagS, P = kDiagP;
  typedef T v2 __attribute__((ext_vector_type(2)));
  const int i = threadIdx.x;  // this lane's row (factor) / column (loads, stores, inverse)
  T* l = sm.l;  // row-major L, published a panel at a time
  // lane i loads its own row; masked entries read the block's first element
  // (always in bounds) so that the 64 loads carry no branches and issue together
  T r[N];
  {
    const TS* row = a + (int64_t)min(i, b - 1) * lda;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const bool in = i < b && j <= i;
      const T v = T(*(in ? row + j : a));
      r[j] = in ? v : ((i >= b && i == j) ? T(1) : T(0));
    }
  }
  // (no early exit: a failure only turns the remaining steps off, so every loop
  // keeps its constant trip count and r stays in registers)
  int fail = -1;
  static_for(std::make_integer_sequence<int, N / P>{}, [&](auto pi) {
    constexpr int j0 = decltype(pi)::value * P;
#pragma unroll
    for (int c = 0; c < P; ++c) {
      const int j = j0 + c;
      if (fail < 0) {
        const T d = bcast(r[j], j);
        if (d <= T(0)) {  // cholesky.rs:69 (NaN passes)
          fail = j;
        } else {
          const T piv = sqrt_rn(d);
          const T rp = T(1) / piv;
          r[j] = (i == j) ? piv : r[j] * rp;
          if (i == 0) sm.rinv[j] = rp;
#pragma unroll
          for (int q = c + 1; q < P; ++q) r[j0 + q] -= r[j] * bcast(r[j], j0 + q);
        }
      }
    }
    if (fail < 0) {
      // publish the panel (this lane's row), then the rank-8 update to its right
#pragma unroll
      for (int q = 0; q < P; q += 2) *reinterpret_cast<v2*>(&l[i * S + j0 + q]) = v2{r[j0 + q], r[j0 + q + 1]};
      __builtin_amdgcn_wave_barrier();
      // (panel column pairs outermost: the pair stays in registers while the
      // rows stream past, which keeps the register count bounded)
#pragma unroll
      for (int q = 0; q < P; q += 2) {
        const T p0 = r[j0 + q], p1 = r[j0 + q + 1];
        rank2_rows<T, j0 + P>(r, l + j0 + q, [&](T& y, const v2& lk) { y = fma_t(-p0, lk[0], fma_t(-p1, lk[1], y)); });
      }
    }
  });
  // write back from registers: rows < f (or all rows) with their lower part,
  // row f up to f-1
  {
    const int last = (fail < 0 || i < fail) ? i : (i == fail ? i - 1 : -1);
    TS* row = a + (int64_t)i * lda;
#pragma unroll
    for (int j = 0; j < N; ++j)
      if (i < b && j <= last) row[j] = TS(r[j]);
  }
  if (fail >= 0) {
    if (i == 0) atomicMin(info, (int32_t)(info_base + fail));
    return;
  }
  // inv(L): lane c solves L x = e_c in panels of 8
  const int c = i;
  T x[N];
#pragma unroll
  for (int k = 0; k < N; ++k) x[k] = (k == c) ? T(1) : T(0);
  static_for(std::make_integer_sequence<int, N / P>{}, [&](auto pi) {
    constexpr int k0 = decltype(pi)::value * P;
#pragma unroll
    for (int q = 0; q < P; ++q) {
      const int k = k0 + q;
      x[k] *= sm.rinv[k];
#pragma unroll
      for (int q2 = q + 1; q2 < P; ++q2) x[k0 + q2] -= l[(k0 + q2) * S + k] * x[k];
    }
#pragma unroll
    for (int q = 0; q < P; q += 2) {
      const T p0 = x[k0 + q], p1 = x[k0 + q + 1];
      rank2_rows<T, k0 + P>(x, l + k0 + q, [&](T& y, const v2& li) { y = fma_t(-li[0], p0, fma_t(-li[1], p1, y)); });
    }
  });
#pragma unroll
  for (int ii = 0; ii < N; ++ii) inv[ii * N + c] = TS(x[ii]);
}

template <typename TS, typename T>
__global__ __launch_bounds__(64) void diag_fast_kernel(TS* __restrict__ a, int64_t lda, int b,
                                                       TS* __restrict__ inv, int32_t* __restrict__ info,
                                                       int64_t info_base, const int32_t* __restrict__ gate) {
  if (gate != nullptr && *gate != kNoFail) return;
  __shared__ __attribute__((aligned(16))) DiagLds<T> sm;
  diag_wave_block<TS, T>(a, lda, b, inv, info, info_base, sm);
}

// Grouped form for the batch planner: work w = (item, k0) factors and inverts
// the block at (k0, k0) of its system; a failure sets the system's status.
template <typename T>
__global__ __launch_bounds__(64) void grp_diag_fast_kernel(const BatchBase<T>* __restrict__ base,
                                                           const BatchItem* __restrict__ items,
                                                           const StepWork* __restrict__ work, int count) {
  if ((int)blockIdx.x >= count) return;
  const StepWork w = work[blockIdx.x];
  const BatchItem it = items[w.item];
  const BatchBase<T> bb = *base;
  int32_t* st = bb.status + it.id;
  if (*st != kNoFail) return;
  __shared__ __attribute__((aligned(16))) DiagLds<T> sm;
  diag_wave_block<T, T>(bb.a + it.a_off + (int64_t)w.k0 * it.lda + w.k0, it.lda, min(kNB, it.n - w.k0),
                        bb.linv + it.linv_off + (int64_t)(w.k0 / kNB) * kNB * kNB, st, w.k0, sm);
}

}  // namespace

template <typename T>
hipError_t launch_grp_diag_fast(const BatchBase<T>* base, const BatchItem* items, const StepWork* work,
                                int count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL((grp_diag_fast_kernel<T>), dim3(count), dim3(64), 0, s, base, items, work, count);
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
  hipLaunchKernelGGL((diag_fast_kernel<TS, T>), dim3(1), dim3(64), 0, s, a, lda, b, inv, info, info_base, gate);
  return hipGetLastError();
}

template hipError_t launch_diag_fast<double, double>(double*, int64_t, int, double*, int32_t*, int64_t,
                                                     const int32_t*, hipStream_t);
template hipError_t launch_diag_fast<float, float>(float*, int64_t, int, float*, int32_t*, int64_t,
                                                   const int32_t*, hipStream_t);
template hipError_t launch_diag_fast<float, double>(float*, int64_t, int, float*, int32_t*, int64_t,
                                                    const int32_t*, hipStream_t);

}  // namespace llh

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

#include "llh_common.h"
#include "llh_kernels.h"

namespace llh {
namespace {

// Packed column-major lower triangle: entry (j, k), j >= k.
__device__ __forceinline__ int pk(int n, int j, int k) { return k * n - ((k * (k - 1)) >> 1) + (j - k); }

// Factors the n x n matrix at `a` (row-major, lda) in LDS: Ap receives the
// lower triangle of A, W the pending sums and then L (both packed column-major).
// Returns -1 on success or the first row j whose pivot fails `d <= 0`.  On
// return W holds every final column (all columns on success, columns < fail
// otherwise) and Dg the square roots of the passed pivots.
template <typename TS, typename T, int NT, int NMAX>
__device__ int factor_exact(const TS* __restrict__ a, int64_t lda, int n, T* W, T* Ap, T* Dg) {
  constexpr int KG = 8;  // columns whose loads are issued together
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  constexpr int NW = NT / 64;
  const int tot = (n * (n + 1)) >> 1;
  for (int e = tid; e < tot; e += NT) W[e] = T(0);
  for (int j = wave; j < n; j += NW)
    for (int k = lane; k <= j; k += 64) Ap[pk(n, j, k)] = T(a[(int64_t)j * lda + k]);
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    // Every thread evaluates the pivot identically, so no broadcast is needed.
    const T d = Ap[pk(n, k, k)] - W[pk(n, k, k)];
    if (d <= T(0)) return k;  // cholesky.rs:69; NaN passes
    const T piv = sqrt_rn(d);
    if (tid == 0) Dg[k] = piv;
    for (int j = k + 1 + tid; j < n; j += NT) {
      const int o = pk(n, j, k);
      W[o] = (Ap[o] - W[o]) / piv;  // cholesky.rs:62
    }
    __syncthreads();
    // Pending sums: S[j,k'] += L[k',k] * L[j,k] for k < k' <= j (k' == j is d_j).
    // Thread (lane, wave) owns rows j = lane + 64 r and columns k' = k+1+wave+NW i;
    // the loads of a group of KG columns are issued together (they never alias
    // the column-k values read), then the group is updated.
    const int colk = pk(n, k, k) - k;  // W[colk + i] == L[i, k]
#pragma unroll
    for (int r = 0; r < (NMAX + 63) / 64; ++r) {
      const int j = lane + 64 * r;
      if (j > k && j < n) {
        const T ljk = W[colk + j];
        for (int kp0 = k + 1 + wave; kp0 <= j; kp0 += KG * NW) {
          T lk[KG], sv[KG];
#pragma unroll
          for (int g = 0; g < KG; ++g) {
            const int kp = kp0 + g * NW;
            const bool on = kp <= j;
            lk[g] = on ? W[colk + kp] : T(0);
            sv[g] = on ? W[pk(n, j, kp)] : T(0);
          }
#pragma unroll
          for (int g = 0; g < KG; ++g) {
            const int kp = kp0 + g * NW;
            if (kp <= j) W[pk(n, j, kp)] = sv[g] + lk[g] * ljk;
          }
        }
      }
    }
    __syncthreads();
  }
  return -1;
}

// Writes the factor back into `a` (lower triangle only; the strict upper
// triangle is never touched, which is the dirty form).  On failure at row f it
// writes rows < f and L[f, 0..f-1], exactly what the reference has written when
// it returns NotPositiveDefinite.
template <typename TS, typename T, int NT>
__device__ void write_back(TS* __restrict__ a, int64_t lda, int n, int fail, const T* W, const T* Dg) {
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  constexpr int NW = NT / 64;
  const int rows = fail < 0 ? n : fail + 1;
  for (int j = wave; j < rows; j += NW) {
    const int kend = (j == fail) ? j : j + 1;
    for (int k = lane; k < kend; k += 64) a[(int64_t)j * lda + k] = TS((k == j) ? Dg[j] : W[pk(n, j, k)]);
  }
}

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
  const SmallDesc<T> ds = descs[bi];
  const int n = ds.n;
  if (n > NMAX) return;  // the host never routes a larger matrix here
  __shared__ T W[(NMAX * (NMAX + 1)) / 2];
  __shared__ T Ap[(NMAX * (NMAX + 1)) / 2];
  __shared__ T Dg[NMAX];
  __shared__ T Bs[NMAX * MCH];
  __shared__ T Xs[NMAX * MCH];
  const int fail = factor_exact<T, T, NT, NMAX>(ds.a, ds.lda, n, W, Ap, Dg);
  __syncthreads();
  write_back<T, T, NT>(ds.a, ds.lda, n, fail, W, Dg);
  if (threadIdx.x == 0) ds.info[0] = fail < 0 ? kNoFail : fail;
  if (fail >= 0 || !do_solve) return;
  const int tid = threadIdx.x;
  // Solve in chunks of MCH right-hand-side columns; b is row-major n x m.
  for (int c0 = 0; c0 < ds.m; c0 += MCH) {
    const int mc = min(MCH, ds.m - c0);
    const int tot = n * mc;
    __syncthreads();
    for (int e = tid; e < tot; e += NT) {
      const int r = e / mc, c = e - r * mc;
      Bs[r * MCH + c] = ds.b[(int64_t)r * ds.ldb + c0 + c];
    }
    __syncthreads();
    // L y = b (Lower: rows 0..n, a[r, i] = L[r, i] for r > i), triangular.rs:142.
    for (int i = 0; i < n; ++i) {
      for (int e = tid; e < (n - i) * mc; e += NT) {
        const int rr = e / mc, c = e - rr * mc, r = i + rr;
        const T coeff = Bs[i * MCH + c] / Dg[i];
        if (r == i) Xs[i * MCH + c] = coeff;
        else Bs[r * MCH + c] = Bs[r * MCH + c] + ((-coeff) * W[pk(n, r, i)]);
      }
      __syncthreads();
    }
    // L^T x = y (Upper on L.t(): rows n-1..0, a[r, i] = L[i, r] for r < i),
    // cholesky.rs:142 with triangular.rs:140.
    for (int i = n - 1; i >= 0; --i) {
      for (int e = tid; e < (i + 1) * mc; e += NT) {
        const int r = e / mc, c = e - r * mc;
        const T coeff = Xs[i * MCH + c] / Dg[i];
        if (r == i) Bs[i * MCH + c] = coeff;
        else Xs[r * MCH + c] = Xs[r * MCH + c] + ((-coeff) * W[pk(n, i, r)]);
      }
      __syncthreads();
    }
    for (int e = tid; e < tot; e += NT) {
      const int r = e / mc, c = e - r * mc;
      ds.b[(int64_t)r * ds.ldb + c0 + c] = Bs[r * MCH + c];
    }
  }
}

}  // namespace

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

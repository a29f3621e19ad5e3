// Device code of the exact-order small-matrix factor and solve, shared by the
// single-matrix kernels (k_small.hip) and the batched kernels (k_batch.hip).
// Files that include it are compiled with -ffp-contract=off and
// `#pragma clang fp contract(off)`: every `s += x*y` must stay a separately
// rounded multiply and add.
#pragma once

#include "llh_common.h"
#include "llh_kernels.h"

namespace llh {

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
  (void)lane;
  for (int e = tid; e < tot; e += NT) W[e] = T(0);
  // Stage the lower triangle of A: groups of LG loads are issued before any of
  // them is stored, so their latencies overlap.
  {
    constexpr int LG = 16;
    const int sq = n * n;
    for (int e0 = 0; e0 < sq; e0 += NT * LG) {
      TS v[LG];
#pragma unroll
      for (int g = 0; g < LG; ++g) {
        const int e = e0 + g * NT + tid, j = e / n, k = e - j * n;
        v[g] = (e < sq && k <= j) ? a[(int64_t)j * lda + k] : TS(0);
      }
#pragma unroll
      for (int g = 0; g < LG; ++g) {
        const int e = e0 + g * NT + tid, j = e / n, k = e - j * n;
        if (e < sq && k <= j) Ap[pk(n, j, k)] = T(v[g]);
      }
    }
  }
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

// LDS of one small factor + solve.
template <typename T, int NMAX, int MCH>
struct SmallLds {
  T W[(NMAX * (NMAX + 1)) / 2];
  T Ap[(NMAX * (NMAX + 1)) / 2];
  T Dg[NMAX];
  T Bs[NMAX * MCH];
  T Xs[NMAX * MCH];
};

// Factors ds.a (dirty, in place) and, when do_solve, solves ds.b in place in the
// reference's operation order; ds.info receives kNoFail or the failing row.
// do_solve == 2: ds.a already holds the factor (a solve against it, as the
// refinement steps of a batched mixed-precision solve need); ds.info is read:
// a matrix whose factor failed is skipped.
template <typename T, int NMAX, int NT, int MCH>
__device__ void small_solve_one(const SmallDesc<T>& ds, int do_solve, SmallLds<T, NMAX, MCH>& lds) {
  const int n = ds.n;
  if (n > NMAX) return;  // the host never routes a larger matrix here
  T* W = lds.W;
  T* Ap = lds.Ap;
  T* Dg = lds.Dg;
  T* Bs = lds.Bs;
  T* Xs = lds.Xs;
  const int tid = threadIdx.x;
  if (do_solve == 2) {
    if (ds.info[0] != kNoFail) return;
    for (int e = tid; e < n * n; e += NT) {
      const int j = e / n, k = e - j * n;
      if (k < j) W[pk(n, j, k)] = ds.a[(int64_t)j * ds.lda + k];
      else if (k == j) Dg[j] = ds.a[(int64_t)j * ds.lda + j];
    }
  } else {
    const int fail = factor_exact<T, T, NT, NMAX>(ds.a, ds.lda, n, W, Ap, Dg);
    __syncthreads();
    write_back<T, T, NT>(ds.a, ds.lda, n, fail, W, Dg);
    if (threadIdx.x == 0) ds.info[0] = fail < 0 ? kNoFail : fail;
    if (fail >= 0 || !do_solve) return;
  }
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

}  // namespace llh

// Device code shared by the batched substitutions: the 64 x 64 by 64 x 8 tile
// product on MFMA, and the forward-substitution steps that the grouped
// factor's diagonal and panel launches carry when a plan gives them a
// workspace (BatchBase::y).  Both forms run the same products in the same
// order, so a plan's results do not depend on where the forward sweep runs.
#pragma once

#include "k_gemm_impl.h"
#include "llh_common.h"
#include "llh_kernels.h"

namespace llh {

// Rows [16w, 16w + 16) of A Y for a 64 x 64 A in ts (LDS, row stride 65) and a
// 64 x kSolveCols Y in ys (LDS, padded to 16 columns with zeros), w =
// threadIdx.x / 64.  Result register q holds row 16w + Mfma<T>::row(lane, q),
// column lane % 16.  The sum runs over k = 4s + lane / 16, s = 0 .. 15, the
// order of the substitution kernel's own tile products (solve_system), so the
// forward steps carried by the factor give its bits.
template <typename T>
__device__ __forceinline__ typename Mfma<T>::acc_t subst_tile(const T* __restrict__ ts, const T* __restrict__ ys) {
  using MF = Mfma<T>;
  constexpr int S = kNB + 1, MC = kSolveCols;
  const int lane = threadIdx.x & 63, w16 = (threadIdx.x >> 6) * 16, fr = lane & 15, fk = lane >> 4;
  typename MF::acc_t acc = {0, 0, 0, 0};
#pragma unroll
  for (int s = 0; s < kNB / 4; ++s) {
    const int k = 4 * s + fk;
    acc = MF::mma(ts[(w16 + fr) * S + k], fr < MC ? ys[k * MC + fr] : T(0), acc);
  }
  return acc;
}

// Stages the 64 x 64 tile at src (rows ld apart; rows >= rows and columns >=
// cols read as zero) into ts (row stride kNB + 1), as the substitution kernel
// stages its tiles: all sixteen loads of a thread are issued before the first
// store.
template <typename T>
__device__ __forceinline__ void subst_stage(const T* __restrict__ src, int64_t ld, int rows, int cols,
                                            T* __restrict__ ts) {
  T pre[kNB * kNB / 256];
#pragma unroll
  for (int g = 0; g < kNB * kNB / 256; ++g) {
    const int e = threadIdx.x + 256 * g, i = e >> 6, j = e & 63;
    pre[g] = src[(int64_t)min(i, rows - 1) * ld + min(j, cols - 1)];
  }
#pragma unroll
  for (int g = 0; g < kNB * kNB / 256; ++g) {
    const int e = threadIdx.x + 256 * g, i = e >> 6, j = e & 63;
    ts[i * (kNB + 1) + j] = (i < rows && j < cols) ? pre[g] : T(0);
  }
}

// Right-hand-side rows the forward steps read: Y (B's layout) from the first
// block's steps on; those steps read B itself and write Y, so Y needs no copy.
template <typename T>
__device__ __forceinline__ const T* fwd_src(const BatchBase<T>& bb, const BatchItem& it, int k0) {
  return (k0 == 0 ? bb.b : bb.y) + it.b_off;
}

// Y_k (rows k0 .., the first kSolveCols columns) loaded before the block's
// factor, so the load overlaps it: element e = threadIdx.x + 256 g is row
// e / 8, column e % 8, from clamped indices (masked when staged).
template <typename T>
struct FwdDiagPrefetch {
  T yk[kNB * kSolveCols / 256];
};
template <typename T>
__device__ __forceinline__ FwdDiagPrefetch<T> fwd_diag_prefetch(const BatchBase<T>& bb, const BatchItem& it,
                                                                int k0) {
  constexpr int MC = kSolveCols;
  const int n = it.n, m = it.m, b = min(kNB, n - k0), mc = min(MC, m);
  const T* src = fwd_src(bb, it, k0);
  FwdDiagPrefetch<T> f;
#pragma unroll
  for (int g = 0; g < kNB * MC / 256; ++g) {
    const int e = threadIdx.x + 256 * g, k = e / MC, c = e % MC;
    f.yk[g] = mc > 0 ? src[(int64_t)(k0 + min(k, b - 1)) * m + min(c, mc - 1)] : T(0);
  }
  return f;
}

// Forward-substitution step of the diagonal block at k0 of system `it` in the
// grouped factor: Y_k = inv(L_kk) Y_k.  Called by all 256 threads once the
// block's inverse is stored; ts (64 x 65) and ys (64 x kSolveCols) are LDS the
// caller no longer needs, f what fwd_diag_prefetch loaded.  (The substitution
// kernel's forward sweep does the same per block.)
template <typename T>
__device__ void grp_forward_diag(const BatchBase<T>& bb, const BatchItem& it, int k0, T* __restrict__ ts,
                                 T* __restrict__ ys, const FwdDiagPrefetch<T>& f) {
  using MF = Mfma<T>;
  constexpr int MC = kSolveCols;
  const int tid = threadIdx.x, n = it.n, m = it.m, b = min(kNB, n - k0);
  const int lane = tid & 63, w16 = (tid >> 6) * 16, fr = lane & 15;
  const T* src = fwd_src(bb, it, k0) + (int64_t)k0 * m;
  T* yk = bb.y + it.b_off + (int64_t)k0 * m;
  subst_stage<T>(bb.linv + it.linv_off + (int64_t)(k0 / kNB) * kNB * kNB, kNB, kNB, kNB, ts);
  for (int c0 = 0; c0 < m; c0 += MC) {
    const int mc = min(MC, m - c0);
#pragma unroll
    for (int g = 0; g < kNB * MC / 256; ++g) {
      const int e = tid + 256 * g, k = e / MC, c = e % MC;
      ys[e] = (k < b && c < mc) ? (c0 == 0 ? f.yk[g] : src[(int64_t)k * m + c0 + c]) : T(0);
    }
    __syncthreads();
    const auto prod = subst_tile<T>(ts, ys);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int row = w16 + MF::row(lane, q);
      if (row < b && fr < mc) yk[(int64_t)row * m + c0 + fr] = prod[q];
    }
    __syncthreads();  // ys is restaged for the next columns
  }
}

// Forward-substitution update of a panel tile of the grouped factor: Y rows
// [i0, i0 + 64) -= L_ik Y_k with i0 = k0 + 64 + m0.  The caller has just
// stored L_ik (all 64 columns: a block with rows below is full) and left it in
// ts (row stride 65, zero below the system's last row); ys is free LDS.  f
// holds what fwd_panel_prefetch loaded for the first kSolveCols columns before
// the tile's product, so those loads overlap it.
template <typename T>
struct FwdPrefetch {
  T yk[kNB * kSolveCols / 256];  // Y_k, element e = threadIdx.x + 256 g: row e / 8, column e % 8
  T yi[4];                       // the rows this tile updates, at the rows and column of this lane's products
};
template <typename T>
__device__ __forceinline__ FwdPrefetch<T> fwd_panel_prefetch(const BatchBase<T>& bb, const BatchItem& it, int k0,
                                                             int m0) {
  using MF = Mfma<T>;
  constexpr int MC = kSolveCols;
  const int tid = threadIdx.x, lane = tid & 63, w16 = (tid >> 6) * 16, fr = lane & 15;
  const int n = it.n, m = it.m, i0 = k0 + kNB + m0, rows = min(kNB, n - i0), mc = min(MC, m);
  const T* y = bb.y + it.b_off;
  const T* src = fwd_src(bb, it, k0);
  FwdPrefetch<T> f;
#pragma unroll
  for (int g = 0; g < kNB * MC / 256; ++g) {
    const int e = tid + 256 * g, k = e / MC, c = e % MC;
    f.yk[g] = mc > 0 ? y[(int64_t)(k0 + k) * m + min(c, mc - 1)] : T(0);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int row = min(w16 + MF::row(lane, q), rows - 1);
    f.yi[q] = mc > 0 ? src[(int64_t)(i0 + row) * m + min(fr, mc - 1)] : T(0);
  }
  return f;
}
template <typename T>
__device__ void grp_forward_panel(const BatchBase<T>& bb, const BatchItem& it, int k0, int m0, const T* ts,
                                  T* __restrict__ ys, const FwdPrefetch<T>& f) {
  using MF = Mfma<T>;
  constexpr int MC = kSolveCols;
  const int tid = threadIdx.x, lane = tid & 63, w16 = (tid >> 6) * 16, fr = lane & 15;
  const int n = it.n, m = it.m, i0 = k0 + kNB + m0, rows = min(kNB, n - i0);
  T* y = bb.y + it.b_off;
  const T* src = fwd_src(bb, it, k0);
  for (int c0 = 0; c0 < m; c0 += MC) {
    const int mc = min(MC, m - c0);
    const T* yk = y + (int64_t)k0 * m + c0;
#pragma unroll
    for (int g = 0; g < kNB * MC / 256; ++g) {
      const int e = tid + 256 * g, k = e / MC, c = e % MC;
      ys[e] = c < mc ? (c0 == 0 ? f.yk[g] : yk[(int64_t)k * m + c]) : T(0);
    }
    __syncthreads();
    const auto prod = subst_tile<T>(ts, ys);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int row = w16 + MF::row(lane, q);
      if (row < rows && fr < mc) {
        const int64_t at = (int64_t)(i0 + row) * m + c0 + fr;
        y[at] = (c0 == 0 ? f.yi[q] : src[at]) - prod[q];
      }
    }
    __syncthreads();  // ys is restaged for the next columns
  }
}

}  // namespace llh

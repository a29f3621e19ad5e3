// Grouped kernels of the batched blocked solve.  One launch does one step of
// the blocked algorithm for every matrix of the batch that is still active at
// that step: each workgroup reads its (matrix, block[, tile]) from a work list
// the planner built once, so a whole batch of uneven sizes runs as a fixed
// sequence of launches (captured in one hipGraph).  A matrix whose status word
// records a failed pivot is skipped by every later launch.
#include <algorithm>
#include <climits>

#include "llh_common.h"
#include "k_diag_impl.h"
#include "k_gemm_impl.h"
#include "k_subst_impl.h"
#include "k_trsm_impl.h"
#include "llh_kernels.h"

namespace llh {
namespace {

template <typename T>
__device__ __forceinline__ bool failed(const BatchBase<T>& bb, const BatchItem& it) {
  return bb.status[it.id] != kNoFail;
}

// Work item of workgroup b in a launch of `count` workgroups, chosen so that
// each XCD takes a contiguous run of the work list: workgroups are dealt to the
// 8 XCDs round-robin (b to XCD b % 8), and XCD x gets items
// [x q + min(x, r), ...) with q = count / 8, r = count % 8 -- a bijection on
// [0, count).  The tiles of one system sit together in the list, so they run
// on one XCD and read their panel blocks from its L2, not from HBM.
__device__ __forceinline__ int xcd_item(int b, int count) {
  const int x = b & 7, q = count >> 3, r = count & 7;
  return x * q + min(x, r) + (b >> 3);
}

// inv(L_kk) of the block at (k0, k0) into the matrix's inverse slot k0 / 64.
template <typename T>
__global__ __launch_bounds__(256) void grp_trtri_kernel(const BatchBase<T>* __restrict__ base,
                                                        const BatchItem* __restrict__ items,
                                                        const StepWork* __restrict__ work, int count) {
  if ((int)blockIdx.x >= count) return;
  const StepWork w = work[blockIdx.x];
  const BatchItem it = items[w.item];
  const BatchBase<T> bb = *base;
  if (failed(bb, it)) return;
  __shared__ T Ms[kNB * (kNB + 1)];
  __shared__ T Xs[kNB * (kNB + 1)];
  __shared__ T Ts[kNB * (kNB + 1)];
  const T* a = bb.a + it.a_off + (int64_t)w.k0 * it.lda + w.k0;
  trtri64_block<T, T, T>(min(kNB, it.n - w.k0), a, it.lda, false, false,
                      bb.linv + it.linv_off + (int64_t)(w.k0 / kNB) * kNB * kNB, Ms, Xs, Ts);
}

// Panel tile: L21[m0.., :] = A21[m0.., :] * inv(L11)^T, in place (the tile holds
// all b <= 64 columns of its rows, so the rows it reads are the rows it writes);
// with a forward workspace (bb.y), then Y of its rows -= L21 Y_k.
template <typename T>
__global__ __launch_bounds__(256) void grp_panel_kernel(const BatchBase<T>* __restrict__ base,
                                                        const BatchItem* __restrict__ items,
                                                        const TileWork* __restrict__ work, int count) {
  if ((int)blockIdx.x >= count) return;
  const TileWork w = work[xcd_item(blockIdx.x, count)];
  const BatchItem it = items[w.item];
  const BatchBase<T> bb = *base;
  if (failed(bb, it)) return;
  const int b = min(kNB, it.n - w.k0), below = it.n - w.k0 - b;
  T* p = bb.a + it.a_off + (int64_t)(w.k0 + b) * it.lda + w.k0;
  const T* li = bb.linv + it.linv_off + (int64_t)(w.k0 / kNB) * kNB * kNB;
  // the product's slabs, then (forward step) the tile and Y_k: 37 KB in f64,
  // four workgroups per CU either way
  __shared__ __attribute__((aligned(16))) union {
    struct {
      NtSlab<T, 64> as[2], bs[2];
    } t;
    struct {
      T ts[kNB * (kNB + 1)];
      T ys[kNB * kSolveCols];
    } f;
  } sm;
  if (bb.y == nullptr) {
    nt_tile_in<T, 64, 64, 2, 2, false, true>(below, b, b, T(1), p, it.lda, li, kNB, T(0), p, it.lda, w.m0, 0,
                                             sm.t.as, sm.t.bs);
    return;
  }
  const FwdPrefetch<T> f = fwd_panel_prefetch<T>(bb, it, w.k0, w.m0);
  nt_tile_in<T, 64, 64, 2, 2, false, true, false, T, true>(below, b, b, T(1), p, it.lda, li, kNB, T(0), p, it.lda,
                                                          w.m0, 0, sm.t.as, sm.t.bs, sm.f.ts);
  __syncthreads();  // the tile is in ts
  grp_forward_panel<T>(bb, it, w.k0, w.m0, sm.f.ts, sm.f.ys, f);
}

// C prefetch in the trailing tiles (f64): with k = 64 a tile has four slabs,
// and the C tile's loads otherwise wait until after its MFMAs.  Not for f32,
// where folding C into the sum early costs the refinement accuracy.
template <typename T>
constexpr bool kGrpSyrkPre = sizeof(T) == 8;

// Trailing tile: A22[m0.., n0..] -= L21 L21^T (lower part), L21 the kb columns
// from k0 (kb = 64, or 128 for the pair of blocks of one step).
template <typename T>
__global__ __launch_bounds__(256) void grp_syrk_kernel(const BatchBase<T>* __restrict__ base,
                                                       const BatchItem* __restrict__ items,
                                                       const TileWork* __restrict__ work, int count) {
  if ((int)blockIdx.x >= count) return;
  const TileWork w = work[xcd_item(blockIdx.x, count)];
  const BatchItem it = items[w.item];
  const BatchBase<T> bb = *base;
  if (failed(bb, it)) return;
  const int b = w.kb, m2 = it.n - w.k0 - b;
  T* p = bb.a + it.a_off + (int64_t)(w.k0 + b) * it.lda + w.k0;
  nt_tile<T, 64, 64, 2, 2, true, true, kGrpSyrkPre<T>>(m2, m2, b, T(-1), p, it.lda, p, it.lda, T(1), p + b, it.lda,
                                                        w.m0, w.n0);
}

// Forward-solve step of block k0: Y = inv(L_kk) B_k, then B_rest -= L21 Y.
// B columns are processed 8 at a time; thread (r, c) owns row r of the block.
template <typename T>
__global__ __launch_bounds__(256) void grp_fwd_kernel(const BatchBase<T>* __restrict__ base,
                                                      const BatchItem* __restrict__ items,
                                                      const StepWork* __restrict__ work, int count) {
  constexpr int MC = 8, S = kNB + 1;
  if ((int)blockIdx.x >= count) return;
  const StepWork w = work[blockIdx.x];
  const BatchItem it = items[w.item];
  const BatchBase<T> bb = *base;
  if (failed(bb, it)) return;
  __shared__ T Li[kNB * S];
  __shared__ T Ys[kNB * MC];
  const int tid = threadIdx.x, n = it.n, k0 = w.k0, b = min(kNB, n - k0);
  const T* li = bb.linv + it.linv_off + (int64_t)(k0 / kNB) * kNB * kNB;
  const T* a = bb.a + it.a_off;
  T* bm = bb.b + it.b_off;
  for (int e = tid; e < kNB * kNB; e += 256) Li[(e / kNB) * S + e % kNB] = li[e];
  for (int c0 = 0; c0 < it.m; c0 += MC) {
    const int mc = min(MC, it.m - c0);
    __syncthreads();
    for (int e = tid; e < kNB * MC; e += 256) {
      const int r = e / MC, c = e % MC;
      Ys[e] = (r < b && c < mc) ? bm[(int64_t)(k0 + r) * it.m + c0 + c] : T(0);
    }
    __syncthreads();
    // Y = Li * Bk (thread: row r, columns c, c + 4)
    const int r = tid >> 2, c = tid & 3;
    T y0 = T(0), y1 = T(0);
    for (int k = 0; k <= r && k < b; ++k) {
      const T l = Li[r * S + k];
      y0 += l * Ys[k * MC + c];
      y1 += l * Ys[k * MC + c + 4];
    }
    __syncthreads();
    Ys[r * MC + c] = y0;
    Ys[r * MC + c + 4] = y1;
    if (r < b) {
      if (c < mc) bm[(int64_t)(k0 + r) * it.m + c0 + c] = y0;
      if (c + 4 < mc) bm[(int64_t)(k0 + r) * it.m + c0 + c + 4] = y1;
    }
    __syncthreads();
    // B[k0+b+i, :] -= L[k0+b+i, k0..k0+b] * Y
    const int below = n - k0 - b;
    for (int e = tid; e < below * MC; e += 256) {
      const int i = e / MC, cc = e % MC;
      if (cc >= mc) continue;
      const T* lr = a + (int64_t)(k0 + b + i) * it.lda + k0;
      T s = T(0);
      for (int k = 0; k < b; ++k) s += lr[k] * Ys[k * MC + cc];
      bm[(int64_t)(k0 + b + i) * it.m + c0 + cc] -= s;
    }
  }
}

// Back-solve step of block k0 (op = L^T): X = inv(L_kk)^T Y_k, then
// B[0:k0] -= L[k0:k0+b, 0:k0]^T X.
template <typename T>
__global__ __launch_bounds__(256) void grp_bwd_kernel(const BatchBase<T>* __restrict__ base,
                                                      const BatchItem* __restrict__ items,
                                                      const StepWork* __restrict__ work, int count) {
  constexpr int MC = 8, S = kNB + 1;
  if ((int)blockIdx.x >= count) return;
  const StepWork w = work[blockIdx.x];
  const BatchItem it = items[w.item];
  const BatchBase<T> bb = *base;
  if (failed(bb, it)) return;
  __shared__ T Li[kNB * S];
  __shared__ T Ys[kNB * MC];
  const int tid = threadIdx.x, k0 = w.k0, b = min(kNB, it.n - k0);
  const T* li = bb.linv + it.linv_off + (int64_t)(k0 / kNB) * kNB * kNB;
  const T* a = bb.a + it.a_off;
  T* bm = bb.b + it.b_off;
  for (int e = tid; e < kNB * kNB; e += 256) Li[(e / kNB) * S + e % kNB] = li[e];
  for (int c0 = 0; c0 < it.m; c0 += MC) {
    const int mc = min(MC, it.m - c0);
    __syncthreads();
    for (int e = tid; e < kNB * MC; e += 256) {
      const int r = e / MC, c = e % MC;
      Ys[e] = (r < b && c < mc) ? bm[(int64_t)(k0 + r) * it.m + c0 + c] : T(0);
    }
    __syncthreads();
    // X = Li^T * Yk: X[r] = sum_{k >= r} Li[k][r] Y[k]
    const int r = tid >> 2, c = tid & 3;
    T x0 = T(0), x1 = T(0);
    for (int k = r; k < b; ++k) {
      const T l = Li[k * S + r];
      x0 += l * Ys[k * MC + c];
      x1 += l * Ys[k * MC + c + 4];
    }
    __syncthreads();
    Ys[r * MC + c] = x0;
    Ys[r * MC + c + 4] = x1;
    if (r < b) {
      if (c < mc) bm[(int64_t)(k0 + r) * it.m + c0 + c] = x0;
      if (c + 4 < mc) bm[(int64_t)(k0 + r) * it.m + c0 + c + 4] = x1;
    }
    __syncthreads();
    // B[i, :] -= sum_k L[k0+k][i] X[k] for i < k0
    for (int e = tid; e < k0 * MC; e += 256) {
      const int i = e / MC, cc = e % MC;
      if (cc >= mc) continue;
      T s = T(0);
      for (int k = 0; k < b; ++k) s += a[(int64_t)(k0 + k) * it.lda + i] * Ys[k * MC + cc];
      bm[(int64_t)i * it.m + c0 + cc] -= s;
    }
  }
}

// Forward and back substitution of one system per workgroup, all of it in one
// launch (n <= kSolveLdsMaxN): the system's right-hand sides (8 columns at a
// time) stay in LDS from the first block to the last, and L streams through a
// 64 x 64 LDS tile that is loaded with whole rows per wave (coalesced) while the
// previous tile is being used (register prefetch).
//   forward, block k:  Y_k = inv(L_kk) B_k, then B_i -= L_ik Y_k for the
//                      64-row chunks i below (L read along its rows);
//   backward, block k: X_k = inv(L_kk)^T Y_k, then B_i -= L_ki^T X_k for the
//                      64-column chunks i left of k (L's block row k, read
//                      along its rows, the transpose taken in LDS).
// The 64 x 64 by 64 x 8 tile products run on MFMA (tile_mul).  When the
// grouped factor's steps carried the forward sweep (a plan's workspace bb.y),
// only the back sweep runs here, starting from Y.
// solve_system runs it for system `it` (B's rows ldb apart) with the caller's
// LDS: Bs holds kSolveLdsMaxN x kSolveCols values, Ts a 64 x 65 tile.
template <typename T>
__device__ __forceinline__ void solve_system(const BatchBase<T>& bb, const BatchItem& it, T* __restrict__ Bs,
                                             T* __restrict__ Ts, int64_t ldb) {
  constexpr int MC = kSolveCols, S = kNB + 1;
  const int tid = threadIdx.x, n = it.n, m = it.m, nb = (n + kNB - 1) / kNB;
  const int64_t lda = it.lda;
  const T* a = bb.a + it.a_off;
  const T* linv = bb.linv + it.linv_off;
  T* bm = bb.b + it.b_off;
  T pre[kNB * kNB / 256];
  // The tile products on MFMA: wave w forms rows [16w, 16w + 16) of op(Ts) Y,
  // Y the 64 x 8 right-hand-side block at y (LDS, padded to 16 columns with
  // zeros), op(Ts) = Ts or its transpose; result register q holds row
  // 16w + MF::row(lane, q), column lane % 16.  (One thread per output pair with
  // scalar FMAs read three LDS values per two FMAs: 393 KB of LDS reads per
  // tile against 64 KB here.)
  using MF = Mfma<T>;
  using acc_t = typename MF::acc_t;
  const int lane = tid & 63, w16 = (tid >> 6) * 16, fr = lane & 15, fk = lane >> 4;
  // (subst_tile's product, with the transposed form: the same k order, so the
  // forward steps that the grouped factor carries give the same bits)
  auto tile_mul = [&](const T* y, bool trans) -> acc_t {
    acc_t acc = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < kNB / 4; ++s) {
      const int k = 4 * s + fk;
      const T av = trans ? Ts[k * S + w16 + fr] : Ts[(w16 + fr) * S + k];
      const T bv = fr < MC ? y[k * MC + fr] : T(0);
      acc = MF::mma(av, bv, acc);
    }
    return acc;
  };
  // a tile: rows x cols of src (ld), zero outside; element e = tid + 256 g is
  // (row e / 64, column e % 64), so each wave reads one row segment
  auto fetch = [&](const T* src, int64_t ld, int rows, int cols) {
#pragma unroll
    for (int g = 0; g < kNB * kNB / 256; ++g) {
      const int e = tid + 256 * g, i = e >> 6, j = e & 63;
      pre[g] = (i < rows && j < cols) ? src[(int64_t)i * ld + j] : T(0);
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int g = 0; g < kNB * kNB / 256; ++g) {
      const int e = tid + 256 * g;
      Ts[(e >> 6) * S + (e & 63)] = pre[g];
    }
  };
  for (int c0 = 0; c0 < m; c0 += MC) {
    const int mc = min(MC, m - c0);
    __syncthreads();
    // with a forward workspace the factor's steps have done the forward sweep:
    // start from Y (B's layout, rows m apart)
    const T* src = bb.y != nullptr ? bb.y + it.b_off : bm;
    const int64_t ld_src = bb.y != nullptr ? (int64_t)m : ldb;
    for (int e = tid; e < nb * kNB * MC; e += 256) {
      const int row = e / MC, cc = e % MC;
      Bs[e] = (row < n && cc < mc) ? src[(int64_t)row * ld_src + c0 + cc] : T(0);
    }
    // forward: tile (kb, i0), i0 < 0 the inverse of block kb, else rows i0.. below it
    if (bb.y == nullptr) {
      auto src = [&](int kb, int i0) {
        const int k0 = kb * kNB;
        if (i0 < 0) fetch(linv + (int64_t)kb * kNB * kNB, kNB, kNB, kNB);
        else fetch(a + (int64_t)i0 * lda + k0, lda, min(kNB, n - i0), min(kNB, n - k0));
      };
      auto advance = [&](int& kb, int& i0) -> bool {
        const int k0 = kb * kNB, nx = i0 < 0 ? k0 + min(kNB, n - k0) : i0 + kNB;
        if (nx < n) {
          i0 = nx;
          return true;
        }
        if (kb + 1 < nb) {
          kb += 1;
          i0 = -1;
          return true;
        }
        return false;
      };
      int kb = 0, i0 = -1;
      src(kb, i0);
      for (;;) {
        __syncthreads();
        stash();
        __syncthreads();
        int nkb = kb, ni0 = i0;
        const bool more = advance(nkb, ni0);
        if (more) src(nkb, ni0);
        const int k0 = kb * kNB, b = min(kNB, n - k0);
        const acc_t prod = tile_mul(Bs + (int64_t)k0 * MC, false);
        if (i0 < 0) {
          __syncthreads();  // every wave has read Y_k before it is overwritten
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int row = w16 + MF::row(lane, q);
            if (row < b && fr < MC) Bs[(k0 + row) * MC + fr] = prod[q];
          }
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int row = w16 + MF::row(lane, q);
            if (i0 + row < n && fr < MC) Bs[(i0 + row) * MC + fr] -= prod[q];
          }
        }
        if (!more) break;
        kb = nkb;
        i0 = ni0;
      }
    }
    // backward: tile (kb, i0), i0 < 0 the inverse of block kb, else columns
    // i0 .. i0 + 63 of L's block row kb
    {
      auto src = [&](int kb, int i0) {
        const int k0 = kb * kNB;
        if (i0 < 0) fetch(linv + (int64_t)kb * kNB * kNB, kNB, kNB, kNB);
        else fetch(a + (int64_t)k0 * lda + i0, lda, min(kNB, n - k0), kNB);
      };
      auto advance = [&](int& kb, int& i0) -> bool {
        const int k0 = kb * kNB, nx = i0 < 0 ? 0 : i0 + kNB;
        if (nx < k0) {
          i0 = nx;
          return true;
        }
        if (kb > 0) {
          kb -= 1;
          i0 = -1;
          return true;
        }
        return false;
      };
      int kb = nb - 1, i0 = -1;
      src(kb, i0);
      for (;;) {
        __syncthreads();
        stash();
        __syncthreads();
        int nkb = kb, ni0 = i0;
        const bool more = advance(nkb, ni0);
        if (more) src(nkb, ni0);
        const int k0 = kb * kNB, b = min(kNB, n - k0);
        const acc_t prod = tile_mul(Bs + (int64_t)k0 * MC, true);
        if (i0 < 0) {
          __syncthreads();  // every wave has read X_k before it is overwritten
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int row = w16 + MF::row(lane, q);
            if (row < b && fr < MC) Bs[(k0 + row) * MC + fr] = prod[q];
          }
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int row = w16 + MF::row(lane, q);
            if (fr < MC) Bs[(i0 + row) * MC + fr] -= prod[q];
          }
        }
        if (!more) break;
        kb = nkb;
        i0 = ni0;
      }
    }
    __syncthreads();
    for (int e = tid; e < n * MC; e += 256) {
      const int row = e / MC, cc = e % MC;
      if (cc < mc) bm[(int64_t)row * ldb + c0 + cc] = Bs[e];
    }
  }
}

// A workgroup barrier that orders LDS accesses only: __syncthreads() also
// waits for every global load in flight (one counter covers loads and stores),
// which would drain the tile prefetch at every diagonal block.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The same two sweeps for one system with L streamed (solve_lds_kernel).  The
// 64 x 64 tiles come in a fixed order -- forward: per block k the inverse, then
// the row chunks below; backward: per block k the transposed inverse, then the
// column chunks to its left -- and no load depends on a result, so three tiles
// are in flight at any time.  Each wave works on its own 16 rows of a tile
// (16 rows of L, or 16 columns for a transposed tile): it loads them with
// consecutive lanes on consecutive addresses, stages them in its own slice of
// LDS and reads them back as lane (fr, fk) = (row fr, k values 16fk .. 16fk +
// 15), so no workgroup barrier sits between tiles; the MCT partial sums of a
// row are reduced over the four fk lanes by two xor shuffles.  Only the
// diagonal blocks synchronise the workgroup (all of B_k is read before Y_k
// replaces it).  VALU dot products: with MCT <= 8 columns an MFMA tile would be
// at least half padding.  (Loading the dot layout straight from global memory
// puts 64 cache lines behind every load instruction: 2.0k against ~0.7k cycles
// per tile.)
template <typename T>
constexpr int kStageElems = sizeof(T) == 8 ? 64 * 18 : 64 * 20;  // per wave

template <typename T, int MCT, bool VEC>
__device__ __forceinline__ void solve_stream(int n, int m, const T* __restrict__ a, int64_t lda,
                                             const T* __restrict__ linv, T* b, int64_t ldb, T* __restrict__ Bs,
                                             T* __restrict__ stage) {
  constexpr int KS = 16, VL = 16 / sizeof(T), NQ = KS / VL;
  constexpr int SR = kNB + VL, SC = KS + VL;  // staged row strides: 16 x 64 rows, 64 x 16 columns
  static_assert(16 * SR <= kStageElems<T> && kNB * SC <= kStageElems<T>, "a staged tile fits the slice");
  typedef T vec_t __attribute__((ext_vector_type(VL)));
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, fr = lane & 15, fk = lane >> 4;
  const int r = 16 * w + fr, kq = KS * fk;
  T* const ws = stage + w * kStageElems<T>;
  const int nb = (n + kNB - 1) / kNB, per = nb * (nb + 1) / 2, total = 2 * per;
  // tile t: block kb and chunk i (-1: the diagonal block's inverse)
  auto decode = [&](int t, int& kb, int& i) {
    if (t < per) {
      for (int k = 0;; ++k) {
        if (t < nb - k) {
          kb = k;
          i = t == 0 ? -1 : k + t;
          return;
        }
        t -= nb - k;
      }
    }
    t -= per;
    for (int k = nb - 1;; --k) {
      if (t < k + 1) {
        kb = k;
        i = t - 1;
        return;
      }
      t -= k + 1;
    }
  };
  // VEC: a, lda and linv keep VL-element groups 16-byte aligned (else scalar
  // loads; a run-time choice per load would put a branch join after every
  // load, where the compiler waits for all of them)
  auto load_vl = [&](const T* q) -> vec_t {
    if constexpr (VEC) {
      return *reinterpret_cast<const vec_t*>(q);
    } else {
      vec_t x;
#pragma unroll
      for (int e = 0; e < VL; ++e) x[e] = q[e];
      return x;
    }
  };
  // Global loads only, no masking (a value that depends on a lane's bounds
  // test makes the compiler wait for the load where it is issued).  Rows
  // beyond n are read from row n - 1 (finite values of L's lower triangle) and
  // come to nothing: they meet rows of B that stay zero, or their sums are not
  // written.  A tile is either the wave's 16 rows of 64 columns (rows) or all
  // 64 rows of the wave's 16 columns: lane positions for both, chosen per tile
  // by a select, so that every fetch is the same straight-line run of NQ loads
  // (with a branch per tile kind the compiler cannot tell which loads a later
  // use waits for, and waits for all of them).
  constexpr int LR = kNB / VL, LC = KS / VL;  // lanes per staged row: 64 or 16 columns
  const int row_r = 16 * w + lane / LR, col_r = VL * (lane % LR);
  const int row_c = lane / LC, col_c = 16 * w + VL * (lane % LC);
  auto fetch = [&](int t, vec_t* g) {
    int kb, i;
    decode(min(t, total - 1), kb, i);  // past the end: the last tile again (not used)
    const bool rows = t < per;
    const int k0 = kb * kNB;
    const T* p;
    int64_t ld;
    int rmax;
    if (i < 0) {
      p = linv + (int64_t)kb * kNB * kNB;
      ld = kNB;
      rmax = kNB - 1;
    } else if (rows) {
      p = a + (int64_t)(i * kNB) * lda + k0;
      ld = lda;
      rmax = n - 1 - i * kNB;
    } else {
      p = a + (int64_t)k0 * lda + i * kNB;
      ld = lda;
      rmax = n - 1 - k0;
    }
    const int row0 = rows ? row_r : row_c, dq = rows ? VL : kNB / LC, col = rows ? col_r : col_c;
#pragma unroll
    for (int q = 0; q < NQ; ++q) g[q] = load_vl(p + (int64_t)min(row0 + q * dq, rmax) * ld + col);
  };
  auto process = [&](int t, const vec_t* g) {
    int kb, i;
    decode(t, kb, i);
    const bool rows = t < per;
    // stage the wave's part of the tile, then read it as (row fr, k = kq ..)
    T v[KS];
    {
      const int base = rows ? (lane / LR) * SR + col_r : (lane / LC) * SC + (col_c - 16 * w);
      const int step = rows ? VL * SR : (kNB / LC) * SC;
#pragma unroll
      for (int q = 0; q < NQ; ++q) *reinterpret_cast<vec_t*>(ws + base + q * step) = g[q];
    }
    if (rows) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const vec_t x = *reinterpret_cast<const vec_t*>(ws + fr * SR + kq + q * VL);
#pragma unroll
        for (int e = 0; e < VL; ++e) v[q * VL + e] = x[e];
      }
    } else {
#pragma unroll
      for (int s = 0; s < KS; ++s) v[s] = ws[(kq + s) * SC + fr];
    }
    const T* x = Bs + (kb * kNB + kq) * MCT;
    if (i < 0) lds_barrier();  // every wave's updates of B_k are done
    T acc[MCT];
#pragma unroll
    for (int c = 0; c < MCT; ++c) acc[c] = T(0);
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
      for (int c = 0; c < MCT; ++c) acc[c] = fma_t(v[s], x[s * MCT + c], acc[c]);
#pragma unroll
    for (int c = 0; c < MCT; ++c) {
      acc[c] += __shfl_xor(acc[c], 16);
      acc[c] += __shfl_xor(acc[c], 32);
    }
    if (i < 0) {
      lds_barrier();  // every wave has read B_k
#pragma unroll
      for (int c = 0; c < MCT; ++c)
        if ((c & 3) == fk && kb * kNB + r < n) Bs[(kb * kNB + r) * MCT + c] = acc[c];
      lds_barrier();  // Y_k (X_k) is complete
    } else {
#pragma unroll
      for (int c = 0; c < MCT; ++c)
        if ((c & 3) == fk && i * kNB + r < n) Bs[(i * kNB + r) * MCT + c] -= acc[c];
    }
  };
  for (int c0 = 0; c0 < m; c0 += MCT) {
    const int mc = min(MCT, m - c0);
    lds_barrier();  // the previous chunk's results are stored
    for (int e = tid; e < nb * kNB * MCT; e += 256) {
      const int row = e / MCT, cc = e % MCT;
      Bs[e] = (row < n && cc < mc) ? b[(int64_t)row * ldb + c0 + cc] : T(0);
    }
    // process t while t + 1 and t + 2 are in flight (the scheduling barriers
    // keep the compiler from sinking the loads to their use)
    vec_t g0[NQ], g1[NQ], g2[NQ];
    fetch(0, g0);
    fetch(1, g1);
    for (int t = 0; t < total; t += 3) {
      fetch(t + 2, g2);
      __builtin_amdgcn_sched_barrier(0);
      process(t, g0);
      if (t + 1 >= total) break;
      fetch(t + 3, g0);
      __builtin_amdgcn_sched_barrier(0);
      process(t + 1, g1);
      if (t + 2 >= total) break;
      fetch(t + 4, g1);
      __builtin_amdgcn_sched_barrier(0);
      process(t + 2, g2);
    }
    lds_barrier();
    for (int e = tid; e < n * MCT; e += 256) {
      const int row = e / MCT, cc = e % MCT;
      if (cc < mc) b[(int64_t)row * ldb + c0 + cc] = Bs[e];
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void grp_solve_kernel(const BatchBase<T>* __restrict__ base,
                                                        const BatchItem* __restrict__ items, int count) {
  if ((int)blockIdx.x >= count) return;
  const BatchItem it = items[blockIdx.x];
  const BatchBase<T> bb = *base;
  if (failed(bb, it)) return;
  __shared__ T Bs[kSolveLdsMaxN * kSolveCols];
  __shared__ T Ts[kNB * (kNB + 1)];
  solve_system<T>(bb, it, Bs, Ts, it.m);
}

// One system's substitutions (launch_solve_lds): one workgroup, MCT columns of
// B at a time.
template <typename T, int MCT, bool VEC>
__global__ __launch_bounds__(256) void solve_lds_kernel(int n, int m, const T* __restrict__ a, int64_t lda,
                                                        const T* __restrict__ linv, T* b, int64_t ldb,
                                                        const int32_t* __restrict__ gate) {
  if (gate != nullptr && *gate != kNoFail) return;
  __shared__ __attribute__((aligned(16))) T Bs[kSolveLdsMaxN * MCT];
  __shared__ __attribute__((aligned(16))) T stage[4 * kStageElems<T>];
  solve_stream<T, MCT, VEC>(n, m, a, lda, linv, b, ldb, Bs, stage);
}

// The phases of the fused per-system kernel run one after another, so they
// share one LDS buffer.
template <typename T>
union MedLds {
  DiagLds<T> diag;
  struct {
    NtSlab<T, 64> as[2];
    NtSlab<T, 64> bs[2];
  } tile;
  struct {
    T bs[kSolveLdsMaxN * kSolveCols];
    T ts[kNB * (kNB + 1)];
  } solve;
};

// The medium bucket fused per system (every n <= kSolveLdsMaxN): one workgroup
// takes systems from a work queue, largest first (the items are sorted by
// decreasing n), and runs each through its whole blocked factor -- per 64-column
// step the diagonal block with its inverse, the panel as one MFMA product with
// that inverse, the trailing update on MFMA over 64 x 64 lower tiles -- and,
// with SOLVE, its forward and back substitution.  It is the device code of the
// grouped launches, in the same order for each system, so the results are bit
// for bit those of the grouped schedule; the difference is the traversal: the
// system stays in the workgroup's L2 between its steps instead of every step
// streaming the whole batch through HBM.  A failed diagonal block ends the
// system (status holds the row, as in the grouped path).
template <typename T, bool SOLVE>
__global__ __launch_bounds__(256, 2) void med_kernel(const BatchBase<T>* __restrict__ base,
                                                  const BatchItem* __restrict__ items, int count, int counter) {
  __shared__ __attribute__((aligned(16))) MedLds<T> sm;
  __shared__ int s_next;
  const BatchBase<T> bb = *base;
  NtSlab<T, 64>* const as = sm.tile.as;
  NtSlab<T, 64>* const bs = sm.tile.bs;
  for (;;) {
    __syncthreads();  // the previous system's LDS use and the read of s_next are done
    if (threadIdx.x == 0) s_next = atomicAdd(bb.counters + counter, 1);
    __syncthreads();
    const int q = s_next;
    if (q >= count) return;
    const BatchItem it = items[q];
    const int n = it.n;
    const int64_t lda = it.lda;
    T* a = bb.a + it.a_off;
    bool ok = true;
    // the diagonal block at k0, then the panel below it, as one grouped step
    auto block = [&](int k0) -> bool {
      const int b = min(kNB, n - k0);
      T* li = bb.linv + it.linv_off + (int64_t)(k0 / kNB) * kNB * kNB;
      if (diag_quad_block<T, T>(a + (int64_t)k0 * lda + k0, lda, b, li, bb.status + it.id, k0, sm.diag))
        return false;  // uniform: every thread read the same failing column
      __syncthreads();  // L_kk and its inverse written; the LDS is free for the tiles
      const int below = n - k0 - b;
      T* p = a + (int64_t)(k0 + b) * lda + k0;
      for (int m0 = 0; m0 < below; m0 += 64)
        nt_tile_in<T, 64, 64, 2, 2, false, true>(below, b, b, T(1), p, lda, li, kNB, T(0), p, lda, m0, 0, as, bs);
      __syncthreads();  // the whole panel is written before an update reads it
      return true;
    };
    // trailing tiles of the rows below the panel [k0, k0 + kb), in the first
    // max_cols columns
    auto update = [&](int k0, int kb, int max_cols) {
      const int m2 = n - k0 - kb;
      T* p = a + (int64_t)(k0 + kb) * lda + k0;
      for (int m0 = 0; m0 < m2; m0 += 64)
        for (int n0 = 0; n0 <= m0 && n0 < max_cols; n0 += 64)
          nt_tile_in<T, 64, 64, 2, 2, true, true, kGrpSyrkPre<T>>(m2, m2, kb, T(-1), p, lda, p, lda, T(1), p + kb,
                                                                  lda, m0, n0, as, bs);
      __syncthreads();
    };
    // the grouped schedule's order (kBatchPanel)
    constexpr int nbp = kBatchPanel / kNB;
    for (int k0 = 0; k0 < n && ok; k0 += kBatchPanel)
      for (int j = 0; j < nbp; ++j) {
        const int kj = k0 + kNB * j;
        if (kj >= n) break;
        if (!(ok = block(kj))) break;
        if (kj + kNB >= n) break;
        if (j + 1 == nbp) {
          update(k0, kBatchPanel, INT_MAX);
        } else {
          const int r = lowbit(j + 1);
          update(k0 + kNB * (j + 1 - r), kNB * r, kNB * min(r, nbp - j - 1));
        }
      }
    if (SOLVE && ok) solve_system<T>(bb, it, sm.solve.bs, sm.solve.ts, it.m);
  }
}

// Elementwise steps of the mixed-precision batch before the factor (pointers
// read from the device base, so a replayed graph follows the caller's buffers).
__global__ void mixed_flat_kernel(const MixedBase* __restrict__ base, int op, int64_t elems) {
  const MixedBase mb = *base;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < elems;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (op == 0) {
      mb.a32[i] = (float)mb.a64[i];
    } else {
      const double v = mb.b64[i];
      mb.x32[i] = (float)v;
      mb.bk64[i] = v;
    }
  }
}

// b64 = x32 (add false) or b64 += x32 (add true) for rows [r0, r0 + 64) of one
// system; a system whose factor failed keeps the caller's B.
__global__ __launch_bounds__(256) void grp_update_kernel(const MixedBase* __restrict__ base,
                                                         const BatchItem* __restrict__ items,
                                                         const StepWork* __restrict__ work, int count, bool add) {
  if ((int)blockIdx.x >= count) return;
  const StepWork w = work[blockIdx.x];
  const BatchItem it = items[w.item];
  const MixedBase mb = *base;
  if (mb.status[it.id] != kNoFail) return;
  const int m = it.m, rows = min(64, it.n - w.k0);
  const int64_t o = it.b_off + (int64_t)w.k0 * m;
  for (int e = threadIdx.x; e < rows * m; e += 256) {
    const double v = (double)mb.x32[o + e];
    mb.b64[o + e] = add ? mb.b64[o + e] + v : v;
  }
}

// r = bk - A x for rows [r0, r0 + 64) of one system, in f64, from the lower
// triangle of A (A[i][k] for k > i is read as A[k][i]), stored as f32 into x32
// (the right-hand side of the correction solve).  Thread (row, q) takes
// the columns q, q + 4, ... of its row, two at a time.
__global__ __launch_bounds__(256) void grp_residual_kernel(const MixedBase* __restrict__ base,
                                                           const BatchItem* __restrict__ items,
                                                           const StepWork* __restrict__ work, int count) {
  if ((int)blockIdx.x >= count) return;
  const StepWork w = work[blockIdx.x];
  const BatchItem it = items[w.item];
  const MixedBase mb = *base;
  if (mb.status[it.id] != kNoFail) return;
  const int row = w.k0 + (threadIdx.x >> 2), q = threadIdx.x & 3;
  if (row >= it.n) return;
  const double* a = mb.a64 + it.a_off;
  const double* x = mb.b64 + it.b_off;
  const double* b = mb.bk64 + it.b_off;
  float* r = mb.x32 + it.b_off;
  const int m = it.m;
  for (int c0 = q; c0 < m; c0 += 8) {
    const int c1 = c0 + 4;
    double s0 = 0.0, s1 = 0.0;
    for (int k = 0; k < it.n; ++k) {
      const double av = k <= row ? a[(int64_t)row * it.lda + k] : a[(int64_t)k * it.lda + row];
      s0 += av * x[(int64_t)k * m + c0];
      if (c1 < m) s1 += av * x[(int64_t)k * m + c1];
    }
    r[(int64_t)row * m + c0] = (float)(b[(int64_t)row * m + c0] - s0);
    if (c1 < m) r[(int64_t)row * m + c1] = (float)(b[(int64_t)row * m + c1] - s1);
  }
}

__global__ void fill_status_kernel(int32_t* st, int64_t count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) st[i] = kNoFail;
}

}  // namespace

#define LLH_GRP_LAUNCH(name, kernel, WorkT)                                                             \
  template <typename T>                                                                                \
  hipError_t name(const BatchBase<T>* base, const BatchItem* items, const WorkT* work, int count,        \
                  hipStream_t s) {                                                                     \
    if (count <= 0) return hipSuccess;                                                                 \
    hipLaunchKernelGGL((kernel<T>), dim3(count), dim3(256), 0, s, base, items, work, count);            \
    return hipGetLastError();                                                                          \
  }                                                                                                    \
  template hipError_t name<double>(const BatchBase<double>*, const BatchItem*, const WorkT*, int,       \
                                   hipStream_t);                                                       \
  template hipError_t name<float>(const BatchBase<float>*, const BatchItem*, const WorkT*, int, hipStream_t);

LLH_GRP_LAUNCH(launch_grp_trtri, grp_trtri_kernel, StepWork)
LLH_GRP_LAUNCH(launch_grp_panel, grp_panel_kernel, TileWork)
LLH_GRP_LAUNCH(launch_grp_syrk, grp_syrk_kernel, TileWork)
LLH_GRP_LAUNCH(launch_grp_fwd, grp_fwd_kernel, StepWork)
LLH_GRP_LAUNCH(launch_grp_bwd, grp_bwd_kernel, StepWork)

template <typename T>
hipError_t launch_grp_solve(const BatchBase<T>* base, const BatchItem* items, int count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL((grp_solve_kernel<T>), dim3(count), dim3(256), 0, s, base, items, count);
  return hipGetLastError();
}
template hipError_t launch_grp_solve<double>(const BatchBase<double>*, const BatchItem*, int, hipStream_t);
template hipError_t launch_grp_solve<float>(const BatchBase<float>*, const BatchItem*, int, hipStream_t);

template <typename T>
hipError_t launch_solve_lds(int n, int m, const T* a, int64_t lda, const T* linv, T* b, int64_t ldb,
                            const int32_t* gate, hipStream_t s) {
  if (n <= 0 || m <= 0) return hipSuccess;
  if (n > kSolveLdsMaxN || lda > INT32_MAX) return hipErrorInvalidValue;
  const bool vec = (uintptr_t)a % 16 == 0 && (lda * (int64_t)sizeof(T)) % 16 == 0 && (uintptr_t)linv % 16 == 0;
  auto go = [&](auto mct, auto vc) {
    hipLaunchKernelGGL((solve_lds_kernel<T, decltype(mct)::value, decltype(vc)::value>), dim3(1), dim3(256), 0, s,
                       n, m, a, lda, linv, b, ldb, gate);
  };
  auto pick = [&](auto vc) {
    if (m == 1) go(std::integral_constant<int, 1>{}, vc);
    else if (m == 2) go(std::integral_constant<int, 2>{}, vc);
    else if (m <= 4) go(std::integral_constant<int, 4>{}, vc);
    else go(std::integral_constant<int, kSolveCols>{}, vc);
  };
  if (vec) pick(std::true_type{});
  else pick(std::false_type{});
  return hipGetLastError();
}
template hipError_t launch_solve_lds<double>(int, int, const double*, int64_t, const double*, double*, int64_t,
                                             const int32_t*, hipStream_t);
template hipError_t launch_solve_lds<float>(int, int, const float*, int64_t, const float*, float*, int64_t,
                                            const int32_t*, hipStream_t);

template <typename T>
int med_resident_per_cu(bool solve) {
  int nb = 0;
  const hipError_t e = solve ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, med_kernel<T, true>, 256, 0)
                             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, med_kernel<T, false>, 256, 0);
  return e == hipSuccess && nb > 0 ? nb : 1;
}
template int med_resident_per_cu<double>(bool);
template int med_resident_per_cu<float>(bool);

template <typename T>
hipError_t launch_med(const BatchBase<T>* base, const BatchItem* items, int count, int counter, int grid,
                      bool solve, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  grid = std::max(1, std::min(grid, count));
  if (solve)
    hipLaunchKernelGGL((med_kernel<T, true>), dim3(grid), dim3(256), 0, s, base, items, count, counter);
  else
    hipLaunchKernelGGL((med_kernel<T, false>), dim3(grid), dim3(256), 0, s, base, items, count, counter);
  return hipGetLastError();
}
template hipError_t launch_med<double>(const BatchBase<double>*, const BatchItem*, int, int, int, bool,
                                       hipStream_t);
template hipError_t launch_med<float>(const BatchBase<float>*, const BatchItem*, int, int, int, bool, hipStream_t);

hipError_t launch_mixed_flat(const MixedBase* mb, int op, int64_t elems, hipStream_t s) {
  if (elems <= 0) return hipSuccess;
  const int64_t blocks = std::min<int64_t>((elems + 255) / 256, 8192);
  hipLaunchKernelGGL(mixed_flat_kernel, dim3((unsigned)blocks), dim3(256), 0, s, mb, op, elems);
  return hipGetLastError();
}

hipError_t launch_grp_residual(const MixedBase* mb, const BatchItem* items, const StepWork* work, int count,
                               hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(grp_residual_kernel, dim3(count), dim3(256), 0, s, mb, items, work, count);
  return hipGetLastError();
}

hipError_t launch_grp_update(const MixedBase* mb, const BatchItem* items, const StepWork* work, int count, bool add,
                             hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(grp_update_kernel, dim3(count), dim3(256), 0, s, mb, items, work, count, add);
  return hipGetLastError();
}

hipError_t launch_fill_status(int32_t* status, int64_t count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_status_kernel, dim3((count + 255) / 256), dim3(256), 0, s, status, count);
  return hipGetLastError();
}

}  // namespace llh

// Device code of the fast factor + inverse of one 64 x 64 diagonal block,
// shared by the blocked paths' kernels (k_diag.hip) and the fused batched
// kernel (k_batch.hip).  The trailing updates of the blocked paths round
// differently from the reference anyway, so the block may too; small whole
// matrices keep the exact-order kernel.
//
// Four waves, lane i = row i in every wave; wave w holds columns
// [16w, 16w + 16) of its row in registers.  The block is factored in four
// steps; in step w wave w factors its two 8-column panels with the pivot and
// each L[k, j] taken from their lanes by v_readlane (no barrier inside a wave's
// step, and no branch on the data: a failing pivot is recorded, not branched
// on) and writes them to a row-major copy of L in LDS.  After one barrier every
// later wave applies the step's rank-16 update to its own columns, so the
// trailing work is spread over the four waves.  The inverse goes by 16-row
// blocks: each wave inverts one 16 x 16 diagonal block, then the blocks below
// the diagonal come from 16x16x16 MFMA products, one level of block rows at a
// time.  Rows beyond b are identity rows, so the block
// factors to diag(L, I) and its inverse likewise.  Pivot test, failure row and
// write-back follow cholesky.rs:54-74: on a failure at row f the block holds
// rows < f and L[f, ..f].
#pragma once

#include <type_traits>
#include <utility>

#include "k_gemm_impl.h"
#include "llh_common.h"

namespace llh {


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

// rp = 1 / sqrt(d) and piv = sqrt(d) = d * rp for d > 0: the hardware
// reciprocal square root refined by Newton steps y <- y (3/2 - (d/2) y^2)
// (two in f64, one in f32, each doubling the correct bits), a chain of about
// eight dependent operations per column where the correctly rounded square
// root followed by a division takes about twenty-five.  Both results are
// within an ulp or two, which the blocked factor's tolerance admits (the
// exact-order kernel keeps the reference's square root and division).
__device__ __forceinline__ void pivot_rsqrt(double d, double& piv, double& rp) {
  double y = __builtin_amdgcn_rsq(d);
  const double h = 0.5 * d;
  y = y * __builtin_fma(-h, y * y, 1.5);
  y = y * __builtin_fma(-h, y * y, 1.5);
  rp = y;
  piv = d * y;
}
__device__ __forceinline__ void pivot_rsqrt(float d, float& piv, float& rp) {
  float y = __builtin_amdgcn_rsqf(d);
  y = y * __builtin_fmaf(-0.5f * d, y * y, 1.5f);
  rp = y;
  piv = d * y;
}

constexpr int kDiagS = kNB + 2;  // LDS row stride of L (keeps rows 16-byte aligned)
constexpr int kDiagP = 8;        // panel width
constexpr int kDiagW = 16;       // columns (factor) / rows (inverse) per wave
constexpr int kDiagT = 256;      // threads of a diagonal-block workgroup

// Packed lower triangle of a 16 x 16 block: entry (q, c), q >= c.
__device__ __forceinline__ int xd_at(int q, int c) { return ((q * (q + 1)) >> 1) + c; }

// 38.6 KB in f64 (four workgroups per CU).
template <typename T>
struct DiagLds {
  T l[kNB * kDiagS];             // L (row-major), written a panel at a time
  T xd[4][137];                  // inverses of the four 16 x 16 diagonal sub-blocks, packed (xd_at), + a sink
  T rinv[kNB];                   // 1 / L[j, j]
  int fail[kNB / kDiagP];        // per step: failing column or -1
  int skip;                      // the system has already failed: nothing to do
};

// TS: storage type of the block and of the inverse; T: arithmetic type.
// Called by all kDiagT threads of a workgroup; returns true (in every thread)
// when the block failed, after recording the failing row in *info.  inv (64 x
// 64, row-major) is stored by column pairs: it must be 2 * sizeof(TS)-aligned
// (every caller's inverse slots are whole 64 x 64 blocks of an allocation).
template <typename TS, typename T>
__device__ __forceinline__ bool diag_quad_block(TS* __restrict__ a, int64_t lda, int b, TS* __restrict__ inv,
                                                int32_t* __restrict__ info, int64_t info_base, DiagLds<T>& sm) {
  constexpr int N = kNB, S = kDiagS, P = kDiagP, WC = kDiagW;
  typedef T v2 __attribute__((ext_vector_type(2)));
  const int i = threadIdx.x & 63;     // row (factor) / column of inv(L) (inverse)
  const int w = threadIdx.x >> 6;     // wave: columns (factor) / rows (inverse) [16w, 16w + 16)
  const int c0 = w * WC;
  T* l = sm.l;
  if (threadIdx.x < N / P) sm.fail[threadIdx.x] = -1;
  // this lane's 16 columns of row i; masked entries read the block's first
  // element (always in bounds) so that the loads carry no branches.  (Loading
  // through LDS with consecutive lanes on consecutive columns measured slower:
  // these loads overlap the first panel, a staged copy cannot.)
  T r[WC];
  {
    const TS* row = a + (int64_t)min(i, b - 1) * lda;
#pragma unroll
    for (int kk = 0; kk < WC; ++kk) {
      const int k = c0 + kk;
      const bool in = i < b && k <= i;
      const T v = T(*(in ? row + k : a));
      r[kk] = in ? v : ((i >= b && i == k) ? T(1) : T(0));
    }
  }
  __syncthreads();
  // No branch depends on the data: a failing pivot only records its column in
  // f and the factor runs on through NaNs, which land in columns >= f, where
  // the write-back does not look; every loop keeps its constant trip count and
  // r stays in registers.
  T rpl = T(0);  // 1 / L[i, i] once this lane's diagonal column is factored
  // Factors the 8-column panel of this wave's columns starting at lo (block
  // column j0): pivots and L[k, j] by v_readlane; f becomes the first failing
  // column.
  auto factor8 = [&](auto lo_c, int j0, int& f) {
    constexpr int lo = decltype(lo_c)::value;
#pragma unroll
    for (int c = 0; c < P; ++c) {
      const int j = j0 + c;
      const T d = bcast(r[lo + c], j);
      f = (f < 0 && d <= T(0)) ? j : f;  // cholesky.rs:69 (NaN passes)
      T piv, rp;
      pivot_rsqrt(d, piv, rp);
      r[lo + c] = (i == j) ? piv : r[lo + c] * rp;
      rpl = (i == j) ? rp : rpl;
#pragma unroll
      for (int q = c + 1; q < P; ++q) r[lo + q] -= r[lo + c] * bcast(r[lo + c], j0 + q);
    }
#pragma unroll
    for (int q = 0; q < P; q += 2) *reinterpret_cast<v2*>(&l[i * S + j0 + q]) = v2{r[lo + q], r[lo + q + 1]};
  };
  // Step ws: wave ws factors its two panels (the second after the first's
  // update of it, from the LDS copy this wave has just written: one wave's LDS
  // accesses complete in order, so no barrier), then every later wave applies
  // the rank-16 update to its own columns.
  static_for(std::make_integer_sequence<int, N / WC>{}, [&](auto wsi) {
    constexpr int ws = decltype(wsi)::value, j0 = ws * WC;
    if (w == ws) {
      int f = -1;
      factor8(std::integral_constant<int, 0>{}, j0, f);
      __builtin_amdgcn_wave_barrier();
      T pr[P];
#pragma unroll
      for (int q = 0; q < P; q += 2) {
        const v2 v = *reinterpret_cast<const v2*>(&l[i * S + j0 + q]);
        pr[q] = v[0];
        pr[q + 1] = v[1];
      }
#pragma unroll
      for (int kk = P; kk < WC; ++kk) {
        T t = r[kk];
#pragma unroll
        for (int q = 0; q < P; q += 2) {
          const v2 lk = *reinterpret_cast<const v2*>(&l[(j0 + kk) * S + j0 + q]);
          t = fma_t(-pr[q], lk[0], fma_t(-pr[q + 1], lk[1], t));
        }
        r[kk] = t;
      }
      factor8(std::integral_constant<int, P>{}, j0 + P, f);
      if (i == 0) sm.fail[ws] = f;
      if (i >= j0 && i < j0 + WC) sm.rinv[i] = rpl;
    }
    __syncthreads();
    if (w > ws) {
      // rank-16 update of this wave's columns: r[k] -= sum_q L[i, j0+q] L[k, j0+q]
      T pr[WC];
#pragma unroll
      for (int q = 0; q < WC; q += 2) {
        const v2 v = *reinterpret_cast<const v2*>(&l[i * S + j0 + q]);
        pr[q] = v[0];
        pr[q + 1] = v[1];
      }
#pragma unroll
      for (int kk = 0; kk < WC; ++kk) {
        const int k = c0 + kk;
        T t = r[kk];
#pragma unroll
        for (int q = 0; q < WC; q += 2) {
          const v2 lk = *reinterpret_cast<const v2*>(&l[k * S + j0 + q]);
          t = fma_t(-pr[q], lk[0], fma_t(-pr[q + 1], lk[1], t));
        }
        r[kk] = t;
      }
    }
  });
  // the first failing step's column (the steps after it ran on garbage)
  int fail = -1;
#pragma unroll
  for (int ws = 0; ws < N / WC; ++ws) fail = fail < 0 ? sm.fail[ws] : fail;
  // write back from the LDS copy of L (every panel up to the failing one was
  // published there), consecutive lanes on consecutive column pairs of a row
  // (one 16-byte LDS read, and in f64 one 16-byte store when the pair is whole
  // and a's rows keep pairs aligned): rows < f (or all rows) with their lower
  // part, row f up to f-1
  typedef TS s2 __attribute__((ext_vector_type(2)));
  {
    const bool a2 = (uintptr_t)a % (2 * sizeof(TS)) == 0 && lda % 2 == 0;
#pragma unroll
    for (int e = 0; e < N * N / (2 * kDiagT); ++e) {
      const int idx = threadIdx.x + kDiagT * e, row = idx / (N / 2), col = 2 * (idx % (N / 2));
      const int last = (fail < 0 || row < fail) ? row : (row == fail ? row - 1 : -1);
      const v2 v = *reinterpret_cast<const v2*>(&l[row * S + col]);
      TS* dst = a + (int64_t)row * lda + col;
      if (row < b) {
        if (a2 && col + 1 <= last) {
          *reinterpret_cast<s2*>(dst) = s2{TS(v[0]), TS(v[1])};
        } else {
          if (col <= last) dst[0] = TS(v[0]);
          if (col + 1 <= last) dst[1] = TS(v[1]);
        }
      }
    }
  }
  if (fail >= 0) {
    if (threadIdx.x == 0) atomicMin(info, (int32_t)(info_base + fail));
    return true;
  }
  // inv(L) by 16-row blocks.  Wave bb forms X_bb, the inverse of the 16 x 16
  // diagonal block bb (lane c = column c, a column sweep reading L from the LDS
  // copy as uniform-address loads).  Then, for block rows bi = 1..3, wave bj < bi
  // forms X_ij = -X_ii sum_{k=j..i-1} L_ik X_kj as 16x16x16 MFMA products; the
  // sum's result registers feed the second product's B operand directly (result
  // register s holds the rows that step s takes), and X_ij goes to the free upper
  // triangle of the LDS copy, block (j, i).  (Row panels of 8 with a barrier and a
  // rank-8 update each took 16k of the block's 52k cycles.)
  {
    using MF = Mfma<T>;
    using acc_t = typename MF::acc_t;
    const int cb = 16 * w, cc = i & 15, fr = i & 15;
    {
      T x[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) x[q] = q == cc ? T(1) : T(0);
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        x[q] *= sm.rinv[cb + q];
#pragma unroll
        for (int q2 = q + 1; q2 < 16; ++q2) x[q2] = fma_t(-l[(cb + q2) * S + cb + q], x[q], x[q2]);
      }
      if (i < 16) {
#pragma unroll
        for (int q = 0; q < 16; ++q) sm.xd[w][q >= cc ? xd_at(q, cc) : 136] = x[q];  // 136: a sink for q < cc
      }
    }
    __syncthreads();
    for (int bi = 1; bi < 4; ++bi) {
      const int bj = w;
      if (bj < bi) {
        acc_t sacc = acc_t{0, 0, 0, 0};
        for (int bk = bj; bk < bi; ++bk) {
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const int k = MF::row(i, s);
            const T xb = bk == bj ? (k >= fr ? sm.xd[bj][xd_at(k, fr)] : T(0))
                                  : l[(16 * bj + k) * S + 16 * bk + fr];
            sacc = MF::mma(l[(16 * bi + fr) * S + 16 * bk + k], xb, sacc);
          }
        }
        acc_t xc = acc_t{0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int k = MF::row(i, s);
          xc = MF::mma(fr >= k ? -sm.xd[bi][xd_at(fr, k)] : T(0), sacc[s], xc);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) l[(16 * bj + MF::row(i, q)) * S + 16 * bi + fr] = xc[q];
      }
      __syncthreads();
    }
  }
  // inv, row-major 64 x 64 (column pairs, as above): X_ii from xd, X_ij (i > j)
  // from block (j, i), zero above
#pragma unroll
  for (int e = 0; e < N * N / (2 * kDiagT); ++e) {
    const int idx = threadIdx.x + kDiagT * e, row = idx / (N / 2), col = 2 * (idx % (N / 2));
    const int bi = row >> 4, bj = col >> 4, rr = row & 15, cc = col & 15;
    T v0 = T(0), v1 = T(0);
    if (bi == bj) {
      if (cc <= rr) v0 = sm.xd[bi][xd_at(rr, cc)];
      if (cc + 1 <= rr) v1 = sm.xd[bi][xd_at(rr, cc + 1)];
    } else if (bi > bj) {
      const v2 u = *reinterpret_cast<const v2*>(&l[(16 * bj + rr) * S + 16 * bi + cc]);
      v0 = u[0];
      v1 = u[1];
    }
    *reinterpret_cast<s2*>(&inv[row * N + col]) = s2{TS(v0), TS(v1)};
  }
  return false;
}

}  // namespace llh

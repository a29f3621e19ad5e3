// Device code shared by the MFMA GEMM kernels (k_gemm.hip) and the grouped
// batch kernels (k_batch.hip): the MFMA wrappers and one NT tile,
//   C[m0.., n0..] = beta * C + alpha * A * B^T  (TRI: only row >= col),
// with both operands k-contiguous.
//
// MFMA lane layouts (measured on gfx950):
//   A operand: lane l holds op(A)[l%16][k0 + l/16]
//   B operand: lane l holds op(B)[k0 + l/16][l%16]
//   f64 result reg i: C[4*i + l/16][l%16];  f32 result reg i: C[4*(l/16) + i][l%16]
// The f16 form (v_mfma_f32_16x16x32_f16) takes 8 consecutive k per lane,
// op(A)[l%16][k0 + 8*(l/16) + 0..7] and likewise for B, and returns the f32
// layout.  A sum over k does not depend on which k a lane group holds, as long
// as A and B are fed the same way, which the tile below guarantees.
#pragma once

#include "llh_common.h"

namespace llh {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  using acc_t = v4d;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int i) { return 4 * i + (lane >> 4); }
  static constexpr int kRowStep = 4;  // row(lane, i) = row(lane, 0) + kRowStep * i
};
template <>
struct Mfma<float> {
  using acc_t = v4f;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int i) { return 4 * (lane >> 4) + i; }
  static constexpr int kRowStep = 1;
};
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
template <>
struct Mfma<_Float16> {
  using acc_t = v4f;
  static __device__ __forceinline__ acc_t mma(v8h a, v8h b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int i) { return 4 * (lane >> 4) + i; }
  static constexpr int kRowStep = 1;
};

// Operand elements one lane feeds into one MFMA: 1 for f64 / f32, 8 for f16.
template <typename TI>
constexpr int mfma_lane_k() { return sizeof(TI) == 2 ? 8 : 1; }

// One BM x BN tile of C = beta*C + alpha*A*B^T on WM x WN waves.
//
// Slabs of SK k are staged in double-buffered LDS, row-major ([row][k]) with
// the row stride padded to 16 mod 128 bytes, so the 8 lanes of one 128-bit LDS
// access phase read 8 rows from 8 different bank groups.  Each lane's MFMA
// fragment is KPL = 16/sizeof(T) consecutive k of one row (one 128-bit read):
// within a sub-chunk of 4*KPL k, MFMA s takes k = sub + KPL*(lane/16) + s from
// every lane, for A and B alike, so the KPL MFMAs of a sub-chunk together sum all
// of its k.  One barrier per slab: slab t+1 is loaded into registers while slab
// t is multiplied, then written to the other LDS buffer.
//
// Rows beyond M (or N) are clamped to the last valid row, so interior loads
// carry no bounds checks; their products land in rows the epilogue discards.
// With TOUT, the stored tile is also left in tile_out (LDS, row stride BN + 1,
// free once the last slab is used: it may alias the slabs), zero outside C.
// Only a final partial slab of k is loaded with masking.  A may alias C only when
// each tile's rows of A are its own rows of C (every read precedes the writes).
//
// PRE (beta != 0, alpha != 0, at least FM*FN/CPS slabs): C is read during the
// main loop instead of after it.  The last FM*FN/CPS slabs each issue the loads
// of CPS accumulator fragments of C, and the next slab folds them in as
// acc += (beta/alpha) C, so the epilogue only stores alpha * acc.  Without it
// every workgroup of a rank-k update with small k reads its C tile at the same
// moment, after its MFMAs, and the launch waits on one HBM burst per tile.
//
// TI is the operand type: T itself, or _Float16 with T = float (f16 products
// accumulated in f32; one MFMA then takes a lane's whole 8-element fragment).
//
// nt_tile_in takes the two double-buffered LDS slabs from the caller (the
// types NtSlab<TI, BM> for A and NtSlab<TI, BN> for B, two buffers each), so a
// kernel that runs other phases between tiles can share the memory; nt_tile
// declares them itself.
template <typename TI, int ROWS>
using NtSlab = TI[ROWS * 9 * (16 / sizeof(TI))];

template <typename T, int BM, int BN, int WM, int WN, bool TRI, bool VEC, bool PRE = false, typename TI = T,
          bool TOUT = false>
__device__ __forceinline__ void nt_tile_in(int M, int N, int K, T alpha, const TI* A, int64_t lda,
                                           const TI* __restrict__ B, int64_t ldb, T beta, T* C, int64_t ldc,
                                           int m0, int n0, NtSlab<TI, BM>* As, NtSlab<TI, BN>* Bs,
                                           T* tile_out = nullptr) {
  constexpr int NTHR = WM * WN * 64;
  constexpr int TM = BM / WM, TN = BN / WN;
  constexpr int FM = TM / 16, FN = TN / 16;
  constexpr int KPL = 16 / sizeof(TI);            // k per lane per 128-bit read
  constexpr int OPK = mfma_lane_k<TI>();          // of which one MFMA takes OPK
  static_assert(OPK == 1 || OPK == KPL, "an MFMA takes one element or the whole fragment");
  static_assert(sizeof(TI) == sizeof(T) || (sizeof(TI) == 2 && sizeof(T) == 4), "f16 operands accumulate in f32");
  constexpr int SUB = 4 * KPL;                    // k covered by KPL / OPK MFMAs
  constexpr int SK = 2 * SUB;                     // k per slab
  constexpr int LDK = SK + KPL;                   // padded LDS row stride (elements)
  constexpr int LPR = SK / KPL;                   // lanes per slab row
  constexpr int RPP = NTHR / LPR;                 // rows per load pass
  constexpr int PA = BM / RPP, PB = BN / RPP;     // load passes
  static_assert(BM % RPP == 0 && BN % RPP == 0, "tile rows must divide the load passes");
  using MF = Mfma<TI>;
  using acc_t = typename MF::acc_t;
  typedef TI vec_t __attribute__((ext_vector_type(KPL)));

  static_assert(sizeof(NtSlab<TI, BM>) == BM * LDK * sizeof(TI) && sizeof(NtSlab<TI, BN>) == BN * LDK * sizeof(TI),
                "LDS slab types match the tile");

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm0 = (wave / WN) * TM, wn0 = (wave % WN) * TN;
  const int lrow = t / LPR, lk = (t % LPR) * KPL;

  // Clamped source rows of this thread's loads.
  const TI* pa[PA];
  const TI* pb[PB];
#pragma unroll
  for (int p = 0; p < PA; ++p) pa[p] = A + (int64_t)min(m0 + lrow + p * RPP, M - 1) * lda + lk;
#pragma unroll
  for (int p = 0; p < PB; ++p) pb[p] = B + (int64_t)min(n0 + lrow + p * RPP, N - 1) * ldb + lk;

  acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  vec_t ra[PA], rb[PB];
  auto load = [&](int k0) {
    if (VEC && k0 + SK <= K) {
#pragma unroll
      for (int p = 0; p < PA; ++p) ra[p] = *reinterpret_cast<const vec_t*>(pa[p] + k0);
#pragma unroll
      for (int p = 0; p < PB; ++p) rb[p] = *reinterpret_cast<const vec_t*>(pb[p] + k0);
    } else {
#pragma unroll
      for (int p = 0; p < PA; ++p)
#pragma unroll
        for (int e = 0; e < KPL; ++e) ra[p][e] = (k0 + lk + e < K) ? pa[p][k0 + e] : TI(0);
#pragma unroll
      for (int p = 0; p < PB; ++p)
#pragma unroll
        for (int e = 0; e < KPL; ++e) rb[p][e] = (k0 + lk + e < K) ? pb[p][k0 + e] : TI(0);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int p = 0; p < PA; ++p) *reinterpret_cast<vec_t*>(&As[buf][(lrow + p * RPP) * LDK + lk]) = ra[p];
#pragma unroll
    for (int p = 0; p < PB; ++p) *reinterpret_cast<vec_t*>(&Bs[buf][(lrow + p * RPP) * LDK + lk]) = rb[p];
  };

  const int nslab = (K + SK - 1) / SK;
  const int fr = lane & 15, fq = (lane >> 4) * KPL;

  // C prefetch: fragment q = (i, j) -> group q / CPS, loaded in slab pre0 + group
  constexpr int NF = FM * FN;
  constexpr int CPS = sizeof(T) == 8 ? 1 : 2;  // fragments per slab (8 VGPRs each)
  constexpr int NG = NF / CPS;
  static_assert(NF % CPS == 0, "fragments split evenly over slabs");
  const bool pre = PRE && alpha != T(0) && beta != T(0) && nslab >= NG;
  const int pre0 = nslab - NG;
  const T cscale = pre ? beta / alpha : T(0);
  acc_t cbuf[CPS];
  const int ci0 = m0 + wm0 + MF::row(lane, 0), cj0 = n0 + wn0 + fr;
  // The loads are unconditional, from indices clamped into the matrix (an
  // entry above the diagonal of a square C is still C's memory); the entries
  // outside the tile's part of C are dropped when folded in.  Masked loads here
  // made the compiler zero cbuf first and wait for every outstanding load,
  // the next slab's operands included, before the MFMAs.
  auto cload = [&](int g) {
#pragma unroll
    for (int c = 0; c < CPS; ++c) {
      const int q = g * CPS + c, i = q / FN, j = q % FN;  // uniform
      const int gj = min(cj0 + j * 16, N - 1);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = min(ci0 + i * 16 + MF::kRowStep * r, M - 1);
        cbuf[c][r] = C[(int64_t)gi * ldc + gj];
      }
    }
  };
  auto cadd = [&](int g) {
#pragma unroll
    for (int q = 0; q < NF; ++q)
      if (q / CPS == g) {
        const int i = q / FN, j = q % FN, gj = cj0 + j * 16;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gi = ci0 + i * 16 + MF::kRowStep * r;
          const bool in = gi < M && gj < N && (!TRI || gi >= gj);
          acc[i][j][r] += cscale * (in ? cbuf[q % CPS][r] : T(0));
        }
      }
  };

  load(0);
  store(0);
  __syncthreads();
  for (int sl = 0; sl < nslab; ++sl) {
    const int cur = sl & 1;
    const bool more = sl + 1 < nslab;
    if (PRE && pre && sl > pre0) cadd(sl - 1 - pre0);
    if (more) load((sl + 1) * SK);
    if (PRE && pre && sl >= pre0) cload(sl - pre0);
    const TI* as = As[cur];
    const TI* bs = Bs[cur];
#pragma unroll
    for (int c = 0; c < SK; c += SUB) {
      vec_t af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) af[i] = *reinterpret_cast<const vec_t*>(&as[(wm0 + i * 16 + fr) * LDK + c + fq]);
#pragma unroll
      for (int j = 0; j < FN; ++j) bf[j] = *reinterpret_cast<const vec_t*>(&bs[(wn0 + j * 16 + fr) * LDK + c + fq]);
      if constexpr (OPK == 1) {
#pragma unroll
        for (int s = 0; s < KPL; ++s)
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) acc[i][j] = MF::mma(af[i][s], bf[j][s], acc[i][j]);
      } else {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) acc[i][j] = MF::mma(af[i], bf[j], acc[i][j]);
      }
    }
    if (more) store(cur ^ 1);
    __syncthreads();
  }
  if (PRE && pre) cadd(NG - 1);

#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = m0 + wm0 + i * 16 + MF::row(lane, r);
        const int gj = n0 + wn0 + j * 16 + fr;
        T out = T(0);
        if (gi < M && gj < N && (!TRI || gi >= gj)) {
          T* c = C + (int64_t)gi * ldc + gj;
          const T v = alpha * acc[i][j][r];
          out = (beta == T(0) || (PRE && pre)) ? v : beta * *c + v;
          *c = out;
        }
        if constexpr (TOUT) tile_out[(wm0 + i * 16 + MF::row(lane, r)) * (BN + 1) + wn0 + j * 16 + fr] = out;
      }
}

template <typename T, int BM, int BN, int WM, int WN, bool TRI, bool VEC, bool PRE = false, typename TI = T>
__device__ __forceinline__ void nt_tile(int M, int N, int K, T alpha, const TI* A, int64_t lda,
                                        const TI* __restrict__ B, int64_t ldb, T beta, T* C, int64_t ldc,
                                        int m0, int n0) {
  __shared__ __attribute__((aligned(16))) NtSlab<TI, BM> As[2];
  __shared__ __attribute__((aligned(16))) NtSlab<TI, BN> Bs[2];
  nt_tile_in<T, BM, BN, WM, WN, TRI, VEC, PRE, TI>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, m0, n0, As, Bs);
}

}  // namespace llh

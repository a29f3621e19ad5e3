// MFMA GEMM for the trailing SYRK update of the blocked POTRF and the block
// updates of the triangular solves.
//
//   C = beta * C + alpha * op(A) * op(B)        (TRI: only entries with row >= col)
//
// op(A) is M x K: TA = false reads A[i*lda + k] (row-major M x K), TA = true
// reads A[k*lda + i] (row-major K x M).  op(B) is K x N: TB = true reads
// B[j*ldb + k] (row-major N x K, i.e. C = A * B^T as in SYRK), TB = false reads
// B[k*ldb + j] (row-major K x N).
//
// Tiling for CDNA4: a BM x BN workgroup tile over WM x WN wave64 waves, each
// wave accumulating (BM/WM) x (BN/WN) in 16x16 MFMA blocks
// (v_mfma_f64_16x16x4_f64 / v_mfma_f32_16x16x4_f32).  K advances in slabs of
// 16 staged through LDS in k-major order, so a fragment read is 16 consecutive
// elements of one LDS row; the next slab is prefetched into registers while the
// current one is multiplied.  Lane layouts (measured on gfx950):
//   A operand: lane l holds op(A)[l%16][k0 + l/16]
//   B operand: lane l holds op(B)[k0 + l/16][l%16]
//   f64 result reg i: C[4*i + l/16][l%16];  f32 result reg i: C[4*(l/16) + i][l%16]
#include "llh_common.h"
#include "llh_kernels.h"
#include "k_gemm_impl.h"

namespace llh {
namespace {

constexpr int BK = 16;

// Loads one BK-slab of an operand tile (R rows of op(X), i.e. R output rows or
// columns) from global memory into registers.  KMAJOR_SRC: the source is stored
// with k as its row index (contiguous along the R direction); otherwise the
// source rows are the R index and k is contiguous.
template <typename T, int R, int NTHR, bool KMAJOR_SRC>
struct SlabLoader {
  static constexpr int PER = (R * BK) / NTHR;  // elements per thread
  static_assert((R * BK) % NTHR == 0, "tile must divide over the threads");
  T v[PER];

  __device__ __forceinline__ void load(const T* __restrict__ X, int64_t ldx, int r0, int rmax, int k0,
                                       int kmax) {
    const int t = threadIdx.x;
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      int rr, kk;
      if constexpr (KMAJOR_SRC) {  // consecutive threads walk the contiguous r
        const int e = t + p * NTHR;
        rr = e % R;
        kk = e / R;
      } else {  // consecutive threads walk the contiguous k
        const int e = t + p * NTHR;
        kk = e % BK;
        rr = e / BK;
      }
      const int gr = r0 + rr, gk = k0 + kk;
      T x = T(0);
      if (gr < rmax && gk < kmax) x = KMAJOR_SRC ? X[(int64_t)gk * ldx + gr] : X[(int64_t)gr * ldx + gk];
      v[p] = x;
    }
  }
  template <int LDR>
  __device__ __forceinline__ void store(T* S) const {
    const int t = threadIdx.x;
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      int rr, kk;
      const int e = t + p * NTHR;
      if constexpr (KMAJOR_SRC) {
        rr = e % R;
        kk = e / R;
      } else {
        kk = e % BK;
        rr = e / BK;
      }
      S[kk * LDR + rr] = v[p];
    }
  }
};

template <typename T, int BM, int BN, int WM, int WN, bool TA, bool TB, bool TRI>
// A and C may alias (the in-place panel solve X := X * Linv^T): every read of
// A precedes the epilogue's writes, and a workgroup only touches its own rows.
__global__ __launch_bounds__(WM* WN * 64) void gemm_kernel(int M, int N, int K, T alpha, const T* A, int64_t lda,
                                                           const T* __restrict__ B, int64_t ldb, T beta, T* C,
                                                           int64_t ldc, const int32_t* __restrict__ gate) {
  constexpr int NTHR = WM * WN * 64;
  constexpr int TM = BM / WM, TN = BN / WN;
  constexpr int FM = TM / 16, FN = TN / 16;
  constexpr int PAD = sizeof(T) == 8 ? 2 : 4;
  constexpr int LDA_S = BM + PAD, LDB_S = BN + PAD;
  using MF = Mfma<T>;
  using acc_t = typename MF::acc_t;

  if (gate != nullptr && *gate != kNoFail) return;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  if (TRI && m0 + BM - 1 < n0) return;  // tile entirely above the diagonal

  __shared__ T As[BK * LDA_S];
  __shared__ T Bs[BK * LDB_S];

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm0 = (wave / WN) * TM, wn0 = (wave % WN) * TN;

  acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  SlabLoader<T, BM, NTHR, TA> la;
  SlabLoader<T, BN, NTHR, !TB> lb;
  la.load(A, lda, m0, M, 0, K);
  lb.load(B, ldb, n0, N, 0, K);
  la.template store<LDA_S>(As);
  lb.template store<LDB_S>(Bs);
  __syncthreads();

  const int fr = lane & 15, fk = lane >> 4;
  for (int k0 = 0; k0 < K; k0 += BK) {
    const bool more = k0 + BK < K;
    if (more) {
      la.load(A, lda, m0, M, k0 + BK, K);
      lb.load(B, ldb, n0, N, k0 + BK, K);
    }
#pragma unroll
    for (int s = 0; s < BK / 4; ++s) {
      T af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) af[i] = As[(s * 4 + fk) * LDA_S + wm0 + i * 16 + fr];
#pragma unroll
      for (int j = 0; j < FN; ++j) bf[j] = Bs[(s * 4 + fk) * LDB_S + wn0 + j * 16 + fr];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = MF::mma(af[i], bf[j], acc[i][j]);
    }
    __syncthreads();
    if (more) {
      la.template store<LDA_S>(As);
      lb.template store<LDB_S>(Bs);
      __syncthreads();
    }
  }

#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = m0 + wm0 + i * 16 + MF::row(lane, r);
        const int gj = n0 + wn0 + j * 16 + fr;
        if (gi < M && gj < N && (!TRI || gi >= gj)) {
          T* c = C + (int64_t)gi * ldc + gj;
          const T v = alpha * acc[i][j][r];
          *c = (beta == T(0)) ? v : beta * *c + v;
        }
      }
}

// ---------------------------------------------------------------------------
// C = beta*C + alpha*A*B^T with both operands k-contiguous (the SYRK of the
// trailing update and the inner-panel update), one nt_tile per workgroup.
//
// Tile order: WG ids are dealt round-robin to the 8 XCDs, so the id is first
// turned into (xcd, slot) and each XCD receives a contiguous run of tiles; tiles
// are ordered in groups of kGroupM tile rows, column by column within a group, so
// the WGs resident on one XCD share a few A and B row panels in its L2.
//
// For the square triangular case (M == N) only the tiles that meet the lower
// triangle are enumerated: with BM = R * BN, tile row tm needs the columns
// tn < R * (tm + 1).  Enumerating the full rectangle instead would hand the XCDs
// that own the top tile rows mostly empty tiles and the last ones only full tiles.
constexpr int kGroupM = 8;

__host__ __device__ inline int tri_group_tiles(int g0, int gm, int tiles_n, int R) {
  const int cg = min(tiles_n, R * (g0 + gm));
  int cnt = R * g0 * gm;  // columns left of the group's diagonal band: all gm rows
  for (int c = R * g0; c < cg; ++c) cnt += g0 + gm - c / R;
  return cnt;
}

__host__ __device__ inline int tri_total(int tiles_m, int tiles_n, int R) {
  int total = 0;
  for (int g0 = 0; g0 < tiles_m; g0 += kGroupM) total += tri_group_tiles(g0, min(kGroupM, tiles_m - g0), tiles_n, R);
  return total;
}

__device__ inline void tri_tile(int lin, int tiles_m, int tiles_n, int R, int& tm, int& tn) {
  int rem = lin, g0 = 0;
  for (;;) {
    const int cnt = tri_group_tiles(g0, min(kGroupM, tiles_m - g0), tiles_n, R);
    if (rem < cnt) break;
    rem -= cnt;
    g0 += kGroupM;
  }
  const int gm = min(kGroupM, tiles_m - g0);
  if (rem < R * g0 * gm) {
    tn = rem / gm;
    tm = g0 + rem % gm;
    return;
  }
  rem -= R * g0 * gm;
  int c = R * g0;  // column c of the band starts at tile row c / R
  while (rem >= g0 + gm - c / R) {
    rem -= g0 + gm - c / R;
    ++c;
  }
  tn = c;
  tm = c / R + rem;
}

// SQ: the square triangular case (TRI and M == N), i.e. the trailing SYRK.
// TI: operand type (T, or _Float16 for the f32 SYRK on f16 operands).
template <typename T, int BM, int BN, int WM, int WN, bool TRI, bool SQ, bool VEC, typename TI = T>
__global__ __launch_bounds__(WM* WN * 64, (512 / (WM * WN * 64)) > 0 ? 512 / (WM * WN * 64) : 1) void gemm_nt_kernel(int M, int N, int K, T alpha,
                                                                 const TI* A, int64_t lda,
                                                                 const TI* __restrict__ B, int64_t ldb, T beta,
                                                                 T* C, int64_t ldc,
                                                                 const int32_t* __restrict__ gate, int tiles_m,
                                                                 int tiles_n, int total) {
  static_assert(BM % BN == 0, "tile rows are a multiple of tile columns");
  if (gate != nullptr && *gate != kNoFail) return;

  // XCD-aware linear tile index (the launch has exactly 8 * per WGs)
  const int nx = 8;
  const int id = blockIdx.x;
  const int per = (total + nx - 1) / nx;
  const int lin = (id % nx) * per + id / nx;
  if (lin >= total) return;
  int tm, tn;
  if (SQ) {
    tri_tile(lin, tiles_m, tiles_n, BM / BN, tm, tn);
  } else {
    const int per_group = kGroupM * tiles_n;
    const int g = lin / per_group;
    const int first = g * kGroupM;
    const int gm = min(kGroupM, tiles_m - first);
    const int r = lin - g * per_group;
    tm = first + r % gm;
    tn = r / gm;
  }
  const int m0 = tm * BM, n0 = tn * BN;
  if (TRI && m0 + BM - 1 < n0) return;

  // C prefetch for the triangular updates (the trailing SYRK and the look-ahead
  // panel's columns): f64 on 256 x 128 tiles (the 128 x 128 f64 tile has no
  // registers left) and on 64 x 64 tiles (the small blocked factors' updates:
  // one n = 512 solvec 403 -> 383 us, n = 16384 40.9 -> 40.3 ms), and f16
  // operands on 128 x 128 tiles.  Measured on the SYRK
  // at n = 32256, k = 512: f64 53 -> 57, f16 199 -> 348 TFLOP/s; on the
  // look-ahead update as well, n = 32768: f64 solve 219.7 -> 217.4 ms, FP16+IR
  // 81.0 -> 79.6 ms.  Not
  // for f32 operands (88 -> 107 TFLOP/s): folding C into the f32 sum mid-way
  // costs the accuracy that refinement then pays for (11 instead of 3 steps at
  // cond ~4e6), and holding C to the end does not fit the registers.  Seeding
  // the f32 accumulators with C before the first slab is no better: 119 instead
  // of 104 TFLOP/s alone, no gain inside the factor (149.6 -> 148.4 ms at
  // n = 32768), and 29 instead of 4 refinement steps at lambda = 1e-6.
  constexpr bool PRE = TRI && (((BM == 256 || BM == 64) && sizeof(T) == 8) || sizeof(TI) == 2);
  nt_tile<T, BM, BN, WM, WN, TRI, VEC, PRE, TI>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, m0, n0);
}

// SQ launches the square lower-triangle form (M == N, TRI), which only the
// trailing SYRK uses (launch_syrk_lower), so that it is its own instantiation.
template <typename T, int BM, int BN, int WM, int WN, bool TRI, bool SQ = false, typename TI = T>
hipError_t launch_nt(int M, int N, int K, T alpha, const TI* A, int64_t lda, const TI* B, int64_t ldb, T beta,
                     T* C, int64_t ldc, const int32_t* gate, hipStream_t s) {
  static_assert(!SQ || TRI, "the square form is triangular");
  if (SQ && M != N) return hipErrorInvalidValue;
  const int tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
  const int total = SQ ? tri_total(tiles_m, tiles_n, BM / BN) : tiles_m * tiles_n;
  const bool vec = (reinterpret_cast<uintptr_t>(A) % 16 == 0) && (reinterpret_cast<uintptr_t>(B) % 16 == 0) &&
                   ((lda * sizeof(TI)) % 16 == 0) && ((ldb * sizeof(TI)) % 16 == 0);
  // The kernel deals WG ids to XCDs as lin = (id % 8) * per + id / 8, which is a
  // bijection onto [0, 8 * per) only when the grid has exactly 8 * per WGs.
  const int per = (total + 7) / 8;
  const dim3 grid(8 * per), block(WM * WN * 64);
#define LLH_NT_GO(SQ, VEC)                                                                                    \
  hipLaunchKernelGGL((gemm_nt_kernel<T, BM, BN, WM, WN, TRI, SQ, VEC, TI>), grid, block, 0, s, M, N, K, alpha, \
                     A, lda, B, ldb, beta, C, ldc, gate, tiles_m, tiles_n, total)
  if (vec) LLH_NT_GO(SQ, true);
  else LLH_NT_GO(SQ, false);
#undef LLH_NT_GO
  return hipGetLastError();
}

template <typename T, int BM, int BN, int WM, int WN, bool TA, bool TB, bool TRI>
hipError_t launch_cfg(int M, int N, int K, T alpha, const T* A, int64_t lda, const T* B, int64_t ldb,
                      T beta, T* C, int64_t ldc, const int32_t* gate, hipStream_t s) {
  dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM);
  hipLaunchKernelGGL((gemm_kernel<T, BM, BN, WM, WN, TA, TB, TRI>), grid, dim3(WM * WN * 64), 0, s, M, N,
                     K, alpha, A, lda, B, ldb, beta, C, ldc, gate);
  return hipGetLastError();
}

template <typename T, bool TA, bool TB, bool TRI>
hipError_t launch_shape(int M, int N, int K, T alpha, const T* A, int64_t lda, const T* B, int64_t ldb,
                        T beta, T* C, int64_t ldc, const int32_t* gate, hipStream_t s) {
  if (N <= 32)
    return launch_cfg<T, 128, 32, 4, 1, TA, TB, TRI>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, gate, s);
  if (N <= 64)
    return launch_cfg<T, 128, 64, 2, 2, TA, TB, TRI>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, gate, s);
  // 16 waves (32 x 32 per wave): invc at n = 8192 45.5 -> 36.6 ms against 4
  return launch_cfg<T, 128, 128, 4, 4, TA, TB, TRI>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, gate, s);
}

}  // namespace

template <typename T>
hipError_t launch_gemm(bool ta, bool tb, bool tri, int M, int N, int K, T alpha, const T* A, int64_t lda,
                       const T* B, int64_t ldb, T beta, T* C, int64_t ldc, const int32_t* gate,
                       hipStream_t s) {
  if (M <= 0 || N <= 0) return hipSuccess;
  if (K <= 0) K = 0;
  if (!ta && tb) {
    // Large updates: 256 x 128 tiles on 8 waves (2 per SIMD) cut the operand
    // traffic per flop by a quarter; smaller ones keep 128 x 128 tiles so the
    // grid still fills the 256 CUs; narrow ones (N <= 64, the panel solve) and
    // small ones that would give fewer than 128 such tiles use 64 x 64 tiles.
    // Triangular updates enumerate the full rectangle; tiles wholly above the
    // diagonal exit at once.  A and C may alias when every tile holds all of
    // its rows' k (the in-place panel solve: N = K <= 64, one tile column).
    const int64_t t128 = (int64_t)((M + 127) / 128) * ((N + 127) / 128);
    if (N <= 64 || t128 < 128) {
      // on 8 waves (32 x 16 per wave): n = 512 device solvec 376 -> 368 us,
      // n = 16384 39.4 -> 39.1 ms against 4 waves
      if (tri) return launch_nt<T, 64, 64, 2, 4, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, gate, s);
      return launch_nt<T, 64, 64, 2, 4, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, gate, s);
    }
    // 256 x 128 only for f64: the f32 instantiation of that tile ran the
    // look-ahead panel update (M = 31744, N = K = 1024) at 5.3 ms per launch
    // against 2.0 ms for f64 on the same shapes, while 128 x 128 is the faster
    // f32 tile for the SYRK too.
    // 128 x 128 tiles on 16 waves, as the SYRK's (f64 solve at n = 32768 215.7
    // -> 212.8 ms, FP32+IR 140.0 -> 136.0 ms against 4 waves).
    const bool big = sizeof(T) == 8 && (int64_t)M * N >= (int64_t)4096 * 4096;
    if (tri) {
      if (big) return launch_nt<T, 256, 128, 4, 2, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, gate, s);
      return launch_nt<T, 128, 128, 4, 4, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, gate, s);
    }
    if (big) return launch_nt<T, 256, 128, 4, 2, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, gate, s);
    return launch_nt<T, 128, 128, 4, 4, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, gate, s);
  }
  if (!ta && tb && tri)
    return launch_shape<T, false, true, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, gate, s);
  if (!ta && tb && !tri)
    return launch_shape<T, false, true, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, gate, s);
  if (!ta && !tb && !tri)
    return launch_shape<T, false, false, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, gate, s);
  if (ta && !tb && !tri)
    return launch_shape<T, true, false, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, gate, s);
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_syrk_lower(int n, int k, T alpha, const T* A, int64_t lda, T beta, T* C, int64_t ldc,
                             const int32_t* gate, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (k < 0) k = 0;
  // f64: 256 x 128 tiles on 8 waves (2 per SIMD) while the grid still fills the
  // chip several times over, 128 x 128 tiles for the last few small updates.
  // f32: 128 x 128 tiles throughout (measured faster than 256 x 128 for f32).
  // The 128 x 128 tiles run on 16 waves (32 x 32 each, 56 VGPRs in f32, eight
  // waves per SIMD): n = 32768 FP32+IR 148.8 -> 140.0 ms (f32 SYRK 88 -> 97
  // TFLOP/s) against the 4-wave form; 8 waves (2 x 4 or 4 x 2) gave 142 ms,
  // 256 x 128 on 16 waves 147 ms.  For f64, 16 waves on 256 x 128 (a 64 x 32
  // tile per wave at 128 VGPRs) took the solve from 216 to 225 ms.
  if (sizeof(T) == 8 && (int64_t)n * n >= (int64_t)4096 * 4096)
    return launch_nt<T, 256, 128, 4, 2, true, true>(n, n, k, alpha, A, lda, A, lda, beta, C, ldc, gate, s);
  return launch_nt<T, 128, 128, 4, 4, true, true>(n, n, k, alpha, A, lda, A, lda, beta, C, ldc, gate, s);
}

hipError_t launch_syrk_lower_h(int n, int k, float alpha, const _Float16* A, int64_t lda, float beta, float* C,
                               int64_t ldc, const int32_t* gate, hipStream_t s, int variant) {
  if (n <= 0) return hipSuccess;
  if (k < 0) k = 0;
  // 128 x 128 tiles: two workgroups per CU (72 KB of LDS each) hide more of the
  // operand latency than one 256 x 128 workgroup, measured 200 against 168
  // TFLOP/s at n = 32256, k = 512 (348 with the C prefetch); on 8 waves (64 x
  // 32 per wave) FP16+IR at n = 32768 77.0 -> 74.7 ms against 4
  if (variant == 1)
    return launch_nt<float, 256, 128, 4, 2, true, true, _Float16>(n, n, k, alpha, A, lda, A, lda, beta, C, ldc,
                                                                  gate, s);
  return launch_nt<float, 128, 128, 2, 4, true, true, _Float16>(n, n, k, alpha, A, lda, A, lda, beta, C, ldc, gate,
                                                                s);
}

hipError_t launch_gemm_nt_tri_h(int M, int N, int K, float alpha, const _Float16* A, int64_t lda,
                                const _Float16* B, int64_t ldb, float beta, float* C, int64_t ldc,
                                const int32_t* gate, hipStream_t s) {
  if (M <= 0 || N <= 0) return hipSuccess;
  if (K < 0) K = 0;
  const int64_t t128 = (int64_t)((M + 127) / 128) * ((N + 127) / 128);
  if (N <= 64 || t128 < 128)
    return launch_nt<float, 64, 64, 2, 2, true, false, _Float16>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, gate,
                                                                  s);
  return launch_nt<float, 128, 128, 2, 2, true, false, _Float16>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, gate,
                                                                  s);
}

template hipError_t launch_syrk_lower<double>(int, int, double, const double*, int64_t, double, double*, int64_t,
                                              const int32_t*, hipStream_t);
template hipError_t launch_syrk_lower<float>(int, int, float, const float*, int64_t, float, float*, int64_t,
                                             const int32_t*, hipStream_t);

template hipError_t launch_gemm<double>(bool, bool, bool, int, int, int, double, const double*, int64_t,
                                        const double*, int64_t, double, double*, int64_t, const int32_t*,
                                        hipStream_t);
template hipError_t launch_gemm<float>(bool, bool, bool, int, int, int, float, const float*, int64_t,
                                       const float*, int64_t, float, float*, int64_t, const int32_t*,
                                       hipStream_t);

}  // namespace llh

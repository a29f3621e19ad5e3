// Triangular solves on small diagonal blocks.
//
// trsm_rows: the panel step of the blocked POTRF, X := X * L11^-T, i.e. for
// every row x of the panel, x[c] = (a[c] - sum_{p<c} x[p] * L11[c,p]) / L11[c,c]
// (the division of cholesky.rs:62 applied to a whole block of rows).  One
// thread owns one row, held in registers; L11 sits in LDS transposed and every
// read of it is a wave-wide broadcast.  The sums are accumulated column-sweep
// style (after x[c] is final, every later x[q] -= x[c] * L11[q,c]), so each
// step issues up to 63 independent FMAs instead of one dependent chain.
//
// trsv_block: the diagonal-block step of the blocked triangular solve: one
// thread owns one right-hand-side column, held in registers, swept exactly as
// triangular.rs:113-125 sweeps it (divide by the pivot, then update the rest of
// the column), restricted to b <= 64 rows.
#include "llh_common.h"
#include "llh_kernels.h"
#include "k_trsm_impl.h"

namespace llh {
namespace {

template <typename T, bool VEC>
__global__ __launch_bounds__(64) void trsm_rows_kernel(int m, int b, T* __restrict__ X, int64_t ldx,
                                                       const T* __restrict__ L, int64_t ldl,
                                                       const int32_t* __restrict__ gate) {
  if (gate != nullptr && *gate != kNoFail) return;
  __shared__ __attribute__((aligned(16))) T Lt[kNB * kNB];  // Lt[c * kNB + q] = L11[q, c]
  const int tid = threadIdx.x;
  for (int e = tid; e < kNB * kNB; e += 64) {
    const int q = e / kNB, c = e % kNB;  // coalesced read of row q
    Lt[c * kNB + q] = (q < b && c <= q) ? L[(int64_t)q * ldl + c] : T(0);
  }
  __syncthreads();
  const int row = blockIdx.x * 64 + tid;
  if (row >= m) return;
  T* xr = X + (int64_t)row * ldx;
  T x[kNB];
  if constexpr (VEC && sizeof(T) == 8) {
#pragma unroll
    for (int c = 0; c < kNB; c += 2) {
      if (c < b) {
        const double2 v = *reinterpret_cast<const double2*>(xr + c);
        x[c] = v.x;
        x[c + 1] = v.y;
      } else {
        x[c] = T(0);
        x[c + 1] = T(0);
      }
    }
  } else {
#pragma unroll
    for (int c = 0; c < kNB; ++c) x[c] = (c < b) ? xr[c] : T(0);
  }
#pragma unroll
  for (int c = 0; c < kNB; ++c) {
    if (c < b) {
      x[c] = x[c] / Lt[c * kNB + c];
#pragma unroll
      for (int q = c + 1; q < kNB; ++q) x[q] -= x[c] * Lt[c * kNB + q];
    }
  }
#pragma unroll
  for (int c = 0; c < kNB; ++c)
    if (c < b) xr[c] = x[c];
}

template <typename T, bool UPPER_OP>
__global__ __launch_bounds__(64) void trsv_block_kernel(int b, bool trans, const T* __restrict__ Tm,
                                                        int64_t ldt, int m, T* __restrict__ Y, int64_t ldy,
                                                        const int32_t* __restrict__ gate) {
  if (gate != nullptr && *gate != kNoFail) return;
  // St[c * kNB + r] = op(T)[r, c], only op(T)'s triangle: column c of op(T) is
  // contiguous, which is what the column sweep reads.
  __shared__ __attribute__((aligned(16))) T St[kNB * kNB];
  const int tid = threadIdx.x;
  for (int e = tid; e < kNB * kNB; e += 64) {
    const int i = e / kNB, j = e % kNB;               // reads T[i, j], coalesced over j
    const int r = trans ? j : i, c = trans ? i : j;   // T[i, j] is op(T)[r, c]
    const bool in = r < b && c < b && (UPPER_OP ? c >= r : c <= r);
    St[c * kNB + r] = in ? Tm[(int64_t)i * ldt + j] : T(0);
  }
  __syncthreads();
  const int col = blockIdx.x * 64 + tid;
  if (col >= m) return;
  T y[kNB];
#pragma unroll
  for (int r = 0; r < kNB; ++r) y[r] = (r < b) ? Y[(int64_t)r * ldy + col] : T(0);
  if constexpr (!UPPER_OP) {
#pragma unroll
    for (int r = 0; r < kNB; ++r) {
      if (r < b) {
        y[r] = y[r] / St[r * kNB + r];
#pragma unroll
        for (int p = r + 1; p < kNB; ++p) y[p] -= y[r] * St[r * kNB + p];
      }
    }
  } else {
#pragma unroll
    for (int r = kNB - 1; r >= 0; --r) {
      if (r < b) {
        y[r] = y[r] / St[r * kNB + r];
#pragma unroll
        for (int p = 0; p < r; ++p) y[p] -= y[r] * St[r * kNB + p];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kNB; ++r)
    if (r < b) Y[(int64_t)r * ldy + col] = y[r];
}

// trtri64: one workgroup per diagonal block (trtri64_block, k_trsm_impl.h).
template <typename TS, typename T>
__global__ __launch_bounds__(256) void trtri64_kernel(int n, const TS* __restrict__ A, int64_t lda, bool trans,
                                                      bool upper, TS* __restrict__ out,
                                                      const int32_t* __restrict__ gate) {
  if (gate != nullptr && *gate != kNoFail) return;
  __shared__ T Ms[kNB * (kNB + 1)];
  __shared__ T Xs[kNB * (kNB + 1)];
  __shared__ T Ts[kNB * (kNB + 1)];
  const int o = blockIdx.x * kNB;
  trtri64_block<TS, TS, T>(min(kNB, n - o), A + (int64_t)o * lda + o, lda, trans, upper,
                       out + (int64_t)blockIdx.x * kNB * kNB, Ms, Xs, Ts);
}

// ---------------------------------------------------------------------------
// trsm_persistent: op(A) X = B in place for up to kSolveChunk right-hand-side
// columns per workgroup, one launch for the whole solve.
//
// Workgroups take tickets in launch order; ticket t is (position p, column chunk
// q).  Position p is block row w = p (forward, op(A) lower) or w = nb-1-p
// (backward).  The workgroup accumulates sum_j op(A)_wj X_j over the blocks j
// that precede w in solve order, then forms X_w = inv(op(A)_ww) (B_w - sum) with
// the precomputed inverse, stores it and raises flag (w, q).  A workgroup only
// ever waits for positions taken before its own, so the chain always advances
// whatever the residency.  Flags hold the epoch of the call that raised them, so
// they never need clearing.
//
// The walk is throughput-bound for every workgroup behind the frontier, so it
// keeps synchronisation off the per-block path: wave v owns columns
// [16v, 16v + 16) of each 64 x 64 tile, and each lane (row r) holds its 16 tile
// values in registers, loaded from global memory one block ahead; X_j goes
// through a double-buffered LDS slab with one barrier per block; and flags are
// checked in windows: one poll and one acquire fence cover every block whose
// flag is already raised.  The wave partial sums are reduced once.
//
// Only the last dependencies lie on the chain from flag to flag, so the
// workgroup does everything else before it waits for them.  With l1 and l2 the
// blocks solved just before w (l1 last), it forms at the start
// G_i = inv(op(A)_ww) op(A)_w,li (in LDS) and B_w; after the other blocks,
// Z = inv(op(A)_ww) (B_w - sum_{j != l1, l2} op(A)_wj X_j); once X_l2 is
// published Z -= G_2 X_l2, and once X_l1 is, X_w = Z - G_1 X_l1.  So a flag is
// followed by one 64 x 64 product and one reduction before the next is raised,
// and the rest of the block's work overlaps the two steps before it.
// The wait is bounded: past spin_limit polls (or once another workgroup has
// given up) the workgroup records an error and returns without raising its flag.
// A negative spin_limit makes every wait give up at once (a host-forced failure
// for tests of the error path; every workgroup still exits normally).  With a
// gate (the factor's status word) that holds a failed row, every workgroup
// returns at once and B is left untouched.
constexpr int kSolveChunk = kSolveChunkHost;
constexpr int kFlagWindow = 64;  // flags checked ahead per wait

// f64: the three 64 x 64 tiles and the work area take 139 KB of LDS, one
// workgroup per CU; f32: two.
template <typename T, bool VEC>
__global__ __launch_bounds__(256, sizeof(T) == 8 ? 1 : 2) void trsm_persistent_kernel(
    int n, int m, const T* __restrict__ A, int64_t lda, bool trans, bool lower_op, const T* __restrict__ Dinv,
    bool dinv_t, T* B, int64_t ldb, int32_t* ticket, int32_t* flags, int32_t epoch, int32_t* err,
    int64_t spin_limit, const int32_t* __restrict__ gate) {
  if (gate != nullptr && *gate != kNoFail) return;
  constexpr int MC = kSolveChunk, KW = kNB / 4, S = kNB + 1, SL = kNB * MC;
  static_assert(MC == 16 && KW == 16, "one 16-column chunk, 16 tile columns per wave");
  static_assert(5 * SL >= kNB * S, "a staged tile fits the work area");
  // Work area, by phase: a staged tile of op(A) (64 x S) for G; the
  // double-buffered X_j slabs [0, 2 SL) of the walk; the 4 wave partials
  // [0, 4 SL) and one slab [4 SL, 5 SL) of the reductions.
  __shared__ __attribute__((aligned(16))) T Wk[5 * SL];
  __shared__ T Dl[kNB * S];              // inv(op(A)_ww)
  __shared__ T Gs[2][kNB * S];           // G_1, G_2
  __shared__ int s_ticket, s_ready, s_abort;
  T* const Part = Wk;
  T* const Slab = Wk + 4 * SL;
  const int tid = threadIdx.x;
  const int nb = (n + kNB - 1) / kNB, nq = (m + MC - 1) / MC;
  if (tid == 0) {
    s_ticket = atomicAdd(ticket, 1);
    s_abort = 0;
  }
  __syncthreads();
  const int tk = s_ticket;
  if (tk >= nb * nq) return;
  const int p = tk / nq, q = tk % nq;
  const int w = lower_op ? p : nb - 1 - p;
  const int w0 = w * kNB, bw = min(kNB, n - w0);
  const int c0 = q * MC, mc = min(MC, m - c0);
  const int r = tid & 63, wv = tid >> 6;
  auto jof = [&](int st) { return lower_op ? st : nb - 1 - st; };

  // the inverse of the diagonal block does not depend on any flag: stage it first
  {
    const T* D = Dinv + (int64_t)w * kNB * kNB;
#pragma unroll
    for (int g = 0; g < 16; ++g) {
      const int e = tid + 256 * g, i = e / kNB, jj = e % kNB;
      Dl[(dinv_t ? jj : i) * S + (dinv_t ? i : jj)] = D[e];
    }
  }
  // B_w[r][4 wv .. 4 wv + 3], the entries this thread finishes
  T bwr[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int c = 4 * wv + e;
    bwr[e] = (r < bw && c < mc) ? B[(int64_t)(w0 + r) * ldb + c0 + c] : T(0);
  }

  // this lane's 16 values of op(A)[w0 + r, j0 + 16 wv + kk] for block j
  typedef T vec_t __attribute__((ext_vector_type(16 / sizeof(T))));
  constexpr int VL = 16 / sizeof(T);
  auto fetch = [&](int st, T* out) {
    const int j = jof(st), j0 = j * kNB, bj = min(kNB, n - j0), k0 = KW * wv;
    if (!trans) {
      const T* row = A + (int64_t)(w0 + min(r, bw - 1)) * lda + j0 + k0;
      if (VEC && k0 + KW <= bj) {
#pragma unroll
        for (int v = 0; v < KW / VL; ++v) {
          const vec_t x = *reinterpret_cast<const vec_t*>(row + v * VL);
#pragma unroll
          for (int e = 0; e < VL; ++e) out[v * VL + e] = x[e];
        }
      } else {
#pragma unroll
        for (int kk = 0; kk < KW; ++kk) out[kk] = (k0 + kk < bj) ? row[kk] : T(0);
      }
    } else {
      const T* col = A + (int64_t)j0 * lda + w0 + min(r, bw - 1);
#pragma unroll
      for (int kk = 0; kk < KW; ++kk) out[kk] = (k0 + kk < bj) ? col[(int64_t)(k0 + kk) * lda] : T(0);
    }
  };
  // X_j rows [0, 64) x chunk columns: 4 values per thread (row e / 16, column e % 16)
  auto loady = [&](int st, T* out) {
    const int j0 = jof(st) * kNB, bj = min(kNB, n - j0);
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int e = tid + 256 * g, k = e / MC, c = e % MC;
      out[g] = (k < bj && c < mc) ? B[(int64_t)(j0 + k) * ldb + c0 + c] : T(0);
    }
  };
  // wait until block st's flag is raised, then take every later flag already
  // raised (up to the window); false when the wait gave up
  int ready = 0;
  auto wait_for = [&](int st) -> bool {
    if (tid == 0) {
      int64_t it = 0;
      const int32_t* f = &flags[jof(st) * nq + q];
      if (spin_limit < 0) {
        atomicExch(err, 1);
        s_abort = 1;
      }
      while (!s_abort && __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
        if (++it > spin_limit || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
          atomicExch(err, 1);
          s_abort = 1;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      int t2 = st + 1;
      while (t2 < p && t2 < st + kFlagWindow &&
             __hip_atomic_load(&flags[jof(t2) * nq + q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch)
        ++t2;
      s_ready = t2;
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    if (s_abort) return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    ready = s_ready;
    return true;
  };

  // Gs[i][r][16 wv + kk] = sum_k inv(op(A)_ww)[r][k] op(A)_w,l[k][16 wv + kk], l at position st
  auto make_g = [&](int st, T* gs) {
    T t[KW];
    fetch(st, t);
#pragma unroll
    for (int kk = 0; kk < KW; ++kk) Wk[r * S + KW * wv + kk] = t[kk];
    __syncthreads();
    T g[KW];
#pragma unroll
    for (int kk = 0; kk < KW; ++kk) g[kk] = T(0);
#pragma unroll 4
    for (int k = 0; k < kNB; ++k) {
      const T d = Dl[r * S + k];
#pragma unroll
      for (int kk = 0; kk < KW; ++kk) g[kk] += d * Wk[k * S + KW * wv + kk];
    }
#pragma unroll
    for (int kk = 0; kk < KW; ++kk) gs[r * S + KW * wv + kk] = g[kk];
    __syncthreads();
  };
  const int l1 = p - 1, l2 = p - 2;  // positions of the last two dependencies
  if (p > 0) make_g(l1, Gs[0]);
  if (p > 1) make_g(l2, Gs[1]);

  // the walk over the blocks before l2
  const int nwalk = max(p - 2, 0);
  T acc[MC];
#pragma unroll
  for (int c = 0; c < MC; ++c) acc[c] = T(0);
  {
    T a_cur[KW], a_nxt[KW], y_cur[4], y_nxt[4];
    bool y_pre = false;
    if (nwalk > 0) fetch(0, a_cur);
    for (int st = 0; st < nwalk; ++st) {
      if (st >= ready) {
        if (!wait_for(st)) return;
        y_pre = false;
      }
      if (y_pre) {
#pragma unroll
        for (int e = 0; e < 4; ++e) y_cur[e] = y_nxt[e];
      } else {
        loady(st, y_cur);
      }
      T* ys = Wk + (st & 1) * SL;
#pragma unroll
      for (int e = 0; e < 4; ++e) ys[tid + 256 * e] = y_cur[e];
      __syncthreads();
      if (st + 1 < nwalk) fetch(st + 1, a_nxt);
      y_pre = st + 1 < nwalk && st + 1 < ready;
      if (y_pre) loady(st + 1, y_nxt);
      const T* yk = ys + KW * wv * MC;
#pragma unroll
      for (int kk = 0; kk < KW; ++kk) {
        const T a = a_cur[kk];
#pragma unroll
        for (int c = 0; c < MC; ++c) acc[c] += a * yk[kk * MC + c];
      }
#pragma unroll
      for (int kk = 0; kk < KW; ++kk) a_cur[kk] = a_nxt[kk];
    }
  }

  // Z = inv(op(A)_ww) (B_w - sum): reduce the wave partials into Slab, then
  // lane r, columns 4 wv .. 4 wv + 3
  __syncthreads();
#pragma unroll
  for (int c = 0; c < MC; ++c) Part[wv * SL + r * MC + c] = acc[c];
  __syncthreads();
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int c = 4 * wv + e, o = r * MC + c;
    const T sum = Part[o] + Part[SL + o] + Part[2 * SL + o] + Part[3 * SL + o];
    Slab[o] = (r < bw && c < mc) ? bwr[e] - sum : T(0);
  }
  __syncthreads();
  T x[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll 8
  for (int k = 0; k < kNB; ++k) {
    const T d = Dl[r * S + k];
#pragma unroll
    for (int e = 0; e < 4; ++e) x[e] += d * Slab[k * MC + 4 * wv + e];
  }

  // x -= G X_l for the dependency at position st, once its flag is raised
  auto apply_g = [&](int st, const T* gs) -> bool {
    if (st >= ready && !wait_for(st)) return false;
    T y[4];
    loady(st, y);
    __syncthreads();  // every wave is done reading Slab and Part
#pragma unroll
    for (int e = 0; e < 4; ++e) Slab[tid + 256 * e] = y[e];
    __syncthreads();
    T gy[MC];
#pragma unroll
    for (int c = 0; c < MC; ++c) gy[c] = T(0);
    const T* yk = Slab + KW * wv * MC;
#pragma unroll
    for (int kk = 0; kk < KW; ++kk) {
      const T gv = gs[r * S + KW * wv + kk];
#pragma unroll
      for (int c = 0; c < MC; ++c) gy[c] += gv * yk[kk * MC + c];
    }
#pragma unroll
    for (int c = 0; c < MC; ++c) Part[wv * SL + r * MC + c] = gy[c];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int o = r * MC + 4 * wv + e;
      x[e] -= Part[o] + Part[SL + o] + Part[2 * SL + o] + Part[3 * SL + o];
    }
    return true;
  };
  if (p > 1 && !apply_g(l2, Gs[1])) return;
  if (p > 0 && !apply_g(l1, Gs[0])) return;
#pragma unroll
  for (int e = 0; e < 4; ++e)
    if (r < bw && 4 * wv + e < mc) B[(int64_t)(w0 + r) * ldb + c0 + 4 * wv + e] = x[e];
  // Each wave releases its own stores at agent scope (a release fence: the
  // seq_cst __threadfence would also invalidate the L2 after every block);
  // thread 0's release store of the flag follows the barrier.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (tid == 0) __hip_atomic_store(&flags[w * nq + q], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace

template <typename T>
hipError_t launch_trsm_rows(int m, int b, T* X, int64_t ldx, const T* L, int64_t ldl, const int32_t* gate,
                            hipStream_t s) {
  if (m <= 0 || b <= 0) return hipSuccess;
  if (b > kNB) return hipErrorInvalidValue;
  const dim3 grid((m + 63) / 64);
  const bool vec = (reinterpret_cast<uintptr_t>(X) % 16 == 0) && (ldx % 2 == 0) && (b % 2 == 0);
  if (vec)
    hipLaunchKernelGGL((trsm_rows_kernel<T, true>), grid, dim3(64), 0, s, m, b, X, ldx, L, ldl, gate);
  else
    hipLaunchKernelGGL((trsm_rows_kernel<T, false>), grid, dim3(64), 0, s, m, b, X, ldx, L, ldl, gate);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_trsv_block(bool upper_op, bool trans, int b, const T* Tm, int64_t ldt, int m, T* Y,
                             int64_t ldy, const int32_t* gate, hipStream_t s) {
  if (m <= 0 || b <= 0) return hipSuccess;
  if (b > kNB) return hipErrorInvalidValue;
  const dim3 grid((m + 63) / 64);
  if (upper_op)
    hipLaunchKernelGGL((trsv_block_kernel<T, true>), grid, dim3(64), 0, s, b, trans, Tm, ldt, m, Y, ldy, gate);
  else
    hipLaunchKernelGGL((trsv_block_kernel<T, false>), grid, dim3(64), 0, s, b, trans, Tm, ldt, m, Y, ldy, gate);
  return hipGetLastError();
}

template <typename TS, typename T>
hipError_t launch_trtri_blocks(int n, const TS* A, int64_t lda, bool trans, bool upper, TS* out,
                               const int32_t* gate, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL((trtri64_kernel<TS, T>), dim3((n + kNB - 1) / kNB), dim3(256), 0, s, n, A, lda, trans,
                     upper, out, gate);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_trsm_persistent(int n, int m, const T* A, int64_t lda, bool trans, bool lower_op,
                                  const T* Dinv, bool dinv_t, T* B, int64_t ldb, int32_t* ticket,
                                  int32_t* flags, int32_t epoch, int32_t* err, int64_t spin_limit,
                                  const int32_t* gate, hipStream_t s) {
  if (n <= 0 || m <= 0) return hipSuccess;
  const int nb = (n + kNB - 1) / kNB, nq = (m + kSolveChunk - 1) / kSolveChunk;
  hipError_t e = launch_fill_i32(ticket, 0, 1, s);
  if (e != hipSuccess) return e;
  const bool vec = (reinterpret_cast<uintptr_t>(A) % 16 == 0) && ((lda * sizeof(T)) % 16 == 0);
  if (vec)
    hipLaunchKernelGGL((trsm_persistent_kernel<T, true>), dim3(nb * nq), dim3(256), 0, s, n, m, A, lda, trans,
                       lower_op, Dinv, dinv_t, B, ldb, ticket, flags, epoch, err, spin_limit, gate);
  else
    hipLaunchKernelGGL((trsm_persistent_kernel<T, false>), dim3(nb * nq), dim3(256), 0, s, n, m, A, lda, trans,
                       lower_op, Dinv, dinv_t, B, ldb, ticket, flags, epoch, err, spin_limit, gate);
  return hipGetLastError();
}

template hipError_t launch_trtri_blocks<double, double>(int, const double*, int64_t, bool, bool, double*,
                                                        const int32_t*, hipStream_t);
template hipError_t launch_trtri_blocks<float, float>(int, const float*, int64_t, bool, bool, float*,
                                                      const int32_t*, hipStream_t);
template hipError_t launch_trtri_blocks<float, double>(int, const float*, int64_t, bool, bool, float*,
                                                       const int32_t*, hipStream_t);
template hipError_t launch_trsm_persistent<double>(int, int, const double*, int64_t, bool, bool, const double*,
                                                   bool, double*, int64_t, int32_t*, int32_t*, int32_t, int32_t*,
                                                   int64_t, const int32_t*, hipStream_t);
template hipError_t launch_trsm_persistent<float>(int, int, const float*, int64_t, bool, bool, const float*, bool,
                                                  float*, int64_t, int32_t*, int32_t*, int32_t, int32_t*, int64_t,
                                                  const int32_t*, hipStream_t);
template hipError_t launch_trsm_rows<double>(int, int, double*, int64_t, const double*, int64_t,
                                             const int32_t*, hipStream_t);
template hipError_t launch_trsm_rows<float>(int, int, float*, int64_t, const float*, int64_t, const int32_t*,
                                            hipStream_t);
template hipError_t launch_trsv_block<double>(bool, bool, int, const double*, int64_t, int, double*, int64_t,
                                              const int32_t*, hipStream_t);
template hipError_t launch_trsv_block<float>(bool, bool, int, const float*, int64_t, int, float*, int64_t,
                                             const int32_t*, hipStream_t);

}  // namespace llh

// Device code of the 64 x 64 triangular-block inverse, shared by the trtri
// kernel (k_trsm.hip) and the grouped batch kernels (k_batch.hip).
#pragma once

#include "llh_common.h"

namespace llh {

// trtri64: inverses of the 64 x 64 diagonal blocks of op(A) (the last block may
// be smaller; it is padded with an identity tail).  One workgroup per block.
// The block is copied to LDS in lower-triangular form M (op(A)'s block, or its
// transpose when that block is upper triangular); inv(M) is built by recursive
// doubling: the four 16 x 16 diagonal blocks are inverted by substitution (one
// thread per column), then pairs are combined with
//   [P 0; C Q]^-1 = [P^-1 0; -Q^-1 C P^-1  Q^-1]
// at 32 and at 64.  The result is written row-major, transposed back when op(A)'s
// block is upper, with exact zeros outside the triangle.
// A points at the block's origin, b <= 64 is its order; dst receives the 64 x 64
// row-major inverse.  Ms, Xs, Ts are LDS arrays of 64 x 65 elements.
// TA: element type of A (global or LDS, read through a generic pointer); TD:
// element type of dst; T: arithmetic type.  A may alias Ms when trans == upper
// (each element is then read and rewritten by the same thread).
template <typename TA, typename TD, typename T>
__device__ void trtri64_block(int b, const TA* A, int64_t lda, bool trans, bool upper, TD* __restrict__ dst,
                              T* Ms, T* Xs, T* Ts) {
  constexpr int S = kNB + 1;
  const int tid = threadIdx.x;
  const int o = 0;
  // Element (r, c) of M, r >= c: op(A)[o+r, o+c] (lower) or op(A)[o+c, o+r] (upper);
  // with op(A) = A^T when trans, M[r][c] is A[o+c][o+r] exactly when upper != trans.
  const bool flip = (upper != trans);
  {
    TA v[16];  // all 16 loads of a thread in flight before any store
#pragma unroll
    for (int g = 0; g < 16; ++g) {
      const int e = tid + 256 * g, i = e / kNB, j = e % kNB;  // coalesced over j in A's storage
      const int r = flip ? j : i, c = flip ? i : j;
      v[g] = (r < b && c < b && c <= r) ? A[(int64_t)(o + i) * lda + (o + j)] : TA(0);
    }
#pragma unroll
    for (int g = 0; g < 16; ++g) {
      const int e = tid + 256 * g, i = e / kNB, j = e % kNB;
      const int r = flip ? j : i, c = flip ? i : j;
      Ms[r * S + c] = (r < b && c < b) ? T(v[g]) : (r == c ? T(1) : T(0));
      Xs[r * S + c] = T(0);
    }
  }
  __syncthreads();
  // 16 x 16 diagonal blocks: thread (blk, c) builds column c of inv(D_blk).
  if (tid < kNB) {
    const int base = (tid >> 4) * 16, c = tid & 15;
    T x[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      T s = T(0);
#pragma unroll
      for (int p = 0; p < r; ++p) s += (p >= c) ? Ms[(base + r) * S + base + p] * x[p] : T(0);
      const T d = Ms[(base + r) * S + base + r];
      x[r] = (r < c) ? T(0) : (r == c ? T(1) / d : -s / d);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) Xs[(base + r) * S + base + c] = x[r];
  }
  __syncthreads();
  // Combine at h = 16 (two pairs) and h = 32 (one pair).
#pragma unroll
  for (int h = 16; h <= 32; h *= 2) {
    const int pairs = kNB / (2 * h);
    // T = C * P^-1, C = M[base+h.., base..base+h), P^-1 = X[base.., base..]
    for (int e = tid; e < pairs * h * h; e += 256) {
      const int pr = e / (h * h), rem = e % (h * h), i = rem / h, j = rem % h;
      const int base = pr * 2 * h;
      T s = T(0);
      for (int k = j; k < h; ++k) s += Ms[(base + h + i) * S + base + k] * Xs[(base + k) * S + base + j];
      Ts[(base + h + i) * S + base + j] = s;
    }
    __syncthreads();
    // X21 = -Q^-1 * T, Q^-1 = X[base+h.., base+h..]
    for (int e = tid; e < pairs * h * h; e += 256) {
      const int pr = e / (h * h), rem = e % (h * h), i = rem / h, j = rem % h;
      const int base = pr * 2 * h;
      T s = T(0);
      for (int k = 0; k <= i; ++k) s += Xs[(base + h + i) * S + base + h + k] * Ts[(base + h + k) * S + base + j];
      Xs[(base + h + i) * S + base + j] = -s;
    }
    __syncthreads();
  }
  for (int e = tid; e < kNB * kNB; e += 256) {
    const int r = e / kNB, c = e % kNB;
    dst[e] = TD(upper ? Xs[c * S + r] : Xs[r * S + c]);
  }
}

}  // namespace llh

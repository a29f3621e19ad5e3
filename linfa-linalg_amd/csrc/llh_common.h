// Shared definitions for the linfa-linalg-hip kernels and host drivers (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace llh {

// Status word written by the factor kernels: the first failing row of the
// `d <= 0` pivot test (cholesky.rs:69), or kNoFail.
constexpr int32_t kNoFail = 0x7fffffff;

// Block sizes of the blocked right-looking POTRF: kNB is the width of the
// diagonal block factored in LDS and of the panel solved row by row; the outer
// panel width, the rank of the trailing SYRK update on MFMA, is kNBO, or
// kNBO_LARGE for an order of at least kNBO_LARGE_N when the SYRK runs on f64 or
// f32 operands (outer_width()).  Wider panels give each output tile more MFMA
// k-slabs per read-modify-write of C but a longer panel chain.  Measured solves:
// n = 32768 f64 229 / 220 / 223 / 226 ms at 512 / 1024 / 1536 / 2048; n = 16384
// f64 41.6 / 42.5 ms at 512 / 1024; the f16-operand factor, whose SYRK is fast and
// whose panel chain sets the pace, 87 / 95 ms at 512 / 1024 (n = 32768), and later
// FP16+IR 88.5 / 78.1 / 70.8 ms at 256 / 384 / 512 (narrower panels pass over the
// f32 trailing matrix more often); the FP32
// refined solve 166 / 158 ms at 512 / 1024 (n = 32768: its rank-512 SYRKs cost
// more than the shorter panel chain saves), later 134.6 / 133.4 / 134.6 ms at
// 768 / 1024 / 1536.
constexpr int kNB = 64;
constexpr int kNBO = 512;
constexpr int kNBO_LARGE = 1024;
constexpr int kNBO_LARGE_N = 24576;
inline int outer_width(int n, bool half_syrk) { return (!half_syrk && n >= kNBO_LARGE_N) ? kNBO_LARGE : kNBO; }

// Right-hand-side columns per workgroup of the one-launch triangular solve, and
// the bound on its flag waits (2^22 polls: well over 0.1 s, far above any real wait).
constexpr int kSolveChunkHost = 16;
constexpr int64_t kSpinLimit = 1ll << 22;

// Largest n the single-workgroup LDS kernel takes (the packed lower triangle of
// an f64 matrix must fit in 160 KiB of LDS with room to spare).
constexpr int kSmallMaxD = 128;
constexpr int kSmallMaxS = 176;

// Exact-order arithmetic.  The files that use these helpers are compiled with
// -ffp-contract=off and open with `#pragma clang fp contract(off)`, so every
// `s += x*y` stays a separately rounded multiply and add, as the reference's
// compiled Rust is.  IEEE division and square root are correctly rounded on
// gfx950 for f64 and, without fast-math flags, for f32.
template <typename T>
__device__ __forceinline__ T sqrt_rn(T a);
template <>
__device__ __forceinline__ double sqrt_rn<double>(double a) { return __builtin_sqrt(a); }
template <>
__device__ __forceinline__ float sqrt_rn<float>(float a) { return __builtin_sqrtf(a); }

}  // namespace llh

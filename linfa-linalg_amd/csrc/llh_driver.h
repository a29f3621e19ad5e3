// Host drivers (driver.hip): launch sequences of the factor and the solves.
#pragma once

#include "llh_common.h"
#include "llh_profile.h"

namespace llh {

// Enqueues the in-place lower Cholesky of a (row-major, lda) on stream s.
// *info receives kNoFail or the first row whose pivot fails `d <= 0`.
template <typename T>
hipError_t potrf_enqueue(T* a, int64_t lda, int n, int32_t* info, bool zero_upper, hipStream_t s,
                         Profiler* prof = nullptr);

// Enqueues op(A) X = B in place (B row-major n x m); upper_op names the triangle
// of op(A) and trans != 0 means op(A) = A^T.
template <typename T>
hipError_t trsm_enqueue(bool upper_op, bool trans, int n, const T* a, int64_t lda, int m, T* b, int64_t ldb,
                        hipStream_t s, Profiler* prof = nullptr);

}  // namespace llh

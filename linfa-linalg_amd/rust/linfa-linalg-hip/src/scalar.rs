//! The element types the library computes in, mapped to its `_d` / `_s` entry
//! points.
use ndarray::NdFloat;
use std::os::raw::c_int;

use crate::ffi::{self, llh_context as Ctx, llh_error as Err, llh_status, llh_uplo};

mod sealed {
    pub trait Sealed {}
    impl Sealed for f32 {}
    impl Sealed for f64 {}
}

/// f32 or f64.  The methods are the C ABI's entry points for that type; their
/// pointer / dimension / stride arguments follow include/llh.h.
#[allow(clippy::too_many_arguments, clippy::missing_safety_doc)]
pub trait HipScalar: NdFloat + sealed::Sealed {
    unsafe fn cholesky_inplace_dirty(c: *mut Ctx, rows: usize, cols: usize, a: *mut Self, rs: isize, cs: isize,
                                     e: *mut Err) -> llh_status;
    unsafe fn cholesky_inplace(c: *mut Ctx, rows: usize, cols: usize, a: *mut Self, rs: isize, cs: isize,
                               e: *mut Err) -> llh_status;
    unsafe fn triangular_inplace(c: *mut Ctx, rows: usize, cols: usize, a: *mut Self, rs: isize, cs: isize,
                                 uplo: llh_uplo, e: *mut Err) -> llh_status;
    unsafe fn is_triangular(c: *mut Ctx, rows: usize, cols: usize, a: *const Self, rs: isize, cs: isize,
                            uplo: llh_uplo, out: *mut c_int, e: *mut Err) -> llh_status;
    unsafe fn solve_triangular_inplace(c: *mut Ctx, rows: usize, cols: usize, a: *const Self, rs: isize, cs: isize,
                                       brows: usize, bcols: usize, b: *mut Self, brs: isize, bcs: isize,
                                       uplo: llh_uplo, e: *mut Err) -> llh_status;
    unsafe fn solve_triangular_system(c: *mut Ctx, rows: usize, cols: usize, a: *const Self, rs: isize, cs: isize,
                                      diag: *const Self, dstride: isize, abs_diag: c_int, brows: usize,
                                      bcols: usize, b: *mut Self, brs: isize, bcs: isize, uplo: llh_uplo,
                                      e: *mut Err) -> llh_status;
    unsafe fn solvec_inplace(c: *mut Ctx, rows: usize, cols: usize, a: *mut Self, rs: isize, cs: isize,
                             brows: usize, bcols: usize, b: *mut Self, brs: isize, bcs: isize, e: *mut Err)
                             -> llh_status;
    unsafe fn invc_inplace(c: *mut Ctx, rows: usize, cols: usize, a: *mut Self, rs: isize, cs: isize,
                           out: *mut Self, ors: isize, ocs: isize, e: *mut Err) -> llh_status;
    unsafe fn dev_potrf(c: *mut Ctx, n: usize, a: *mut Self, lda: usize, zero_upper: c_int, e: *mut Err)
                        -> llh_status;
    unsafe fn dev_solve_triangular(c: *mut Ctx, uplo: llh_uplo, trans: c_int, n: usize, a: *const Self,
                                   lda: usize, m: usize, b: *mut Self, ldb: usize, e: *mut Err) -> llh_status;
    unsafe fn dev_solvec(c: *mut Ctx, n: usize, a: *mut Self, lda: usize, m: usize, b: *mut Self, ldb: usize,
                         e: *mut Err) -> llh_status;
    unsafe fn dev_invc(c: *mut Ctx, n: usize, a: *mut Self, lda: usize, out: *mut Self, ldo: usize, e: *mut Err)
                       -> llh_status;
}

macro_rules! hip_scalar {
    ($t:ty, $cd:ident, $c:ident, $tri:ident, $istri:ident, $st:ident, $ss:ident, $sv:ident, $inv:ident,
     $dp:ident, $dst:ident, $dsv:ident, $dinv:ident) => {
        impl HipScalar for $t {
            unsafe fn cholesky_inplace_dirty(c: *mut Ctx, rows: usize, cols: usize, a: *mut Self, rs: isize,
                                             cs: isize, e: *mut Err) -> llh_status {
                ffi::$cd(c, rows, cols, a, rs, cs, e)
            }
            unsafe fn cholesky_inplace(c: *mut Ctx, rows: usize, cols: usize, a: *mut Self, rs: isize, cs: isize,
                                       e: *mut Err) -> llh_status {
                ffi::$c(c, rows, cols, a, rs, cs, e)
            }
            unsafe fn triangular_inplace(c: *mut Ctx, rows: usize, cols: usize, a: *mut Self, rs: isize,
                                         cs: isize, uplo: llh_uplo, e: *mut Err) -> llh_status {
                ffi::$tri(c, rows, cols, a, rs, cs, uplo, e)
            }
            unsafe fn is_triangular(c: *mut Ctx, rows: usize, cols: usize, a: *const Self, rs: isize, cs: isize,
                                    uplo: llh_uplo, out: *mut c_int, e: *mut Err) -> llh_status {
                ffi::$istri(c, rows, cols, a, rs, cs, uplo, out, e)
            }
            unsafe fn solve_triangular_inplace(c: *mut Ctx, rows: usize, cols: usize, a: *const Self, rs: isize,
                                               cs: isize, brows: usize, bcols: usize, b: *mut Self, brs: isize,
                                               bcs: isize, uplo: llh_uplo, e: *mut Err) -> llh_status {
                ffi::$st(c, rows, cols, a, rs, cs, brows, bcols, b, brs, bcs, uplo, e)
            }
            unsafe fn solve_triangular_system(c: *mut Ctx, rows: usize, cols: usize, a: *const Self, rs: isize,
                                              cs: isize, diag: *const Self, dstride: isize, abs_diag: c_int,
                                              brows: usize, bcols: usize, b: *mut Self, brs: isize, bcs: isize,
                                              uplo: llh_uplo, e: *mut Err) -> llh_status {
                ffi::$ss(c, rows, cols, a, rs, cs, diag, dstride, abs_diag, brows, bcols, b, brs, bcs, uplo, e)
            }
            unsafe fn solvec_inplace(c: *mut Ctx, rows: usize, cols: usize, a: *mut Self, rs: isize, cs: isize,
                                     brows: usize, bcols: usize, b: *mut Self, brs: isize, bcs: isize,
                                     e: *mut Err) -> llh_status {
                ffi::$sv(c, rows, cols, a, rs, cs, brows, bcols, b, brs, bcs, e)
            }
            unsafe fn invc_inplace(c: *mut Ctx, rows: usize, cols: usize, a: *mut Self, rs: isize, cs: isize,
                                   out: *mut Self, ors: isize, ocs: isize, e: *mut Err) -> llh_status {
                ffi::$inv(c, rows, cols, a, rs, cs, out, ors, ocs, e)
            }
            unsafe fn dev_potrf(c: *mut Ctx, n: usize, a: *mut Self, lda: usize, zero_upper: c_int, e: *mut Err)
                                -> llh_status {
                ffi::$dp(c, n, a, lda, zero_upper, e)
            }
            unsafe fn dev_solve_triangular(c: *mut Ctx, uplo: llh_uplo, trans: c_int, n: usize, a: *const Self,
                                           lda: usize, m: usize, b: *mut Self, ldb: usize, e: *mut Err)
                                           -> llh_status {
                ffi::$dst(c, uplo, trans, n, a, lda, m, b, ldb, e)
            }
            unsafe fn dev_solvec(c: *mut Ctx, n: usize, a: *mut Self, lda: usize, m: usize, b: *mut Self,
                                 ldb: usize, e: *mut Err) -> llh_status {
                ffi::$dsv(c, n, a, lda, m, b, ldb, e)
            }
            unsafe fn dev_invc(c: *mut Ctx, n: usize, a: *mut Self, lda: usize, out: *mut Self, ldo: usize,
                               e: *mut Err) -> llh_status {
                ffi::$dinv(c, n, a, lda, out, ldo, e)
            }
        }
    };
}

hip_scalar!(f64, llh_cholesky_inplace_dirty_d, llh_cholesky_inplace_d, llh_triangular_inplace_d,
            llh_is_triangular_d, llh_solve_triangular_inplace_d, llh_solve_triangular_system_d,
            llh_solvec_inplace_d, llh_invc_inplace_d, llh_dev_potrf_d, llh_dev_solve_triangular_d,
            llh_dev_solvec_d, llh_dev_invc_d);
hip_scalar!(f32, llh_cholesky_inplace_dirty_s, llh_cholesky_inplace_s, llh_triangular_inplace_s,
            llh_is_triangular_s, llh_solve_triangular_inplace_s, llh_solve_triangular_system_s,
            llh_solvec_inplace_s, llh_invc_inplace_s, llh_dev_potrf_s, llh_dev_solve_triangular_s,
            llh_dev_solvec_s, llh_dev_invc_s);

//! Matrices resident in HBM, for callers that keep data on the device between
//! calls: the device-resident factor and solves (llh_dev_*), and the
//! mixed-precision solve with iterative refinement (llh_dev_solve_mixed).
//! Storage is row-major with a leading dimension (element (i, j) at
//! `ptr[i * ld + j]`), the layout every device entry point takes.
use ndarray::{Array2, ArrayBase, Data, Ix2};
use std::marker::PhantomData;
use std::os::raw::c_void;

use crate::context::call;
use crate::ffi;
use crate::scalar::HipScalar;
use crate::triangular::UPLO;
use linfa_linalg::Result;

pub struct DeviceMatrix<A: HipScalar> {
    ptr: *mut A,
    rows: usize,
    cols: usize,
    _t: PhantomData<A>,
}

fn hip_check(e: ffi::hipError_t, what: &str) {
    if e != 0 {
        panic!("linfa-linalg-hip: {} failed (hipError_t {})", what, e);
    }
}

impl<A: HipScalar> DeviceMatrix<A> {
    /// Uninitialised rows x cols storage (panics when device memory is exhausted).
    pub fn uninit(rows: usize, cols: usize) -> Self {
        let mut p: *mut c_void = std::ptr::null_mut();
        let bytes = rows.max(1) * cols.max(1) * std::mem::size_of::<A>();
        hip_check(unsafe { ffi::hipMalloc(&mut p, bytes) }, "hipMalloc");
        DeviceMatrix { ptr: p as *mut A, rows, cols, _t: PhantomData }
    }

    /// A copy of `a` (any layout) in HBM.
    pub fn from_host<S: Data<Elem = A>>(a: &ArrayBase<S, Ix2>) -> Self {
        let (rows, cols) = a.dim();
        let m = Self::uninit(rows, cols);
        let std = a.as_standard_layout();
        if rows > 0 && cols > 0 {
            let pitch = cols * std::mem::size_of::<A>();
            hip_check(unsafe {
                ffi::hipMemcpy2D(m.ptr as *mut c_void, pitch, std.as_ptr() as *const c_void, pitch, pitch, rows,
                                 ffi::HIP_MEMCPY_HOST_TO_DEVICE)
            }, "hipMemcpy2D");
        }
        m
    }

    /// A host copy (C order).
    pub fn to_host(&self) -> Array2<A> {
        let mut out = Array2::<A>::zeros((self.rows, self.cols));
        if self.rows > 0 && self.cols > 0 {
            let pitch = self.cols * std::mem::size_of::<A>();
            hip_check(unsafe {
                ffi::hipMemcpy2D(out.as_mut_ptr() as *mut c_void, pitch, self.ptr as *const c_void, pitch, pitch,
                                 self.rows, ffi::HIP_MEMCPY_DEVICE_TO_HOST)
            }, "hipMemcpy2D");
        }
        out
    }

    pub fn dim(&self) -> (usize, usize) {
        (self.rows, self.cols)
    }

    pub fn as_device_ptr(&self) -> *const A {
        self.ptr
    }

    pub fn as_device_mut_ptr(&mut self) -> *mut A {
        self.ptr
    }

    fn square(&self) -> Result<usize> {
        if self.rows != self.cols {
            return Err(linfa_linalg::LinalgError::NotSquare { rows: self.rows, cols: self.cols });
        }
        Ok(self.rows)
    }

    /// In-place lower Cholesky factor (cholesky_inplace_dirty, or
    /// cholesky_inplace with `zero_upper`), blocked on MFMA for large n.
    pub fn potrf(&mut self, zero_upper: bool) -> Result<()> {
        let n = self.square()?;
        let a = self.ptr;
        call(|c, e| unsafe { A::dev_potrf(c, n, a, n, zero_upper as i32, e) })
    }

    /// op(self) x = b in place in `b` (n x m); `uplo` names op(self)'s triangle
    /// and `trans` reads self transposed (how solvec solves with L.t()).
    pub fn solve_triangular(&self, b: &mut DeviceMatrix<A>, uplo: UPLO, trans: bool) -> Result<()> {
        let n = self.square()?;
        if b.rows != n {
            return Err(linfa_linalg::LinalgError::WrongRows { expected: n, actual: b.rows });
        }
        let (a, m, bp) = (self.ptr as *const A, b.cols, b.ptr);
        let u = match uplo {
            UPLO::Upper => ffi::LLH_UPPER,
            UPLO::Lower => ffi::LLH_LOWER,
        };
        call(|c, e| unsafe { A::dev_solve_triangular(c, u, trans as i32, n, a, n, m, bp, m.max(1), e) })
    }

    /// solvec_inplace on device memory: self := its dirty factor, b := A^-1 b.
    /// After NotPositiveDefinite `b` is left as it was.
    pub fn solvec(&mut self, b: &mut DeviceMatrix<A>) -> Result<()> {
        let n = self.square()?;
        if b.rows != n {
            return Err(linfa_linalg::LinalgError::WrongRows { expected: n, actual: b.rows });
        }
        let (a, m, bp) = (self.ptr, b.cols, b.ptr);
        call(|c, e| unsafe { A::dev_solvec(c, n, a, n, m, bp, m.max(1), e) })
    }
}

impl<A: HipScalar> DeviceMatrix<A> {
    /// invc_inplace on device memory: self := its dirty factor, the result is
    /// A^-1 (both triangles; L^-T L^-1 on MFMA above the exact kernel's bound).
    pub fn invc(&mut self) -> Result<DeviceMatrix<A>> {
        let n = self.square()?;
        let out = DeviceMatrix::<A>::uninit(n, n);
        let (a, o) = (self.ptr, out.ptr);
        call(|c, e| unsafe { A::dev_invc(c, n, a, n, o, n, e) })?;
        Ok(out)
    }
}

/// Factor precision of the mixed-precision solve.
#[derive(Debug, Clone, Copy, PartialEq, Eq)]
pub enum Precision {
    /// plain f64 factor
    F64,
    /// f32 trailing updates on MFMA, f64 diagonal blocks, f64 refinement
    F32,
    /// f16 trailing updates (f32 accumulation) of the diagonally scaled matrix, f64 refinement
    F16,
}

/// What the refinement did.
#[derive(Debug, Clone, Copy)]
pub struct RefineInfo {
    pub iterations: i32,
    pub converged: bool,
    pub fell_back: bool,
    pub residual: f64,
}

impl DeviceMatrix<f64> {
    /// x = A^-1 b with A = self (lower triangle read, left unchanged), factored
    /// in `prec` and refined in f64 to f64 accuracy; with `fallback`, a
    /// low-precision breakdown or stall is redone in f64.  `x` and `b` are n x m.
    pub fn solve_mixed(&self, b: &DeviceMatrix<f64>, x: &mut DeviceMatrix<f64>, prec: Precision, fallback: bool)
                       -> Result<RefineInfo> {
        let n = self.square()?;
        if b.rows != n || x.rows != n || x.cols != b.cols {
            return Err(linfa_linalg::LinalgError::WrongRows { expected: n, actual: b.rows });
        }
        let p = match prec {
            Precision::F64 => ffi::LLH_PREC_F64,
            Precision::F32 => ffi::LLH_PREC_F32,
            Precision::F16 => ffi::LLH_PREC_F16,
        };
        let opts = ffi::llh_refine_opts { max_iter: 0, flags: if fallback { ffi::LLH_FACTOR_FALLBACK_F64 } else { 0 } };
        let mut info = ffi::llh_refine_info::default();
        let ip: *mut ffi::llh_refine_info = &mut info;
        let (a, bp, xp, m) = (self.ptr as *const f64, b.ptr as *const f64, x.ptr, b.cols);
        call(|c, e| unsafe { ffi::llh_dev_solve_mixed(c, p, n, a, n, m, bp, m, xp, m, &opts, ip, e) })?;
        Ok(RefineInfo {
            iterations: info.iterations,
            converged: info.converged != 0,
            fell_back: info.fell_back != 0,
            residual: info.residual,
        })
    }
}

impl<A: HipScalar> Drop for DeviceMatrix<A> {
    fn drop(&mut self) {
        unsafe {
            ffi::hipFree(self.ptr as *mut c_void);
        }
    }
}

//! Triangular matrices and solves: the traits of linfa-linalg 0.2.1's
//! `src/triangular.rs:10-187`, same names, signatures and provided methods,
//! with linfa-linalg's own `UPLO`.  A solve reads only the `uplo` triangle and
//! the diagonal of `self`; a zero pivot yields inf/NaN, not an error.
use ndarray::{Array2, ArrayBase, Data, DataMut, Ix1, Ix2};
use std::os::raw::c_int;

use crate::context::call;
use crate::ffi;
use crate::scalar::HipScalar;
use linfa_linalg::Result;

pub use linfa_linalg::triangular::UPLO;

fn uplo_code(uplo: UPLO) -> ffi::llh_uplo {
    match uplo {
        UPLO::Upper => ffi::LLH_UPPER,
        UPLO::Lower => ffi::LLH_LOWER,
    }
}

/// triangular.rs:16-54
pub trait IntoTriangular {
    /// Zeroes the other triangle (strictly below for Upper, strictly above for Lower).
    fn triangular_inplace(&mut self, uplo: UPLO) -> Result<&mut Self>;

    fn into_triangular(self, uplo: UPLO) -> Result<Self>
    where
        Self: Sized;
}

impl<A: HipScalar, S: DataMut<Elem = A>> IntoTriangular for ArrayBase<S, Ix2> {
    fn triangular_inplace(&mut self, uplo: UPLO) -> Result<&mut Self> {
        let (rows, cols) = self.dim();
        let (rs, cs) = (self.strides()[0], self.strides()[1]);
        let a = self.as_mut_ptr();
        call(|c, e| unsafe { A::triangular_inplace(c, rows, cols, a, rs, cs, uplo_code(uplo), e) })?;
        Ok(self)
    }

    fn into_triangular(mut self, uplo: UPLO) -> Result<Self> {
        self.triangular_inplace(uplo)?;
        Ok(self)
    }
}

/// triangular.rs:57-91
pub trait Triangular {
    /// True when the other triangle is all zero; false for a non-square array.
    fn is_triangular(&self, uplo: UPLO) -> bool;
}

impl<A: HipScalar, S: Data<Elem = A>> Triangular for ArrayBase<S, Ix2> {
    fn is_triangular(&self, uplo: UPLO) -> bool {
        let (rows, cols) = self.dim();
        let (rs, cs) = (self.strides()[0], self.strides()[1]);
        let a = self.as_ptr();
        let mut out: c_int = 0;
        let o: *mut c_int = &mut out;
        match call(|c, e| unsafe { A::is_triangular(c, rows, cols, a, rs, cs, uplo_code(uplo), o, e) }) {
            Ok(()) => out != 0,
            Err(_) => false,
        }
    }
}

/// triangular.rs:147-169: solves `self` x = b in place in `b`.
pub trait SolveTriangularInplace<B> {
    fn solve_triangular_inplace<'a>(&self, b: &'a mut B, uplo: UPLO) -> Result<&'a mut B>;

    fn solve_triangular_into(&self, mut b: B, uplo: UPLO) -> Result<B> {
        self.solve_triangular_inplace(&mut b, uplo)?;
        Ok(b)
    }
}

impl<A: HipScalar, Si: Data<Elem = A>, So: DataMut<Elem = A>> SolveTriangularInplace<ArrayBase<So, Ix2>>
    for ArrayBase<Si, Ix2>
{
    fn solve_triangular_inplace<'a>(&self, b: &'a mut ArrayBase<So, Ix2>, uplo: UPLO)
                                    -> Result<&'a mut ArrayBase<So, Ix2>> {
        let (rows, cols) = self.dim();
        let (rs, cs) = (self.strides()[0], self.strides()[1]);
        let (brows, bcols) = b.dim();
        let (brs, bcs) = (b.strides()[0], b.strides()[1]);
        let a = self.as_ptr();
        let bp = b.as_mut_ptr();
        call(|c, e| unsafe {
            A::solve_triangular_inplace(c, rows, cols, a, rs, cs, brows, bcols, bp, brs, bcs, uplo_code(uplo), e)
        })?;
        Ok(b)
    }
}

/// triangular.rs:172-187: on a copy of `b`.
pub trait SolveTriangular<B> {
    type Output;

    fn solve_triangular(&self, b: &B, uplo: UPLO) -> Result<Self::Output>;
}

impl<A: HipScalar, Si: Data<Elem = A>, So: Data<Elem = A>> SolveTriangular<ArrayBase<So, Ix2>>
    for ArrayBase<Si, Ix2>
{
    type Output = Array2<A>;

    fn solve_triangular(&self, b: &ArrayBase<So, Ix2>, uplo: UPLO) -> Result<Self::Output> {
        self.solve_triangular_into(b.to_owned(), uplo)
    }
}

/// The crate-private `solve_triangular_system(a, b, uplo, diag_fn)` of
/// triangular.rs:133-144 with the pivots from a vector: `diag_fn(i)` is
/// `|diag[i]|` when `abs_diag`, else `diag[i]`, and `a`'s own diagonal is never
/// read -- what QR's back-substitutions pass (qr.rs:145-150, 172-177).
pub fn solve_triangular_system<A, Sa, Sb, Sd>(a: &ArrayBase<Sa, Ix2>, b: &mut ArrayBase<Sb, Ix2>, uplo: UPLO,
                                              diag: &ArrayBase<Sd, Ix1>, abs_diag: bool) -> Result<()>
where
    A: HipScalar,
    Sa: Data<Elem = A>,
    Sb: DataMut<Elem = A>,
    Sd: Data<Elem = A>,
{
    let (rows, cols) = a.dim();
    assert!(diag.len() >= rows.min(cols), "one pivot per row of a");
    let (rs, cs) = (a.strides()[0], a.strides()[1]);
    let (brows, bcols) = b.dim();
    let (brs, bcs) = (b.strides()[0], b.strides()[1]);
    let ds = diag.strides()[0];
    let (ap, dp) = (a.as_ptr(), diag.as_ptr());
    let bp = b.as_mut_ptr();
    call(|c, e| unsafe {
        A::solve_triangular_system(c, rows, cols, ap, rs, cs, dp, ds, abs_diag as c_int, brows, bcols, bp, brs, bcs,
                                   uplo_code(uplo), e)
    })
}

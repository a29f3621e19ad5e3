//! Cholesky factorisation and SPD solves: the traits of linfa-linalg 0.2.1's
//! `src/cholesky.rs:16-199`, same names, signatures and provided methods.
//!
//! Semantics kept: only the lower triangle is read; the dirty forms leave the
//! strict upper triangle as it was; `solvec*` / `invc_inplace` leave `self`
//! holding the dirty factor, also when the later `WrongRows` check fails; errors
//! come in the reference's order (NotSquare, NotPositiveDefinite, WrongRows);
//! after NotPositiveDefinite at row j the array holds what the reference leaves.
use ndarray::{Array2, ArrayBase, Data, DataMut, Ix2};

use crate::context::call;
use crate::scalar::HipScalar;
use linfa_linalg::Result;

/// cholesky.rs:16-44
pub trait CholeskyInplace {
    /// The lower factor L with A = L L^T, in place; the strict upper triangle is left as it was.
    fn cholesky_inplace_dirty(&mut self) -> Result<&mut Self>;

    fn cholesky_into_dirty(mut self) -> Result<Self>
    where
        Self: Sized,
    {
        self.cholesky_inplace_dirty()?;
        Ok(self)
    }

    /// As `cholesky_inplace_dirty`, then the strict upper triangle is zeroed.
    fn cholesky_inplace(&mut self) -> Result<&mut Self>;

    fn cholesky_into(mut self) -> Result<Self>
    where
        Self: Sized,
    {
        self.cholesky_inplace()?;
        Ok(self)
    }
}

impl<A: HipScalar, S: DataMut<Elem = A>> CholeskyInplace for ArrayBase<S, Ix2> {
    fn cholesky_inplace_dirty(&mut self) -> Result<&mut Self> {
        let (rows, cols) = self.dim();
        let (rs, cs) = (self.strides()[0], self.strides()[1]);
        let a = self.as_mut_ptr();
        call(|c, e| unsafe { A::cholesky_inplace_dirty(c, rows, cols, a, rs, cs, e) })?;
        Ok(self)
    }

    fn cholesky_inplace(&mut self) -> Result<&mut Self> {
        let (rows, cols) = self.dim();
        let (rs, cs) = (self.strides()[0], self.strides()[1]);
        let a = self.as_mut_ptr();
        call(|c, e| unsafe { A::cholesky_inplace(c, rows, cols, a, rs, cs, e) })?;
        Ok(self)
    }
}

/// cholesky.rs:86-115: on a copy of `self`.
pub trait Cholesky {
    type Output;

    fn cholesky_dirty(&self) -> Result<Self::Output>;

    fn cholesky(&self) -> Result<Self::Output>;
}

impl<A: HipScalar, S: Data<Elem = A>> Cholesky for ArrayBase<S, Ix2> {
    type Output = Array2<A>;

    fn cholesky_dirty(&self) -> Result<Self::Output> {
        self.to_owned().cholesky_into_dirty()
    }

    fn cholesky(&self) -> Result<Self::Output> {
        self.to_owned().cholesky_into()
    }
}

/// cholesky.rs:118-145: `self` := its dirty factor, `b` := A^-1 b.
pub trait SolveCInplace<B> {
    fn solvec_inplace<'a>(&mut self, b: &'a mut B) -> Result<&'a mut B>;

    fn solvec_into(&mut self, mut b: B) -> Result<B> {
        self.solvec_inplace(&mut b)?;
        Ok(b)
    }
}

impl<A: HipScalar, Si: DataMut<Elem = A>, So: DataMut<Elem = A>> SolveCInplace<ArrayBase<So, Ix2>>
    for ArrayBase<Si, Ix2>
{
    fn solvec_inplace<'a>(&mut self, b: &'a mut ArrayBase<So, Ix2>) -> Result<&'a mut ArrayBase<So, Ix2>> {
        let (rows, cols) = self.dim();
        let (rs, cs) = (self.strides()[0], self.strides()[1]);
        let (brows, bcols) = b.dim();
        let (brs, bcs) = (b.strides()[0], b.strides()[1]);
        let a = self.as_mut_ptr();
        let bp = b.as_mut_ptr();
        call(|c, e| unsafe { A::solvec_inplace(c, rows, cols, a, rs, cs, brows, bcols, bp, brs, bcs, e) })?;
        Ok(b)
    }
}

/// cholesky.rs:148-163: on a copy of `b`.
pub trait SolveC<B> {
    type Output;

    fn solvec(&mut self, b: &B) -> Result<Self::Output>;
}

impl<A: HipScalar, Si: DataMut<Elem = A>, So: Data<Elem = A>> SolveC<ArrayBase<So, Ix2>> for ArrayBase<Si, Ix2> {
    type Output = Array2<A>;

    fn solvec(&mut self, b: &ArrayBase<So, Ix2>) -> Result<Self::Output> {
        self.solvec_into(b.to_owned())
    }
}

/// cholesky.rs:166-183: A^-1 (a new C-order array); `self` := its dirty factor.
pub trait InverseCInplace {
    type Output;

    fn invc_inplace(&mut self) -> Result<Self::Output>;
}

impl<A: HipScalar, S: DataMut<Elem = A>> InverseCInplace for ArrayBase<S, Ix2> {
    type Output = Array2<A>;

    fn invc_inplace(&mut self) -> Result<Self::Output> {
        let (rows, cols) = self.dim();
        let (rs, cs) = (self.strides()[0], self.strides()[1]);
        let mut out = Array2::<A>::zeros((rows, rows));
        let (ors, ocs) = (out.strides()[0], out.strides()[1]);
        let a = self.as_mut_ptr();
        let o = out.as_mut_ptr();
        call(|c, e| unsafe { A::invc_inplace(c, rows, cols, a, rs, cs, o, ors, ocs, e) })?;
        Ok(out)
    }
}

/// cholesky.rs:186-199: on a copy of `self`.
pub trait InverseC {
    type Output;

    fn invc(&self) -> Result<Self::Output>;
}

impl<A: HipScalar, S: Data<Elem = A>> InverseC for ArrayBase<S, Ix2> {
    type Output = Array2<A>;

    fn invc(&self) -> Result<Self::Output> {
        self.to_owned().invc_inplace()
    }
}

//! linfa-linalg's Cholesky / SPD-solve / triangular-solve traits, computed by
//! the hand-written HIP kernels of libllh.so on an MI355X (gfx950).
//!
//! The traits keep linfa-linalg 0.2.1's names, methods, signatures, provided
//! methods and semantics (`src/cholesky.rs:16-199`, `src/triangular.rs:10-187`):
//! switching a caller over is a change of import path,
//! `linfa_linalg::cholesky::*` -> `linfa_linalg_hip::cholesky::*` and
//! `linfa_linalg::triangular::*` -> `linfa_linalg_hip::triangular::*`.
//! Errors are linfa-linalg's own `LinalgError` and `Result`, and `UPLO` is
//! linfa-linalg's own enum, so values pass between the two crates unchanged.
//! (The traits are declared here rather than implemented for linfa-linalg's:
//! a third crate may not implement a foreign trait for ndarray's foreign
//! `ArrayBase`, and linfa-linalg's blanket impls would overlap anyway.)
//!
//! Every method hands the array's pointer, dimensions and signed element
//! strides (`as_ptr`, `dim`, `strides`) to the C ABI, so every ndarray layout,
//! inverted and transposed axes included, goes through unchanged.  A failure of
//! the device itself (a HIP error, device memory exhausted) has no
//! `LinalgError` variant and panics with the library's message, as an
//! allocation failure does in ndarray.
//!
//! `device::DeviceMatrix` keeps matrices in HBM between calls: the
//! device-resident factor and solves, and the mixed-precision solve with
//! iterative refinement (f64 accuracy from an FP32-MFMA factor).
//!
//! Element types: f32 and f64 (linfa-linalg's `NdFloat`).  The reference's
//! `IntoTriangular` / `Triangular` also accept integer arrays; this crate's
//! take f32 and f64 only.
pub mod cholesky;
mod context;
pub mod device;
pub mod ffi;
mod scalar;
pub mod triangular;

pub use linfa_linalg::{LinalgError, Result};
pub use scalar::HipScalar;
pub use triangular::UPLO;

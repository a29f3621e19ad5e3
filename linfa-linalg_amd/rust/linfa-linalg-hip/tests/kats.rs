//! The reference's own known answers (linfa-linalg 0.2.1 src/cholesky.rs:205-277,
//! src/triangular.rs:195-299), through the HIP library.  Needs an MI355X.
use approx::assert_abs_diff_eq;
use linfa_linalg::LinalgError;
use linfa_linalg_hip::cholesky::*;
use linfa_linalg_hip::triangular::*;
use ndarray::{array, Array2};

#[test]
fn decompose() {
    let arr = array![[25f64, 15., -5.], [15., 18., 0.], [-5., 0., 11.]];
    let lower = array![[5.0, 0.0, 0.0], [3.0, 3.0, 0.0], [-1., 1., 3.]];
    let chol = arr.cholesky().unwrap();
    assert_abs_diff_eq!(chol, lower, epsilon = 1e-7);
    assert!(chol.is_triangular(UPLO::Lower));
}

#[test]
fn bad_matrix() {
    let row = array![[1., 2., 3.], [3., 4., 5.]];
    assert!(matches!(row.cholesky(), Err(LinalgError::NotSquare { rows: 2, cols: 3 })));
    let non_pd = array![[1., 2.], [2., 1.]];
    assert!(matches!(non_pd.cholesky(), Err(LinalgError::NotPositiveDefinite)));
    let zeros = array![[0., 0.], [0., 0.]];
    assert!(matches!(zeros.cholesky(), Err(LinalgError::NotPositiveDefinite)));
}

#[test]
fn solvec() {
    let mut arr = array![[25f64, 15., -5.], [15., 18., 0.], [-5., 0., 11.]];
    let x = array![[10., -3., 2.2, 4.], [0., 2.4, -0.9, 1.1], [5.5, 7.6, 8.1, 10.]];
    let b = arr.dot(&x);
    let out = arr.solvec(&b).unwrap();
    assert_abs_diff_eq!(out, x, epsilon = 1e-7);
}

#[test]
fn invc() {
    let arr = array![[25f64, 15., -5.], [15., 18., 0.], [-5., 0., 11.]];
    let inv = arr.invc().unwrap();
    assert_abs_diff_eq!(arr.dot(&inv), Array2::eye(3));
}

#[test]
fn corner_cases() {
    let empty = Array2::<f64>::zeros((0, 0));
    assert_eq!(empty.cholesky().unwrap(), empty);
    assert_eq!(empty.clone().invc().unwrap(), empty);
    let one = array![[1.]];
    assert_eq!(one.cholesky().unwrap(), one);
    assert_eq!(one.clone().invc().unwrap(), one);
}

#[test]
fn solve_triangular_kats() {
    let lower = array![[1.0, 0.0], [3.0, 4.0]];
    assert!(lower.is_triangular(UPLO::Lower));
    let expected = array![[2.2, 3.1, 2.2], [1.0, 0.0, 5.7]];
    let x = lower.solve_triangular_into(lower.dot(&expected), UPLO::Lower).unwrap();
    assert_abs_diff_eq!(x, expected, epsilon = 1e-7);
    let upper = array![[4.4, 2.1], [0.0, 4.3]];
    assert!(upper.is_triangular(UPLO::Upper));
    let x = upper.solve_triangular_into(upper.dot(&expected), UPLO::Upper).unwrap();
    assert_abs_diff_eq!(x, expected, epsilon = 1e-7);
}

#[test]
fn solve_errors_and_corners() {
    let non_square = array![[1.2f64, 3.3]];
    assert!(matches!(non_square.solve_triangular(&non_square, UPLO::Lower), Err(LinalgError::NotSquare { .. })));
    let square = array![[1.1, 2.2], [3.3, 2.1]];
    assert!(matches!(square.solve_triangular(&array![[2.2, 3.3]], UPLO::Upper),
                     Err(LinalgError::WrongRows { expected: 2, actual: 1 })));
    let empty = Array2::<f64>::zeros((0, 0));
    assert_eq!(empty.solve_triangular(&empty, UPLO::Upper).unwrap().dim(), (0, 0));
    let one = Array2::<f64>::ones((1, 1));
    assert_abs_diff_eq!(one.solve_triangular(&one, UPLO::Upper).unwrap(), one);
    let diag_zero = array![[0., 3.], [2., 0.]];
    diag_zero.solve_triangular(&Array2::<f64>::zeros((2, 2)), UPLO::Lower).unwrap();
}

#[test]
fn transposed_and_inverted_layouts() {
    use ndarray::Axis;
    let a = array![[4.0f64, 2.0, 0.6], [2.0, 5.0, 1.0], [0.6, 1.0, 3.0]];
    // a transposed view of a symmetric matrix factors to the same L
    assert_abs_diff_eq!(a.t().cholesky().unwrap(), a.cholesky().unwrap(), epsilon = 1e-12);
    // both axes inverted (negative strides): P A P y = P b gives y = P x
    let rhs = array![[1.0], [2.0], [3.0]];
    let x = a.clone().solvec(&rhs).unwrap();
    let mut v = a.clone();
    v.invert_axis(Axis(0));
    v.invert_axis(Axis(1));
    let mut r = rhs.clone();
    r.invert_axis(Axis(0));
    let mut y = v.solvec(&r).unwrap();
    y.invert_axis(Axis(0));
    assert_abs_diff_eq!(y, x, epsilon = 1e-12);
}

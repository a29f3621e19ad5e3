"""Pins the CPU oracle (the checker for the HIP path) to the reference.

- every known-answer vector held in linfa-linalg's own tests, plus exact
  integer constructions (tests/pins.py);
- the reference's property tests (tests/cholesky.rs, tests/triangular.rs) with
  its value range, dimension range and random memory layouts
  (tests/common.rs:9-43), via hypothesis;
- an independent LAPACK (scipy) cross-check at larger n;
- planted mistakes: every mutant copy of the oracle must fail a pin.
"""
import numpy as np
import pytest
import scipy.linalg as sla
from hypothesis import given, settings, strategies as st

from oracle import oracle as orc
from pins import pin_failures

UPPER, LOWER = orc.UPPER, orc.LOWER


def test_oracle_pinned():
    assert pin_failures(orc.Oracle()) == []


@pytest.mark.parametrize("mutant", orc.MUTANTS)
def test_planted_mistake_is_caught(mutant):
    assert pin_failures(orc.Oracle(mutant)) != []


# --- the reference's proptest strategies (tests/common.rs) -------------------

def layouts(a, invert_rows, invert_cols, transpose):
    """common.rs:19-33: invert axes (negative strides) and/or transpose."""
    if invert_rows:
        a = a[::-1, :]
    if invert_cols:
        a = a[:, ::-1]
    if transpose and a.shape[0] == a.shape[1]:
        a = a.T
    return a


@st.composite
def matrix(draw, rows, cols):
    data = draw(st.lists(st.floats(-100.0, 100.0), min_size=rows * cols, max_size=rows * cols))
    lay = draw(st.tuples(st.booleans(), st.booleans(), st.booleans()))
    return layouts(np.array(data, dtype=np.float64).reshape(rows, cols), *lay)


@st.composite
def hpd_arr(draw):
    """tests/cholesky.rs:9-19: arr^T arr + I, a fresh C-order array."""
    n = draw(st.integers(1, 10))
    a = draw(matrix(n, n))
    m = a.T @ a
    m[np.diag_indices(n)] += 1.0
    return np.ascontiguousarray(m)


@st.composite
def system_of(draw, arrs):
    a = draw(arrs)
    cols = draw(st.integers(1, 10))
    return a, draw(matrix(a.shape[1], cols))


@st.composite
def tri_arr(draw, uplo):
    """tests/triangular.rs:9-20: triangular part, |diag| < 1 replaced by 1."""
    n = draw(st.integers(1, 10))
    a = draw(matrix(n, n))  # a strided view; triangularised in place like the reference
    other = np.tril(np.ones((n, n)), -1) if uplo == UPPER else np.triu(np.ones((n, n)), 1)
    a[other > 0] = 0.0
    d = np.diagonal(a).copy()
    d[np.abs(d) < 1.0] = 1.0
    a[np.diag_indices(n)] = d
    return a


@settings(max_examples=300, deadline=None, derandomize=True)
@given(hpd_arr())
def test_cholesky_property(orig):
    """tests/cholesky.rs:21-54 (tolerance 1e-7)."""
    chol = orc.cholesky(orig)
    assert np.abs(chol @ chol.T - orig).max() <= 1e-7
    assert orc.is_triangular(chol, LOWER)
    dirty = orc.cholesky_dirty(orig)
    assert np.abs(chol - orc.triangular_inplace(dirty, LOWER)).max() <= 1e-7


@settings(max_examples=300, deadline=None, derandomize=True)
@given(system_of(hpd_arr()))
def test_solvec_property(system):
    """tests/cholesky.rs:56-62 (tolerance 1e-5)."""
    a, x = system
    b = a @ x
    assert np.abs(orc.solvec(np.array(a), b) - x).max() <= 1e-5


@settings(max_examples=300, deadline=None, derandomize=True)
@given(hpd_arr())
def test_invc_property(a):
    """tests/cholesky.rs:64-67 (tolerance 1e-7)."""
    assert np.abs(a @ orc.invc(a) - np.eye(a.shape[0])).max() <= 1e-7


@settings(max_examples=300, deadline=None, derandomize=True)
@given(st.sampled_from([UPPER, LOWER]).flatmap(lambda u: st.tuples(st.just(u),
                                                                    system_of(tri_arr(u)))))
def test_solve_triangular_property(case):
    """tests/triangular.rs:22-47 (tolerance 1e-4), with strided layouts."""
    uplo, (a, x) = case
    b = a @ x
    assert np.abs(orc.solve_triangular(a, b, uplo) - x).max() <= 1e-4
    # the in-place form writes into a strided (negative-stride) b as well
    bb = np.array(b)[::-1, :][::-1, :]
    orc.solve_triangular_inplace(a, bb, uplo)
    assert np.abs(bb - x).max() <= 1e-4


# --- independent cross-check against LAPACK ------------------------------------

@pytest.mark.parametrize("n", [16, 64, 200])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_matches_lapack(n, dtype):
    rng = np.random.default_rng(n)
    b = rng.uniform(-1, 1, (n, n))
    a = (b @ b.T / n + np.eye(n)).astype(dtype)
    l_ref = sla.cholesky(a.astype(np.float64), lower=True)
    l = orc.cholesky(a)
    u = np.finfo(dtype).eps
    assert np.abs(l - l_ref).max() <= 50 * n * u * np.abs(l_ref).max()
    rhs = rng.standard_normal((n, 5)).astype(dtype)
    y = orc.solve_triangular(l, rhs, LOWER)
    y_ref = sla.solve_triangular(l.astype(np.float64), rhs.astype(np.float64), lower=True)
    assert np.abs(y - y_ref).max() <= 50 * n * u * np.abs(y_ref).max()
    x = orc.solvec(np.array(a), rhs)
    x_ref = sla.cho_solve((l_ref, True), rhs.astype(np.float64))
    assert np.abs(x - x_ref).max() <= 200 * n * u * np.abs(x_ref).max()


def test_not_pd_reports_first_failing_row_and_partial_state():
    """cholesky.rs:54-72: rows < j are factored, row j holds L[j, :j] and the
    original diagonal, rows > j are untouched."""
    rng = np.random.default_rng(3)
    n, j = 12, 7
    l0, a = integer_factor_like(n, rng)
    a[j, j] = -5.0
    work = np.array(a)
    with pytest.raises(orc.OracleError) as ei:
        orc.cholesky_inplace_dirty(work)
    assert ei.value.code == orc.NOT_POSITIVE_DEFINITE and ei.value.column == j
    assert np.array_equal(np.tril(work[:j, :j]), l0[:j, :j])
    assert np.array_equal(work[j, :j], l0[j, :j]) and work[j, j] == -5.0
    assert np.array_equal(work[j + 1:], a[j + 1:])


def integer_factor_like(n, rng):
    from pins import integer_factor
    return integer_factor(n, rng)


def test_nan_pivot_passes():
    """`d <= 0` is false for NaN (cholesky.rs:69): Ok with NaNs."""
    a = np.array([[4.0, 2.0], [2.0, np.nan]])
    l = orc.cholesky(a)
    assert l[0, 0] == 2.0 and l[1, 0] == 1.0 and np.isnan(l[1, 1])

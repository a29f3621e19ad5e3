"""Parity of the HIP path with the oracle (and, through it, the reference).

- every known-answer and exact-integer pin (tests/pins.py) through the host
  mirror, which calls the C ABI;
- bit-identical factors and solutions for matrices the exact-order LDS kernel
  takes (n <= 128 f64, n <= 176 f32), against the oracle and against the
  committed seeded fixtures;
- the blocked path at larger n: backward error ||A - L L^T||_F / ||A||_F
  <= 1e-13 (f64) and solution error against the oracle's solution;
- the reference's property tests with random strided layouts;
- error paths, including the partially written array after NotPositiveDefinite.
"""
import os

import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

import linfa_linalg_hip as llh
from oracle import oracle as orc
from pins import pin_failures, integer_factor

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
U, L = llh.UPLO.Upper, llh.UPLO.Lower


def spd(n, seed, dtype=np.float64, shift=1.0):
    """A = B B^T / n + shift I with B uniform in [-1, 1] (the C1 construction)."""
    rng = np.random.default_rng(seed)
    b = rng.uniform(-1.0, 1.0, size=(n, n))
    return (b @ b.T / n + shift * np.eye(n)).astype(dtype)


def backward_error(a, l):
    l = np.tril(l).astype(np.float64)
    a = a.astype(np.float64)
    al = np.tril(a)
    r = np.tril(l @ l.T) - al
    return np.linalg.norm(r) / np.linalg.norm(al)


def test_pins_through_hip():
    assert pin_failures(llh) == []


@pytest.mark.parametrize("dtype,nmax", [(np.float64, 128), (np.float32, 176)])
@pytest.mark.parametrize("n", [1, 2, 3, 7, 16, 33, 64, 65, 100, 128, 176])
def test_small_bit_identical(n, dtype, nmax):
    if n > nmax:
        pytest.skip("above the exact kernel's size")
    a = spd(n, 10 + n, dtype)
    assert np.array_equal(llh.cholesky_dirty(a), orc.cholesky_dirty(a))
    assert np.array_equal(llh.cholesky(a), orc.cholesky(a))
    b = np.random.default_rng(n).standard_normal((n, 5)).astype(dtype)
    a1, a2 = np.array(a), np.array(a)
    x = llh.solvec(a1, b)
    x_ref = orc.solvec(a2, b)
    assert np.array_equal(x, x_ref)
    assert np.array_equal(a1, a2)  # both left holding the dirty factor


def test_seeded_fixtures_bit_identical():
    fx = np.load(os.path.join(HERE, "golden", "seeded_oracle.npz"))
    for tag in ("d", "s"):
        for n in (1, 2, 3, 5, 8, 10, 16, 64, 128):
            a, b = fx[f"{tag}{n}_a"], fx[f"{tag}{n}_b"]
            fac = np.array(a)
            x = llh.solvec(fac, b)
            assert np.array_equal(fac, fx[f"{tag}{n}_dirty"]), (tag, n)
            assert np.array_equal(x, fx[f"{tag}{n}_x"]), (tag, n)


@pytest.mark.parametrize("n", [129, 200, 256, 300, 512, 777, 1024, 1111, 1537])
def test_blocked_f64_against_oracle(n):
    a = spd(n, n)
    l_ref = orc.cholesky(a)
    l = llh.cholesky(a)
    assert backward_error(a, l) <= 1e-13
    # both are backward stable factors of the same A: they agree to ~cond * eps
    assert np.abs(l - l_ref).max() <= 1e-12 * np.abs(l_ref).max()
    assert np.all(np.triu(l, 1) == 0)
    rng = np.random.default_rng(n)
    b = rng.standard_normal((n, 3))
    x = llh.solvec(np.array(a), b)
    x_ref = orc.solvec(np.array(a), b)
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) <= 1e-12


def test_c1_config_n512():
    """n=512 f64, A = B B^T/n + I (B uniform [-1,1]), cholesky_into then solvec
    with one N(0,1) right-hand side; oracle = the reference algorithm in FP64."""
    a = spd(512, 2024)
    l = llh.cholesky_into(np.array(a))
    assert backward_error(a, l) <= 1e-13
    b = np.random.default_rng(7).standard_normal((512, 1))
    x = llh.solvec(np.array(a), b)
    x_ref = orc.solvec(np.array(a), b)
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) <= 1e-13


@pytest.mark.parametrize("n", [200, 700, 1111])
def test_blocked_f32(n):
    a = spd(n, n, np.float32)
    l = llh.cholesky(a)
    assert backward_error(a, l) <= 1e-5
    l_ref = orc.cholesky(a)
    assert np.abs(l - l_ref).max() <= 1e-4 * np.abs(l_ref).max()


@pytest.mark.parametrize("n", [5, 64, 65, 130, 333])
@pytest.mark.parametrize("uplo", [U, L])
@pytest.mark.parametrize("m", [1, 7, 64])
def test_solve_triangular_against_oracle(n, uplo, m):
    rng = np.random.default_rng(n * 10 + m)
    t = rng.uniform(-1, 1, (n, n)) / np.sqrt(n)
    t[np.diag_indices(n)] = rng.uniform(1, 2, n)
    t = np.triu(t) if uplo == U else np.tril(t)
    # NaN in the other triangle must never be read
    other = np.triu(np.ones((n, n)), 1) if uplo == L else np.tril(np.ones((n, n)), -1)
    tg = t + np.where(other > 0, np.nan, 0.0)
    b = rng.standard_normal((n, m))
    x = llh.solve_triangular(tg, b, uplo)
    x_ref = orc.solve_triangular(t, b, uplo)
    assert np.abs(x - x_ref).max() <= 1e-12 * max(1.0, np.abs(x_ref).max())


@pytest.mark.parametrize("n", [3, 50, 150, 300])
def test_invc(n):
    a = spd(n, 3 * n)
    inv = llh.invc(a)
    assert np.abs(a @ inv - np.eye(n)).max() <= 1e-12
    inv_ref = orc.invc(a)
    assert np.abs(inv - inv_ref).max() <= 1e-12 * np.abs(inv_ref).max()


@pytest.mark.parametrize("n", [129, 177, 640, 1537])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_invc_mfma_inverse(n, dtype):
    """Above the exact kernel's bound, invc is L^-T L^-1 on MFMA (W = L^-1 by
    halves, then W^T W): against the oracle's solve with the identity
    (cholesky.rs:178-182), to the inverse's condition-scaled accuracy; A keeps
    the dirty factor, as in the reference."""
    a = spd(n, 7 * n).astype(dtype)
    fac = np.array(a)
    inv = llh.invc_inplace(fac)
    fac_ref = np.array(a)
    inv_ref = orc.invc_inplace(fac_ref)
    tol = 1e-12 if dtype == np.float64 else 2e-4
    assert np.abs(inv - inv_ref).max() <= tol * np.abs(inv_ref).max(), n
    if n > (128 if dtype == np.float64 else 176):  # the MFMA inverse mirrors its lower triangle
        assert np.array_equal(inv, inv.T)
    assert np.abs(np.tril(fac) - np.tril(fac_ref)).max() <= tol * np.abs(fac_ref).max()
    assert np.array_equal(np.triu(fac, 1), np.triu(np.array(a), 1))  # dirty: upper untouched
    if dtype == np.float64:
        assert np.abs(a @ inv - np.eye(n)).max() <= 1e-12


def test_invc_strided_and_not_pd():
    n = 300
    a = spd(n, 11)
    big = np.zeros((2 * n, 2 * n))
    view = big[::2, ::-2]  # strided, inverted columns
    view[...] = a
    out = llh.invc(view)
    assert np.abs(out - orc.invc(a)).max() <= 1e-12 * np.abs(out).max()
    bad = np.array(a)
    bad[200, 200] = -1.0
    with pytest.raises(llh.NotPositiveDefinite):
        llh.invc(bad)


@pytest.mark.parametrize("n", [100, 2000])
def test_dev_invc(n):
    import torch
    from linfa_linalg_hip import device as dev
    a = spd(n, 13 + n)
    ta = torch.from_numpy(a).cuda()
    out = torch.full((n, n), float("nan"), dtype=torch.float64, device="cuda")
    dev.invc(ta, out)
    inv_ref = orc.invc(np.array(a))
    inv = out.cpu().numpy()
    if n <= 128:  # the exact-order kernel: bit for bit
        assert np.array_equal(inv, inv_ref)
    assert np.abs(inv - inv_ref).max() <= 1e-12 * np.abs(inv_ref).max()
    bad = np.array(a)
    bad[n // 2, n // 2] = -1.0
    out2 = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    with pytest.raises(llh.NotPositiveDefinite):
        dev.invc(torch.from_numpy(bad).cuda(), out2)
    if n > 128:  # the MFMA inverse leaves out untouched after NotPositiveDefinite
        assert float(out2.abs().max()) == 0.0


# --- error paths ------------------------------------------------------------------

# Blocked orders put the failing column at each position of the 64-column
# diagonal block's four 16-column steps: 300/5 and 300/200 the first step,
# 300/213 and 700/600 the second, 300/172 the third, 700/511 the last column.
@pytest.mark.parametrize("n,j", [(12, 7), (100, 64), (300, 5), (300, 200), (300, 213), (300, 172), (700, 600),
                                 (700, 511)])
def test_not_pd_partial_state_matches_reference(n, j):
    rng = np.random.default_rng(n + j)
    a = spd(n, n + j)
    a[j, j] = -1.0
    work, ref = np.array(a), np.array(a)
    with pytest.raises(llh.NotPositiveDefinite) as ei:
        llh.cholesky_inplace_dirty(work)
    assert ei.value.column == j
    with pytest.raises(orc.OracleError):
        orc.cholesky_inplace_dirty(ref)
    tol = 0 if n <= 128 else 1e-12
    assert np.abs(work[:j + 1] - ref[:j + 1]).max() <= tol
    assert work[j, j] == a[j, j]
    assert np.array_equal(work[j + 1:], a[j + 1:])
    # solvec leaves the same dirty state and reports the same error
    w2 = np.array(a)
    with pytest.raises(llh.NotPositiveDefinite):
        llh.solvec(w2, rng.standard_normal((n, 2)))
    assert np.array_equal(w2, work)
    del rng


@pytest.mark.parametrize("n,j", [(300, 5), (300, 213), (400, 300), (700, 650)])
def test_not_pd_partial_state_f32_blocked(n, j):
    """The f32 blocked path (above the exact kernel's 176) fails at the same
    column as the reference and leaves its partial state; the rows before the
    failure agree with the oracle to f32 rounding."""
    a = spd(n, n + j, dtype=np.float32)
    a[j, j] = -1.0
    work, ref = np.array(a), np.array(a)
    with pytest.raises(llh.NotPositiveDefinite) as ei:
        llh.cholesky_inplace_dirty(work)
    assert ei.value.column == j
    with pytest.raises(orc.OracleError):
        orc.cholesky_inplace_dirty(ref)
    assert np.abs(work[:j + 1] - ref[:j + 1]).max() <= 1e-4
    assert work[j, j] == a[j, j]
    assert np.array_equal(work[j + 1:], a[j + 1:])


def test_error_precedence():
    with pytest.raises(llh.NotSquare) as ei:
        llh.solvec(np.zeros((2, 3)), np.zeros((2, 1)))
    assert (ei.value.rows, ei.value.cols) == (2, 3)
    a = spd(4, 1)
    work = np.array(a)
    with pytest.raises(llh.WrongRows) as ei:
        llh.solvec_inplace(work, np.zeros((3, 1)))
    assert (ei.value.expected, ei.value.actual) == (4, 3)
    # the factor ran before WrongRows (cholesky.rs:140-141)
    assert np.array_equal(work, orc.cholesky_dirty(a))
    with pytest.raises(llh.NotPositiveDefinite):
        llh.solvec(np.array([[1.0, 2.0], [2.0, 1.0]]), np.zeros((5, 1)))


def test_nan_pivot_passes():
    a = np.array([[4.0, 2.0], [2.0, np.nan]])
    l = llh.cholesky(a)
    assert l[0, 0] == 2.0 and l[1, 0] == 1.0 and np.isnan(l[1, 1])


@pytest.mark.parametrize("n,j", [(300, 100), (700, 650)])
def test_nan_pivot_passes_blocked(n, j):
    # cholesky.rs:69 fails only on d <= 0, so a NaN pivot inside a blocked
    # diagonal block is not an error: the NaN reaches exactly the entries that
    # the reference's order reaches (rows and columns >= j of the lower part)
    a = spd(n, n + j)
    a[j, j] = np.nan
    l = np.tril(llh.cholesky_dirty(np.array(a)))
    ref = np.tril(orc.cholesky_dirty(np.array(a)))
    assert np.array_equal(np.isnan(l), np.isnan(ref))
    assert np.isnan(ref[j:, j:][np.tril_indices(n - j)]).all()
    ok = ~np.isnan(ref)
    assert np.abs(l[ok] - ref[ok]).max() <= 1e-12 * np.abs(ref[ok]).max()


def test_zero_columns_and_empty():
    a = spd(5, 1)
    out = llh.solvec(np.array(a), np.zeros((5, 0)))
    assert out.shape == (5, 0)
    assert llh.cholesky(np.zeros((0, 0))).shape == (0, 0)
    with pytest.raises(llh.WrongRows):
        llh.solve_triangular(np.zeros((0, 0)), np.zeros((2, 2)), U)


# --- the reference's property tests with strided layouts ------------------------------

def layouts(a, flags):
    ir, ic, tr = flags
    if ir:
        a = a[::-1, :]
    if ic:
        a = a[:, ::-1]
    if tr and a.shape[0] == a.shape[1]:
        a = a.T
    return a


@st.composite
def matrix(draw, rows, cols):
    data = draw(st.lists(st.floats(-100.0, 100.0), min_size=rows * cols, max_size=rows * cols))
    return layouts(np.array(data).reshape(rows, cols), draw(st.tuples(st.booleans(), st.booleans(),
                                                                        st.booleans())))


@st.composite
def hpd_system(draw):
    n = draw(st.integers(1, 10))
    a = draw(matrix(n, n))
    m = a.T @ a
    m[np.diag_indices(n)] += 1.0
    return m, draw(matrix(n, draw(st.integers(1, 10))))


@settings(max_examples=60, deadline=None, derandomize=True)
@given(hpd_system())
def test_property_cholesky_solvec_invc(system):
    """tests/cholesky.rs:21-67, through strided views; bit-identical here since
    every n is within the exact kernel's range."""
    a, x = system
    chol = llh.cholesky(a)
    assert np.abs(chol @ chol.T - a).max() <= 1e-7
    assert llh.is_triangular(chol, L)
    assert np.array_equal(chol, orc.cholesky(a))
    b = a @ x
    assert np.abs(llh.solvec(np.array(a), b) - x).max() <= 1e-5
    # in place through a negative-stride view of b
    bb = np.array(b)[::-1][::-1]
    aa = np.array(a)[:, ::-1][:, ::-1]
    llh.solvec_inplace(aa, bb)
    assert np.array_equal(bb, orc.solvec(np.array(a), b))
    assert np.abs(a @ llh.invc(a) - np.eye(a.shape[0])).max() <= 1e-7


@settings(max_examples=60, deadline=None, derandomize=True)
@given(st.sampled_from([U, L]), st.integers(1, 10), st.integers(1, 10), st.data())
def test_property_solve_triangular(uplo, n, m, data):
    """tests/triangular.rs:9-47."""
    a = data.draw(matrix(n, n))  # a strided view (inverted axes / transposed)
    # into_triangular on the view, then |diag| < 1 -> 1 (tests/triangular.rs:10-18)
    other = np.tril(np.ones((n, n)), -1) if uplo == U else np.triu(np.ones((n, n)), 1)
    a[other > 0] = 0.0
    d = np.diagonal(a).copy()
    d[np.abs(d) < 1.0] = 1.0
    a[np.diag_indices(n)] = d
    x = data.draw(matrix(n, m))
    b = a @ x
    # The reference's 1e-4 bound on |out - x| holds unless the drawn system is
    # so ill-conditioned that the reference algorithm's own result misses it too
    # (hypothesis finds such draws: diagonal 1, off-diagonal entries near 100);
    # then the bound is a small multiple of the oracle's error.
    ref = orc.solve_triangular(np.array(a), b, uplo)
    tol = max(1e-4, 10 * np.abs(ref - x).max())
    out = llh.solve_triangular(a, b, uplo)
    assert np.abs(out - x).max() <= tol
    out2 = llh.solve_triangular_inplace(a, np.array(b), uplo)
    assert np.abs(out2 - x).max() <= tol


# --- device-resident forms (llh_dev_*) through torch tensors ----------------------

@pytest.mark.parametrize("n,m", [(100, 1), (1000, 16), (2000, 17), (1500, 64), (700, 65)])
@pytest.mark.parametrize("uplo,trans", [(L, False), (U, False), (L, True), (U, True)])
def test_dev_solve_triangular(n, m, uplo, trans):
    """One-launch solve (m <= 64) and blocked solve (m > 64), every (uplo, trans):
    op(A) X = B against the oracle's solve of the same op(A)."""
    import torch
    from linfa_linalg_hip import device as dev
    rng = np.random.default_rng(n + m)
    t = rng.uniform(-1, 1, (n, n)) / np.sqrt(n)
    t[np.diag_indices(n)] = rng.uniform(1, 2, n)
    op_lower = uplo == L
    # op(A)'s triangle is `uplo`; the stored A holds it transposed when trans.
    opa = np.tril(t) if op_lower else np.triu(t)
    stored = opa.T.copy() if trans else opa.copy()
    other = np.triu(np.ones((n, n)), 1) if (op_lower != trans) else np.tril(np.ones((n, n)), -1)
    stored[other > 0] = np.nan  # never read
    b = rng.standard_normal((n, m))
    x_ref = orc.solve_triangular(opa, b, uplo)
    ta = torch.from_numpy(stored).cuda()
    tb = torch.from_numpy(b).cuda()
    dev.solve_triangular(ta, tb, uplo, trans=trans)
    x = tb.cpu().numpy()
    assert np.abs(x - x_ref).max() <= 1e-11 * max(1.0, np.abs(x_ref).max())


@pytest.mark.parametrize("n", [64, 700, 3000])
def test_dev_solvec_and_potrf(n):
    import torch
    from linfa_linalg_hip import device as dev
    a = spd(n, 5 * n)
    b = np.random.default_rng(n).standard_normal((n, 16))
    ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    dev.solvec(ta, tb)
    x = tb.cpu().numpy()
    assert np.linalg.norm(a @ x - b) / (np.linalg.norm(a) * np.linalg.norm(x)) <= 1e-15
    fac = ta.cpu().numpy()
    assert backward_error(a, fac) <= 1e-13
    assert np.array_equal(np.triu(fac, 1), np.triu(a, 1))  # dirty: upper untouched
    tc = torch.from_numpy(a).cuda()
    dev.potrf(tc, zero_upper=True)
    c = tc.cpu().numpy()
    assert np.all(np.triu(c, 1) == 0) and np.array_equal(c, np.tril(fac))


def test_dev_solvec_wide_outer_panels():
    """From n = 24576 on the f64 factor uses 1024-wide outer panels (its SYRK's
    rank); n = 24657 also leaves a ragged last panel and block.  Too large for
    the oracle: the size-independent checks are the backward residual of the
    solve and L L^T = A on a sample of rows, computed on the GPU in f64."""
    import torch
    from linfa_linalg_hip import device as dev
    n, m = 24657, 8
    g = torch.Generator(device="cuda").manual_seed(11)
    bm = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g).sub_(0.5)
    a = torch.matmul(bm, bm.T).div_(n)
    del bm
    a.diagonal().add_(1.0)
    rhs = torch.randn(n, m, dtype=torch.float64, device="cuda", generator=g)
    fac, x = a.clone(), rhs.clone()
    dev.solvec(fac, x)
    res = (torch.matmul(a, x) - rhs).norm() / (a.norm() * x.norm())
    assert float(res) <= 1e-15, float(res)
    rows = torch.tensor([0, 1, 511, 512, 1023, 1024, 1025, 12345, 24575, 24576, n - 2, n - 1], device="cuda")
    l = torch.tril(fac)
    lr = l[rows]
    err = (torch.matmul(lr, l.T) - a[rows]).abs().max() / a.abs().max()
    assert float(err) <= 1e-14, float(err)
    assert torch.equal(torch.triu(fac, 1), torch.triu(a, 1))  # dirty: upper untouched


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-4)])
@pytest.mark.parametrize("n", [5, 70, 300])
def test_solve_triangular_system_matches_oracle(dtype, tol, n):
    """QR's back-substitutions: R x = b with pivots |diag_i| from a separate
    vector (Upper), and R^T x = b through a transposed view (Lower); a's own
    diagonal is poisoned with NaN and must not be read."""
    rng = np.random.default_rng(n)
    r = (np.triu(rng.uniform(-1, 1, (n, n)), 1) / np.sqrt(n)).astype(dtype)
    d = (rng.uniform(0.5, 2.0, n) * rng.choice([-1, 1], n)).astype(dtype)
    r[np.diag_indices(n)] = np.nan
    b = rng.standard_normal((n, 4)).astype(dtype)
    for a, uplo, oracle_uplo in ((r, llh.UPLO.Upper, 0), (r.T, llh.UPLO.Lower, 1)):
        x = llh.solve_triangular_system(a, np.array(b), uplo, d, True)
        x_ref = orc.solve_triangular_system(a, np.array(b), oracle_uplo, d, True)
        assert np.all(np.isfinite(x))
        assert np.abs(x - x_ref).max() <= tol * max(1.0, np.abs(x_ref).max()), (n, uplo)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-4)])
@pytest.mark.parametrize("n", [129, 200, 448, 512, 513])
@pytest.mark.parametrize("m", [1, 2, 3, 8, 9])
def test_dev_solvec_solve_pair_forms(n, m, dtype, tol):
    """solvec's substitutions run in one workgroup with B in LDS up to 512 rows
    and 8 right-hand sides (1, 2, 4 or 8 columns per pass; rows of L 16-byte
    aligned or not), else as the two one-launch triangles: both give the
    oracle's x, B with rows wider than m (strided) keeps its padding, and after a
    failed factor B is untouched (cholesky.rs:140)."""
    import torch
    from linfa_linalg_hip import device as dev
    a = spd(n, 11 * n + m, dtype)
    b = np.random.default_rng(n + m).standard_normal((n, m)).astype(dtype)
    tb_wide = torch.full((n, m + 3), 7.0, dtype=torch.float64 if dtype == np.float64 else torch.float32,
                         device="cuda")
    tb_wide[:, :m] = torch.from_numpy(b).cuda()
    ta = torch.from_numpy(a).cuda()
    dev.solvec(ta, tb_wide[:, :m])
    x = tb_wide[:, :m].cpu().numpy().astype(np.float64)
    x_ref = orc.solvec(np.array(a, dtype=np.float64), np.array(b, dtype=np.float64))
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) <= tol, (n, m)
    assert torch.all(tb_wide[:, m:] == 7.0)
    bad = np.array(a)
    bad[n - 2, n - 2] = -1.0
    tb2 = torch.from_numpy(b).cuda()
    with pytest.raises(llh.NotPositiveDefinite):
        dev.solvec(torch.from_numpy(bad).cuda(), tb2)
    assert np.array_equal(tb2.cpu().numpy(), b)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 5e-4)])
@pytest.mark.parametrize("n", [1025, 1600])
@pytest.mark.parametrize("m", [1, 16, 17, 64])
def test_dev_solvec_fused_forward(n, m, dtype, tol):
    """From two outer panels of 512 columns on, with at most 64 right-hand
    sides, solvec runs the forward substitution beside the factor, one panel at
    a time (a triangle over the panel's rows, then one product for the rows
    below), into a workspace: x still matches the oracle's (ragged last panel,
    1 / 16 / 17 / 64 columns, B strided), B's padding is kept, and after a
    failed factor, early or in the last panel, B is untouched (cholesky.rs:140);
    the host-memory form takes the same path."""
    import torch
    from linfa_linalg_hip import device as dev
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    a = spd(n, 7 * n + m, dtype)
    b = np.random.default_rng(3 * n + m).standard_normal((n, m)).astype(dtype)
    x_ref = orc.solvec(np.array(a, dtype=np.float64), np.array(b, dtype=np.float64))
    tb_wide = torch.full((n, m + 5), 7.0, dtype=tdt, device="cuda")
    tb_wide[:, :m] = torch.from_numpy(b).cuda()
    dev.solvec(torch.from_numpy(a).cuda(), tb_wide[:, :m])
    x = tb_wide[:, :m].cpu().numpy().astype(np.float64)
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) <= tol, (n, m)
    assert torch.all(tb_wide[:, m:] == 7.0)
    xp = np.asarray(llh.solvec(np.array(a), np.array(b)), dtype=np.float64)
    assert np.linalg.norm(xp - x_ref) / np.linalg.norm(x_ref) <= tol, (n, m)
    for row in (100, n - 3):
        bad = np.array(a)
        bad[row, row] = -1.0
        tb2 = torch.from_numpy(b).cuda()
        with pytest.raises(llh.NotPositiveDefinite):
            dev.solvec(torch.from_numpy(bad).cuda(), tb2)
        assert np.array_equal(tb2.cpu().numpy(), b), row

"""Mixed precision: f32 MFMA factor with f64 diagonal blocks and f64 iterative
refinement (llh_dev_solve_mixed / llh_factor_*), checked against the oracle's
f64 solve of the same system."""
import numpy as np
import pytest

import linfa_linalg_hip as llh  # noqa: F401  (loads torch's HIP runtime first)
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def spd_normal(n, lam, seed):
    """A = B B^T / n + lam I with B ~ N(0, 1) (the mixed-precision workload)."""
    rng = np.random.default_rng(seed)
    b = rng.standard_normal((n, n))
    return b @ b.T / n + lam * np.eye(n)


def tt(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


@pytest.mark.parametrize("n", [300, 1000, 2100])
@pytest.mark.parametrize("lam", [1.0, 1e-2, 1e-4])
def test_f32_ir_matches_f64_oracle(n, lam):
    from linfa_linalg_hip import device as dev
    import torch
    a = spd_normal(n, lam, n)
    b = np.random.default_rng(n + 1).standard_normal((n, 16))
    x = torch.zeros((n, 16), dtype=torch.float64, device="cuda")
    info = dev.solve_mixed(tt(a), tt(b), x, prec="f32", fallback=False)
    assert info["converged"] and not info["fell_back"] and info["precision_used"] == "f32"
    assert info["iterations"] >= 1
    xs = x.cpu().numpy()
    x_ref = orc.solvec(np.array(a), b)
    cond = np.linalg.cond(a) if n <= 1000 else 4.0 / lam + 1.0
    err = np.linalg.norm(xs - x_ref) / np.linalg.norm(x_ref)
    # both solutions are backward stable in f64: they agree to ~cond * eps64
    assert err <= max(1e-12, 50 * cond * np.finfo(np.float64).eps), (err, cond, info)
    assert info["residual"] <= 10 * np.sqrt(n) * np.finfo(np.float64).eps
    # ||A||_inf from the split row-sum walk over the lower triangle
    anorm = np.abs(a).sum(axis=1).max()
    assert abs(info["anorm"] - anorm) <= 1e-13 * anorm, (info["anorm"], anorm)


@pytest.mark.parametrize("n", [1500, 3000])
def test_f32_ir_hardest_lambda(n):
    """The hardest refinement case of the mixed-precision workload, lambda =
    1e-6 (cond ~4e6): the f32-updated factor still refines to f64 backward
    accuracy; the solution agrees with the oracle's to the measured
    cond * eps64 (both are backward-stable f64 solutions)."""
    from linfa_linalg_hip import device as dev
    import torch
    lam = 1e-6
    a = spd_normal(n, lam, n + 11)
    b = np.random.default_rng(n + 12).standard_normal((n, 16))
    x = torch.zeros((n, 16), dtype=torch.float64, device="cuda")
    info = dev.solve_mixed(tt(a), tt(b), x, prec="f32", fallback=False)
    assert info["converged"] and not info["fell_back"] and info["precision_used"] == "f32", info
    cond = np.linalg.cond(a)
    assert cond > 1e5, cond  # the workload is as ill-conditioned as intended
    x_ref = orc.solvec(np.array(a), b)
    err = np.linalg.norm(x.cpu().numpy() - x_ref) / np.linalg.norm(x_ref)
    assert err <= max(1e-12, 50 * cond * np.finfo(np.float64).eps), (err, cond, info)
    assert info["residual"] <= 10 * np.sqrt(n) * np.finfo(np.float64).eps


@pytest.mark.parametrize("n", [1600, 3000])
def test_f16_ir_lambda_1e2(n):
    """FP16+IR at the mixed-precision workload's lambda = 1e-2 (cond ~4e2):
    more refinement steps than at lambda = 1, same f64 accuracy."""
    from linfa_linalg_hip import device as dev
    import torch
    a = spd_normal(n, 1e-2, n + 13)
    b = np.random.default_rng(n + 14).standard_normal((n, 16))
    x = torch.zeros((n, 16), dtype=torch.float64, device="cuda")
    info = dev.solve_mixed(tt(a), tt(b), x, prec="f16", fallback=False)
    assert info["converged"] and not info["fell_back"] and info["precision_used"] == "f16", info
    cond = np.linalg.cond(a)
    x_ref = orc.solvec(np.array(a), b)
    err = np.linalg.norm(x.cpu().numpy() - x_ref) / np.linalg.norm(x_ref)
    assert err <= max(1e-12, 50 * cond * np.finfo(np.float64).eps), (err, cond, info)
    assert info["residual"] <= 10 * np.sqrt(n) * np.finfo(np.float64).eps


def test_f64_factor_and_reuse():
    from linfa_linalg_hip import device as dev
    import torch
    n = 900
    a = spd_normal(n, 1.0, 3)
    ta = tt(a)
    for prec in ("f64", "f32"):
        f = dev.Factor(ta, prec=prec)
        assert f.precision == prec
        for seed in (1, 2):
            b = np.random.default_rng(seed).standard_normal((n, 5))
            x = torch.zeros((n, 5), dtype=torch.float64, device="cuda")
            info = f.solve(tt(b), x)
            assert info["converged"]
            x_ref = orc.solvec(np.array(a), b)
            assert np.linalg.norm(x.cpu().numpy() - x_ref) / np.linalg.norm(x_ref) <= 1e-12
        f.close()


def test_fallback_to_f64_on_f32_breakdown():
    """Eigenvalues down to 1e-11: the f32 factor loses positive definiteness (or
    refinement cannot converge); with fallback the f64 factor solves it."""
    from linfa_linalg_hip import device as dev
    import torch
    n = 400
    rng = np.random.default_rng(5)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.logspace(0, -11, n)
    a = (q * ev) @ q.T
    a = (a + a.T) / 2
    b = rng.standard_normal((n, 2))
    x = torch.zeros((n, 2), dtype=torch.float64, device="cuda")
    info = dev.solve_mixed(tt(a), tt(b), x, prec="f32", fallback=True)
    assert info["fell_back"] and info["precision_used"] == "f64"
    assert info["residual"] <= 10 * np.sqrt(n) * np.finfo(np.float64).eps
    with pytest.raises(llh.LinalgError):
        x2 = torch.zeros((n, 2), dtype=torch.float64, device="cuda")
        dev.solve_mixed(tt(a), tt(b), x2, prec="f32", fallback=False, max_iter=5)


@pytest.mark.parametrize("prec", ["f32", "f16"])
def test_fallback_at_fused_size(prec):
    """From two outer panels of 512 columns on, the first forward substitution
    runs beside the low-precision factor; when that factor breaks down (or its
    refinement stalls) the f64 fallback solves from scratch and still reaches
    f64 accuracy, and the well-conditioned case at the same size matches the
    oracle."""
    from linfa_linalg_hip import device as dev
    import torch
    n = 1100
    rng = np.random.default_rng(17)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    a = (q * np.logspace(0, -11, n)) @ q.T
    a = (a + a.T) / 2
    b = rng.standard_normal((n, 3))
    x = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    info = dev.solve_mixed(tt(a), tt(b), x, prec=prec, fallback=True)
    assert info["fell_back"] and info["precision_used"] == "f64", info
    assert info["residual"] <= 10 * np.sqrt(n) * np.finfo(np.float64).eps
    good = spd_normal(n, 1.0, 23)
    x2 = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    info2 = dev.solve_mixed(tt(good), tt(b), x2, prec=prec, fallback=False)
    assert info2["converged"] and not info2["fell_back"], info2
    x_ref = orc.solvec(np.array(good), b)
    assert np.linalg.norm(x2.cpu().numpy() - x_ref) / np.linalg.norm(x_ref) <= 1e-12


def test_not_positive_definite():
    from linfa_linalg_hip import device as dev
    import torch
    n = 300
    a = spd_normal(n, 1.0, 9)
    a[150, 150] = -5.0
    x = torch.zeros((n, 1), dtype=torch.float64, device="cuda")
    with pytest.raises(llh.NotPositiveDefinite) as ei:
        dev.solve_mixed(tt(a), tt(np.ones((n, 1))), x, prec="f32", fallback=True)
    assert ei.value.column == 150


# --- f16 trailing updates (LLH_PREC_F16) --------------------------------------

@pytest.mark.parametrize("n", [1600, 2100, 3000])
@pytest.mark.parametrize("lam", [1.0, 1e-1])
def test_f16_ir_matches_f64_oracle(n, lam):
    """Well-conditioned systems (cond ~5 and ~40): the f16-updated factor
    refines to f64 accuracy, with more steps than the f32 one.  (The first
    512-column panel is updated in f32 by the look-ahead, so f16 products
    appear from the second outer step on: n > 1024 is needed to meet them.)"""
    from linfa_linalg_hip import device as dev
    import torch
    a = spd_normal(n, lam, n + 7)
    b = np.random.default_rng(n + 8).standard_normal((n, 16))
    x = torch.zeros((n, 16), dtype=torch.float64, device="cuda")
    info = dev.solve_mixed(tt(a), tt(b), x, prec="f16", fallback=False)
    assert info["converged"] and not info["fell_back"] and info["precision_used"] == "f16", info
    assert info["iterations"] >= 2
    x_ref = orc.solvec(np.array(a), b)
    err = np.linalg.norm(x.cpu().numpy() - x_ref) / np.linalg.norm(x_ref)
    assert err <= 1e-12, (err, info)
    assert info["residual"] <= 10 * np.sqrt(n) * np.finfo(np.float64).eps
    # the refinement measures the scaled system: ||S A S||_inf, S = diag(A)^-1/2
    sc = 1.0 / np.sqrt(np.diag(a))
    anorm = (np.abs(a) * sc[:, None] * sc[None, :]).sum(axis=1).max()
    assert abs(info["anorm"] - anorm) <= 1e-13 * anorm, (info["anorm"], anorm)


def test_f16_diagonal_scaling():
    """A = D A0 D with D from 1e-4 to 1e4: unscaled, L's entries would span
    ~1e-6 .. 1e4, the small ones below f16's normal range; the factor of
    diag(s) A diag(s) has a unit diagonal and |L| <= 1.  A is ill-conditioned
    only through D, so x is compared with a known solution in both the plain
    and the D-scaled norm, against the oracle's own errors."""
    from linfa_linalg_hip import device as dev
    import torch
    n = 2100
    rng = np.random.default_rng(22)
    d = np.logspace(-4, 4, n)
    a0 = spd_normal(n, 1.0, 21)
    a = (d[:, None] * a0) * d[None, :]
    x_true = rng.standard_normal((n, 4)) / d[:, None]
    b = a @ x_true
    x = torch.zeros((n, 4), dtype=torch.float64, device="cuda")
    info = dev.solve_mixed(tt(a), tt(b), x, prec="f16", fallback=False)
    assert info["converged"] and info["precision_used"] == "f16", info
    assert info["residual"] <= 10 * np.sqrt(n) * np.finfo(np.float64).eps
    xs = x.cpu().numpy()
    x_ref = orc.solvec(np.array(a), b)
    for w in (np.ones(n), d):  # plain and D-scaled errors
        e_gpu = np.linalg.norm(w[:, None] * (xs - x_true)) / np.linalg.norm(w[:, None] * x_true)
        e_orc = np.linalg.norm(w[:, None] * (x_ref - x_true)) / np.linalg.norm(w[:, None] * x_true)
        assert e_gpu <= 10 * e_orc + 1e-13, (e_gpu, e_orc, info)


def test_f16_factor_reuse():
    from linfa_linalg_hip import device as dev
    import torch
    n = 1300
    a = spd_normal(n, 1.0, 31)
    ta = tt(a)
    f = dev.Factor(ta, prec="f16", fallback=False)
    assert f.precision == "f16"
    for seed in (1, 2):
        b = np.random.default_rng(seed).standard_normal((n, 3))
        x = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
        info = f.solve(tt(b), x)
        assert info["converged"], info
        x_ref = orc.solvec(np.array(a), b)
        assert np.linalg.norm(x.cpu().numpy() - x_ref) / np.linalg.norm(x_ref) <= 1e-12
    f.close()


def test_f16_ill_conditioned_falls_back():
    """Eigenvalues down to 1e-12: far beyond what the f16-updated factor
    refines (it breaks down, or the stall test ends the refinement); without
    fallback that is an error, with it the f64 factor produces the result."""
    from linfa_linalg_hip import device as dev
    import torch
    n = 1600
    rng = np.random.default_rng(41)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    a = (q * np.logspace(0, -12, n)) @ q.T
    a = (a + a.T) / 2
    b = rng.standard_normal((n, 2))
    x = torch.zeros((n, 2), dtype=torch.float64, device="cuda")
    with pytest.raises(llh.LinalgError):
        dev.solve_mixed(tt(a), tt(b), x, prec="f16", fallback=False, max_iter=10)
    x = torch.zeros((n, 2), dtype=torch.float64, device="cuda")
    info = dev.solve_mixed(tt(a), tt(b), x, prec="f16", fallback=True)
    assert info["fell_back"] and info["precision_used"] == "f64" and info["converged"], info
    assert info["residual"] <= 10 * np.sqrt(n) * np.finfo(np.float64).eps


def test_f16_not_positive_definite():
    from linfa_linalg_hip import device as dev
    import torch
    n = 1700
    a = spd_normal(n, 1.0, 51)
    a[1600, 1600] = -5.0
    x = torch.zeros((n, 1), dtype=torch.float64, device="cuda")
    with pytest.raises(llh.NotPositiveDefinite) as ei:
        dev.solve_mixed(tt(a), tt(np.ones((n, 1))), x, prec="f16", fallback=False)
    assert ei.value.column == 1600

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

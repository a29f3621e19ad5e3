"""Full-size parity: the seeded C2 / C3 workloads solved on the GPU and compared
with the committed solutions of the multi-threaded oracle (bit-identical to the
reference's operation order, cholesky.rs:136-144 with triangular.rs:95-144).

- The device generator reproduces the oracle's generator bit for bit (checked
  in full at moderate n, and on sampled entries of every full-size input).
- C2 (n = 16384, 64 RHS, lambda = 0.1): the plain f64 solve and the FP32+IR
  solve are within 1e-12 of the oracle's x (solution error
  ||x - x_ref|| / ||x_ref||), with the backward residual and sampled rows of
  L L^T checked too.
- C3 (n = 32768, 16 RHS, lambda = 1 .. 1e-6): the FP32+IR solve (f32 factor
  with 1024-wide outer panels) against the oracle's x, to max(1e-12,
  50 cond eps64) -- two backward-stable f64 solutions agree to about cond eps64,
  cond = ||A|| ||A^-1|| ~ 4 / lambda + 1 for this workload.
"""
import json
import os

import numpy as np
import pytest

import linfa_linalg_hip as llh  # noqa: F401  (loads torch's HIP runtime first)
from oracle import par

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "full_size")
EPS = np.finfo(np.float64).eps


def meta():
    with open(os.path.join(FIX, "meta.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("dist", [0, 1])
@pytest.mark.parametrize("rows,cols", [(1, 1), (300, 257), (64, 1000)])
def test_device_generator_matches_oracle(dist, rows, cols):
    from linfa_linalg_hip import device as dev
    for seed, stream in ((0, 0), (20241202, 1), (2**63 + 5, 7)):
        x = dev.gen_matrix(dist, seed, stream, rows, cols).cpu().numpy()
        assert np.array_equal(x, par.gen_matrix(dist, seed, stream, rows, cols)), (seed, stream)


@pytest.mark.parametrize("n,dist,lam", [(1, 0, 0.1), (63, 1, 1.0), (333, 0, 0.1), (1000, 0, 1e-6), (1029, 1, 1.0)])
def test_device_spd_matches_oracle(n, dist, lam):
    from linfa_linalg_hip import device as dev
    a = dev.gen_spd(dist, 99 + n, n, lam).cpu().numpy()
    assert np.array_equal(a, par.gen_spd(dist, 99 + n, n, lam))


def _inputs(key):
    from linfa_linalg_hip import device as dev
    cfg = meta()[key]
    n, m = cfg["n"], cfg["m"]
    a = dev.gen_spd(0, cfg["seed"], n, cfg["lam"])
    samp = a[[s[0] for s in cfg["a_samples"]], [s[1] for s in cfg["a_samples"]]].cpu().numpy()
    want = np.array([float.fromhex(s[2]) for s in cfg["a_samples"]])
    assert np.array_equal(samp, want), "device generator differs from the oracle's at full size"
    b = dev.gen_matrix(0, cfg["seed"], cfg["rhs_stream"], n, m)
    x_ref = np.load(os.path.join(FIX, cfg["x_ref"]))
    return cfg, a, b, x_ref


def _sol_err(x, x_ref):
    k = x_ref.shape[1]
    return float(np.linalg.norm(x[:, :k] - x_ref) / np.linalg.norm(x_ref))


def _residual(a, x, b):
    import torch
    r = b - a @ x
    anorm = a.abs().sum(dim=1).max()
    return float((r.abs().max(dim=0).values / (anorm * x.abs().max(dim=0).values)).max())


def test_c2_f64_full_size():
    import torch
    from linfa_linalg_hip import device as dev
    cfg, a, b, x_ref = _inputs("c2_lam0.1")
    n = cfg["n"]
    fac = a.clone()
    x = b.clone()
    dev.solvec(fac, x)
    torch.cuda.synchronize()
    err = _sol_err(x.cpu().numpy(), x_ref)
    assert err <= 1e-12, err
    assert _residual(a, x, b) <= 10 * np.sqrt(n) * EPS
    # sampled rows of L L^T against A (backward error of the factor)
    rows = np.random.default_rng(1).choice(n, 32, replace=False).tolist() + [n - 1]
    for r in rows:
        lr = fac[r, : r + 1]  # row r of L: the lower part of the dirty factor's row
        prod = torch.tril(fac[: r + 1, : r + 1]) @ lr
        rel = float((prod - a[r, : r + 1]).norm() / a[r, : r + 1].norm())
        assert rel <= 1e-13, (r, rel)


def test_c2_fp32_ir_full_size():
    import torch
    from linfa_linalg_hip import device as dev
    cfg, a, b, x_ref = _inputs("c2_lam0.1")
    x = torch.zeros_like(b)
    info = dev.solve_mixed(a, b, x, prec="f32", fallback=False)
    assert info["converged"] and not info["fell_back"] and info["precision_used"] == "f32", info
    err = _sol_err(x.cpu().numpy(), x_ref)
    assert err <= 1e-12, (err, info)
    assert _residual(a, x, b) <= 10 * np.sqrt(cfg["n"]) * EPS


@pytest.mark.parametrize("lam", [1.0, 1e-2, 1e-4, 1e-6])
def test_c3_fp32_ir_full_size(lam):
    import torch
    from linfa_linalg_hip import device as dev
    key = f"c3_lam{lam:g}"
    if key not in meta():
        pytest.skip(f"no fixture for {key}")
    cfg, a, b, x_ref = _inputs(key)
    x = torch.zeros_like(b)
    info = dev.solve_mixed(a, b, x, prec="f32", fallback=False)
    assert info["converged"] and not info["fell_back"], info
    cond = 4.0 / lam + 1.0
    err = _sol_err(x.cpu().numpy(), x_ref)
    assert err <= max(1e-12, 50 * cond * EPS), (lam, err, info)
    # the refinement's own stopping test: ||r|| <= ||A|| ||x|| eps64 sqrt(n) per column
    assert _residual(a, x, b) <= 10 * np.sqrt(cfg["n"]) * EPS
    # the plain f64 solve of the same system, for the record of both errors
    fac = a.clone()
    x64 = b.clone()
    dev.solvec(fac, x64)
    err64 = _sol_err(x64.cpu().numpy(), x_ref)
    assert err64 <= max(1e-12, 50 * cond * EPS), (lam, err64)


def test_f32_factor_wide_panels_backward_error():
    """The f32 factor at n >= 24576 (1024-wide outer panels, the path behind the
    FP32+IR figure): sampled rows of L L^T against A, and a solve's residual."""
    import torch
    from linfa_linalg_hip import device as dev
    n = 24657
    a = dev.gen_spd(0, 5, n, 1.0)
    a32 = a.float()
    fac = a32.clone()
    dev.potrf(fac)
    torch.cuda.synchronize()
    rows = np.random.default_rng(2).choice(n, 24, replace=False).tolist() + [n - 1]
    for r in rows:
        lr = fac[r, : r + 1].double()
        prod = torch.tril(fac[: r + 1, : r + 1]).double() @ lr
        rel = float((prod - a32[r, : r + 1].double()).norm() / a32[r, : r + 1].double().norm())
        assert rel <= 2e-6, (r, rel)
    b = dev.gen_matrix(0, 5, 1, n, 3).float()
    x = b.clone()
    dev.solve_triangular(fac, x, llh.UPLO.Lower)
    dev.solve_triangular(fac, x, llh.UPLO.Upper, trans=True)
    r = (b.double() - a32.double() @ x.double()).abs().max()
    assert float(r / (a32.double().abs().sum(1).max() * x.double().abs().max())) <= 1e-5

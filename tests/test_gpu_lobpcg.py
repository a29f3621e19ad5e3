"""LOBPCG's Cholesky consumers (lobpcg/algorithm.rs:60-97) on the device:
orthonormalize and apply_constraints, against the oracle's cholesky and
solve_triangular composed as the reference composes them (the Gram products
V^T V and Y^T V, which the reference takes from ndarray's dot, are formed in
f64 with numpy: parity is to tolerance, not bitwise)."""
import numpy as np
import pytest

import linfa_linalg_hip as llh
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

LOWER = 1


def ref_orthonormalize(v):
    g = v.T.astype(np.float64) @ v.astype(np.float64)
    l = orc.cholesky(g.astype(v.dtype))  # cholesky_into (:81)
    vt = v.T  # v.reversed_axes() (:90), an F-order view
    u = orc.solve_triangular(l, np.array(vt), LOWER).T  # solve_triangular_into(v_t, Lower) (:91-93)
    return u, l


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 5e-3)])
@pytest.mark.parametrize("n,k", [(1000, 1), (3000, 7), (20000, 64), (5000, 96), (700, 130)])
def test_orthonormalize_matches_reference(dtype, tol, n, k):
    import torch
    from linfa_linalg_hip import device as dev
    rng = np.random.default_rng(k)
    v = rng.standard_normal((n, k)).astype(dtype)
    tv = torch.from_numpy(v).cuda()
    u, r = dev.orthonormalize(tv)
    u, r = u.cpu().numpy(), r.cpu().numpy()
    u_ref, l_ref = ref_orthonormalize(v)
    assert np.array_equal(np.triu(r, 1), np.zeros_like(r))  # clean factor
    assert np.abs(r - l_ref).max() <= tol * np.abs(l_ref).max()
    assert np.abs(u - u_ref).max() <= tol * max(1.0, np.abs(u_ref).max())
    # invariants: U^T U = I and U L^T = V
    assert np.abs(u.T.astype(np.float64) @ u - np.eye(k)).max() <= 50 * tol
    assert np.abs(u.astype(np.float64) @ r.T - v).max() <= tol * np.abs(v).max() * 10


def test_orthonormalize_rank_deficient_is_not_pd():
    import torch
    from linfa_linalg_hip import device as dev
    rng = np.random.default_rng(3)
    v = rng.standard_normal((500, 5))
    v[:, 3] = v[:, 1]  # V^T V singular
    tv = torch.from_numpy(v).cuda()
    with pytest.raises(llh.NotPositiveDefinite):
        dev.orthonormalize(tv)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 5e-3)])
@pytest.mark.parametrize("n,k,ky", [(2000, 4, 3), (10000, 16, 40), (3000, 80, 70)])
def test_apply_constraints_matches_reference(dtype, tol, n, k, ky):
    import torch
    from linfa_linalg_hip import device as dev
    rng = np.random.default_rng(n + k)
    y = rng.standard_normal((n, ky)).astype(dtype)
    v = rng.standard_normal((n, k)).astype(dtype)
    l_yy = orc.cholesky((y.T.astype(np.float64) @ y).astype(dtype))
    g = (y.T.astype(np.float64) @ v).astype(dtype)  # y.t().dot(&v) (:66)
    u = orc.solve_triangular(l_yy, g, LOWER)  # (:68-70)
    v_ref = v.astype(np.float64) - y.astype(np.float64) @ u  # general_mat_mul(-1, y, u, 1, v) (:73)
    tv = torch.from_numpy(v).cuda()
    dev.apply_constraints(tv, torch.from_numpy(l_yy).cuda(), torch.from_numpy(y).cuda())
    out = tv.cpu().numpy()
    assert np.abs(out - v_ref).max() <= tol * max(1.0, np.abs(v_ref).max())


def test_apply_constraints_projects_out_orthonormal_y():
    """With orthonormal Y (so L_yy = I) the update is the projection onto Y's
    complement: afterwards Y^T V = 0."""
    import torch
    from linfa_linalg_hip import device as dev
    rng = np.random.default_rng(5)
    n, k, ky = 4000, 6, 9
    y, _ = np.linalg.qr(rng.standard_normal((n, ky)))
    y = np.ascontiguousarray(y)
    v = rng.standard_normal((n, k))
    l_yy = orc.cholesky(y.T @ y)
    tv = torch.from_numpy(v).cuda()
    dev.apply_constraints(tv, torch.from_numpy(l_yy).cuda(), torch.from_numpy(y).cuda())
    out = tv.cpu().numpy()
    assert np.abs(y.T @ out).max() <= 1e-12 * np.sqrt(n) * np.abs(v).max()

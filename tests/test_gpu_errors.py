"""Error paths forced from the host, so every kernel still exits normally.

- A one-launch solve whose bounded flag wait gives up must surface as
  LLH_ERR_HIP from every multi-sweep call (device solvec, the refined mixed
  solve, a server request), never as LLH_OK with an unfinished x; afterwards the
  same context or server must solve correctly again.
- A server step whose descriptor staging or event creation fails finishes its
  own requests with LLH_ERR_HIP and the server carries on.
- A device-resident solvec that meets NotPositiveDefinite leaves B as the
  caller passed it (the reference returns before touching b, cholesky.rs:140).
- An f32 server entry made from a caller's device matrix stays correct after
  the caller frees that matrix (the entry keeps its own copy of A).
"""
import numpy as np
import pytest

import linfa_linalg_hip as llh
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def spd(n, seed, lam=1.0):
    rng = np.random.default_rng(seed)
    b = rng.uniform(-1, 1, (n, n))
    return b @ b.T / n + lam * np.eye(n)


def tt(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


@pytest.fixture
def spin_reset():
    from linfa_linalg_hip import device as dev
    yield dev
    dev.set_solve_spin_limit(0)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_dev_solvec_reports_a_wait_that_gave_up(spin_reset, dtype):
    dev = spin_reset
    import torch
    n, m = 640, 5
    a = spd(n, 1)
    b = np.random.default_rng(2).standard_normal((n, m))
    tdt = torch.float64 if dtype == "f64" else torch.float32
    dev.set_solve_spin_limit(-1)
    ta, tb = tt(a).to(tdt), tt(b).to(tdt)
    with pytest.raises(llh.HipError, match="wait exceeded"):
        dev.solvec(ta, tb)
    # the same context solves correctly once the bound is restored: its flag
    # epochs were not reused
    dev.set_solve_spin_limit(0)
    ta, tb = tt(a).to(tdt), tt(b).to(tdt)
    dev.solvec(ta, tb)
    x_ref = orc.solvec(np.array(a), b)
    err = np.linalg.norm(tb.double().cpu().numpy() - x_ref) / np.linalg.norm(x_ref)
    assert err <= (1e-12 if dtype == "f64" else 1e-4), err


def test_dev_solve_triangular_reports_a_wait_that_gave_up(spin_reset):
    dev = spin_reset
    n = 300
    lo = np.linalg.cholesky(spd(n, 4))
    b = np.random.default_rng(5).standard_normal((n, 3))
    dev.set_solve_spin_limit(-1)
    with pytest.raises(llh.HipError, match="wait exceeded"):
        dev.solve_triangular(tt(lo), tt(b), llh.UPLO.Lower)
    dev.set_solve_spin_limit(0)
    tb = tt(b)
    dev.solve_triangular(tt(lo), tb, llh.UPLO.Lower)
    ref = orc.solve_triangular(np.array(lo), b, llh.UPLO.Lower)
    assert np.linalg.norm(tb.cpu().numpy() - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("prec", ["f32", "f16"])
def test_mixed_solve_reports_a_wait_that_gave_up(spin_reset, prec):
    dev = spin_reset
    import torch
    n, m = 700, 4
    a = spd(n, 6)
    b = np.random.default_rng(7).standard_normal((n, m))
    x = torch.zeros((n, m), dtype=torch.float64, device="cuda")
    dev.set_solve_spin_limit(-1)
    # the refinement stops at the first failed sweep: HipError, not NoConvergence
    # and not a fallback to f64
    with pytest.raises(llh.HipError, match="wait exceeded"):
        dev.solve_mixed(tt(a), tt(b), x, prec=prec, fallback=True)
    dev.set_solve_spin_limit(0)
    info = dev.solve_mixed(tt(a), tt(b), x, prec=prec, fallback=False)
    assert info["converged"]
    x_ref = orc.solvec(np.array(a), b)
    err = np.linalg.norm(x.cpu().numpy() - x_ref) / np.linalg.norm(x_ref)
    assert err <= 1e-12, err


def test_dev_solvec_not_pd_leaves_b_untouched():
    from linfa_linalg_hip import device as dev
    for n in (100, 700):  # exact-order LDS factor and blocked factor
        a = spd(n, 8)
        a[n // 2, n // 2] = -1.0  # NotPositiveDefinite at row n // 2 at the latest
        b = np.random.default_rng(9).standard_normal((n, 3))
        tb = tt(b)
        with pytest.raises(llh.NotPositiveDefinite):
            dev.solvec(tt(a), tb)
        assert np.array_equal(tb.cpu().numpy(), b), n


def _server_case(srv, mid, a, b, prec="f64"):
    x = np.zeros_like(b)
    r = srv.wait(srv.submit(mid, a, b, x, precision=prec))
    return r, x


def test_server_request_reports_a_wait_that_gave_up():
    from linfa_linalg_hip import device as dev
    n = 600  # above the one-workgroup substitution's 512 rows: one-launch sweeps
    a = spd(n, 10)
    b = np.random.default_rng(11).standard_normal((n, 4))
    x_ref = orc.solvec(np.array(a), b)
    srv = dev.Server(streams=2)
    try:
        srv.inject("solve_spin", -1)
        for mid, prec in ((1, "f64"), (2, "f32")):
            r, _ = _server_case(srv, mid, a, b, prec)
            assert r["status"] == 17 and r["error"] == "HipError", (prec, r)
        srv.inject("solve_spin", 0)
        for mid, prec in ((1, "f64"), (2, "f32"), (3, "f64")):
            r, x = _server_case(srv, mid, a, b, prec)
            assert r["status"] == 0, (mid, r)
            err = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
            assert err <= 1e-12, (mid, prec, err)
    finally:
        srv.close()


def test_server_failed_staging_fails_its_step_only():
    from linfa_linalg_hip import device as dev
    srv = dev.Server(streams=2)
    try:
        mats = {k: spd(40 + 20 * k, 20 + k) for k in range(3)}
        rhs = {k: np.random.default_rng(30 + k).standard_normal((mats[k].shape[0], 2)) for k in mats}
        srv.inject("staging", 1)
        r, _ = _server_case(srv, 0, mats[0], rhs[0])
        assert r["status"] == 17 and r["batched"], r
        # the failed entry is not reused: the same id factors again, correctly
        for k in mats:
            r, x = _server_case(srv, k, mats[k], rhs[k])
            assert r["status"] == 0, (k, r)
            assert np.array_equal(x, orc.solvec(np.array(mats[k]), rhs[k])), k
        assert srv.stats()["completed"] == 4
    finally:
        srv.close()


def test_server_failed_event_fails_one_request():
    from linfa_linalg_hip import device as dev
    srv = dev.Server(streams=2)
    try:
        n = 300
        a = spd(n, 40)
        b = np.random.default_rng(41).standard_normal((n, 2))
        x_ref = orc.solvec(np.array(a), b)
        for k in (1, 2):  # the entry's event, then the request's own event
            srv.inject("event", k)
            r, _ = _server_case(srv, 100 + k, a, b)
            assert r["status"] == 17, (k, r)
            srv.inject("event", 0)
            r, x = _server_case(srv, 100 + k, a, b)
            assert r["status"] == 0, (k, r)
            assert np.linalg.norm(x - x_ref) <= 1e-12 * np.linalg.norm(x_ref)
    finally:
        srv.close()


def test_server_f32_hit_after_the_device_matrix_is_freed():
    import torch
    from linfa_linalg_hip import device as dev
    n = 600
    a = spd(n, 50)
    srv = dev.Server(streams=2)
    try:
        ta = tt(a)
        b1 = np.random.default_rng(51).standard_normal((n, 3))
        r, x1 = _server_case(srv, 7, ta, b1, "f32")
        assert r["status"] == 0 and not r["cache_hit"], r
        # the caller drops its matrix; the freed block is reused for garbage
        del ta
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        junk = [torch.full((n, n), float("nan"), dtype=torch.float64, device="cuda") for _ in range(4)]
        torch.cuda.synchronize()
        for seed in (52, 53):
            b = np.random.default_rng(seed).standard_normal((n, 3))
            r, x = _server_case(srv, 7, None, b, "f32")
            assert r["status"] == 0 and r["cache_hit"], r
            x_ref = orc.solvec(np.array(a), b)
            err = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
            assert err <= 1e-12, (seed, err, r)
            assert r["residual"] <= 10 * np.sqrt(n) * np.finfo(np.float64).eps
        del junk
    finally:
        srv.close()

"""Solve server (llh_server_*): continuous batching with a resident factor
cache.  Each result is the same alone and in the mix; results match the
oracle; failures and budget rejections stay with their own requests."""
import numpy as np
import pytest

import linfa_linalg_hip as llh  # noqa: F401
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def spd(n, seed):
    rng = np.random.default_rng(seed)
    b = rng.uniform(-1, 1, (n, n))
    return b @ b.T / n + np.eye(n)


def make_requests(seed=0):
    rng = np.random.default_rng(seed)
    sizes = [16, 40, 100, 128, 129, 200, 300, 450, 600, 1000]
    mats = {i: spd(n, 100 + i) for i, n in enumerate(sizes)}
    reqs = []
    for k in range(40):
        if k >= len(sizes) and rng.random() < 0.4:
            mid = int(rng.integers(0, len(sizes)))  # reuse a matrix id seen before
        else:
            mid = k % len(sizes)
        n = mats[mid].shape[0]
        m = int(rng.integers(1, 9))
        prec = "f32" if (mid % 2) else "f64"
        reqs.append((mid, n, m, prec, rng.standard_normal((n, m))))
    return mats, reqs


def run(server, mats, reqs, alone):
    outs, res, tickets = [], [], []
    for mid, n, m, prec, b in reqs:
        x = np.zeros((n, m))
        t = server.submit(mid, mats[mid], b, x, precision=prec)
        outs.append(x)
        if alone:
            res.append(server.wait(t))
        else:
            tickets.append(t)
    if not alone:
        res = [server.wait(t) for t in tickets]
    return outs, res


def test_alone_equals_mix_and_matches_oracle():
    from linfa_linalg_hip import device as dev
    mats, reqs = make_requests()
    s1 = dev.Server()
    xa, ra = run(s1, mats, reqs, alone=True)
    s1.close()
    s2 = dev.Server()
    xm, rm = run(s2, mats, reqs, alone=False)
    stats = s2.stats()
    s2.close()
    assert all(r["status"] == 0 for r in ra + rm)
    assert stats["cache_hits"] > 0 and stats["batched"] > 0
    for (mid, n, m, prec, b), x1, x2, r1, r2 in zip(reqs, xa, xm, ra, rm):
        assert np.array_equal(x1, x2), (mid, n, prec)
        x_ref = orc.solvec(np.array(mats[mid]), b)
        err = np.linalg.norm(x1 - x_ref) / np.linalg.norm(x_ref)
        assert err <= 1e-12, (mid, n, prec, err, r1)
        if r1["precision_used"] == "f32" and not r1["batched"]:
            assert r1["residual"] <= 10 * np.sqrt(n) * np.finfo(np.float64).eps


def test_reuse_device_matrix_and_cache_hit():
    import torch
    from linfa_linalg_hip import device as dev
    n = 700
    a = spd(n, 9)
    ta = torch.from_numpy(a).cuda()
    s = dev.Server(streams=2)
    hits = []
    for k in range(3):
        b = np.random.default_rng(k).standard_normal((n, 4))
        x = np.zeros((n, 4))
        r = s.wait(s.submit(77, ta if k == 0 else None, b, x, precision="f32"))
        hits.append(r["cache_hit"])
        assert r["status"] == 0
        assert np.linalg.norm(x - orc.solvec(np.array(a), b)) / np.linalg.norm(x) <= 1e-12
    assert hits == [False, True, True]
    s.close()


def test_failure_and_rejection_are_isolated():
    from linfa_linalg_hip import device as dev
    good = spd(300, 1)
    bad = spd(300, 2)
    bad[123, 123] = -1.0
    small_bad = spd(50, 3)
    small_bad[7, 7] = -2.0
    s = dev.Server(budget_bytes=64 << 20)
    xs = [np.zeros((300, 2)), np.zeros((300, 2)), np.zeros((50, 1))]
    bs = [np.ones((300, 2)), np.ones((300, 2)), np.ones((50, 1))]
    ts = [s.submit(1, good, bs[0], xs[0]), s.submit(2, bad, bs[1], xs[1]),
          s.submit(3, small_bad, bs[2], xs[2])]
    r = [s.wait(t) for t in ts]
    assert r[0]["status"] == 0 and r[1]["failed_row"] == 123 and r[2]["failed_row"] == 7
    assert np.abs(xs[0] - orc.solvec(np.array(good), bs[0])).max() <= 1e-12
    with pytest.raises(llh.LinalgError):
        big = np.zeros((3000, 3000))
        s.submit(4, big, np.ones((3000, 1)), np.zeros((3000, 1)))  # 72 MB+ > 64 MB budget
    s.close()


def test_small_burst_with_repeated_ids():
    """Hundreds of small requests submitted at once, ids repeating inside the
    burst, host and device matrices mixed: hits on a system that the pending
    batched step factors wait for that step, and every step's descriptors stay
    intact until its upload has run."""
    import torch
    from linfa_linalg_hip import device as dev
    rng = np.random.default_rng(5)
    sizes = [8, 33, 64, 100, 128]
    mats = {i: spd(sizes[i % len(sizes)], 300 + i) for i in range(20)}
    dmats = {i: torch.from_numpy(a).cuda() for i, a in mats.items() if i % 3 == 0}
    s = dev.Server(streams=4)
    reqs = []
    for k in range(400):
        mid = int(rng.integers(0, 20))
        n = mats[mid].shape[0]
        m = int(rng.integers(1, 17))
        b = rng.standard_normal((n, m))
        x = np.zeros((n, m))
        src = dmats.get(mid, mats[mid])
        reqs.append((mid, b, x, s.submit(mid, src, b, x)))
    res = [s.wait(t) for (_, _, _, t) in reqs]
    stats = s.stats()
    s.close()
    assert stats["cache_hits"] > 0 and stats["batched"] > 0
    for (mid, b, x, _), r in zip(reqs, res):
        assert r["status"] == 0, (mid, r)
        x_ref = orc.solvec(np.array(mats[mid]), b)
        err = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
        assert err <= 1e-12, (mid, err, r)
        if r["batched"]:
            assert np.array_equal(x, x_ref), mid  # exact-order path for new small systems

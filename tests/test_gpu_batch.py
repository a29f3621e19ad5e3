"""Batched solves of uneven SPD systems (llh_batch_*) against the oracle.

Systems taken by the LDS path (n <= 128 f64 / 176 f32) are bit-identical to
the oracle's solvec; the blocked grouped path is compared within backward-error
bounds.  A failing system reports its first failing row and leaves the others
untouched."""
import numpy as np
import pytest

import linfa_linalg_hip as llh  # noqa: F401
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def make_batch(sizes, rhs, dtype, seed):
    rng = np.random.default_rng(seed)
    mats, bs = [], []
    for n, m in zip(sizes, rhs):
        b = rng.uniform(-1, 1, (n, n))
        mats.append((b @ b.T / max(n, 1) + np.eye(n)).astype(dtype))
        bs.append(rng.standard_normal((n, m)).astype(dtype))
    return mats, bs


def pack(plan, mats, bs, dtype):
    a = np.zeros(plan.a_elems, dtype=dtype)
    b = np.zeros(plan.b_elems, dtype=dtype)
    for i, (am, bm) in enumerate(zip(mats, bs)):
        n, ld = am.shape[0], int(plan.lda[i])
        o = int(plan.a_off[i])
        a[o:o + n * ld].reshape(n, ld)[:, :n] = am
        ob = int(plan.b_off[i])
        b[ob:ob + bm.size] = bm.ravel()
    return a, b


def unpack(plan, a, b, mats, bs):
    fac, xs = [], []
    for i, (am, bm) in enumerate(zip(mats, bs)):
        n, ld = am.shape[0], int(plan.lda[i])
        o = int(plan.a_off[i])
        fac.append(a[o:o + n * ld].reshape(n, ld)[:, :n].copy())
        ob = int(plan.b_off[i])
        xs.append(b[ob:ob + bm.size].reshape(bm.shape).copy())
    return fac, xs


def run(plan, mats, bs, dtype):
    import torch
    a, b = pack(plan, mats, bs, dtype)
    ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    st = torch.zeros(len(mats), dtype=torch.int32, device="cuda")
    plan.solve(ta, tb, st)
    return unpack(plan, ta.cpu().numpy(), tb.cpu().numpy(), mats, bs) + (st.cpu().numpy(),)


SIZES = [1, 2, 5, 17, 32, 33, 64, 65, 100, 128, 129, 176, 177, 200, 256, 300, 333, 512, 513, 700]


@pytest.mark.parametrize("dtype,small_limit", [(np.float64, 128), (np.float64, 32), (np.float32, 176)])
def test_batch_against_oracle(dtype, small_limit):
    from linfa_linalg_hip import device as dev
    rng = np.random.default_rng(1)
    sizes = SIZES + list(rng.integers(1, 400, 40))
    rhs = [int(x) for x in rng.integers(0, 9, len(sizes))]
    mats, bs = make_batch(sizes, rhs, dtype, 7)
    plan = dev.BatchPlan(sizes, rhs, prec="f64" if dtype == np.float64 else "f32", small_limit=small_limit)
    fac, xs, st = run(plan, mats, bs, dtype)
    assert np.all(st == -1)
    eps = np.finfo(dtype).eps
    for am, bm, f, x in zip(mats, bs, fac, xs):
        n = am.shape[0]
        fr, xr = np.array(am), np.array(bm)
        x_ref = orc.solvec_inplace(fr, xr)
        if n <= small_limit:
            assert np.array_equal(f, fr) and np.array_equal(x, x_ref), n
        else:
            l = np.tril(f).astype(np.float64)
            bwd = np.linalg.norm(np.tril(l @ l.T - am)) / np.linalg.norm(np.tril(am))
            assert bwd <= 20 * eps, (n, bwd)
            assert np.array_equal(np.triu(f, 1), np.triu(am, 1))  # dirty: upper untouched
            if bm.size:
                assert np.abs(x - x_ref).max() <= 1e3 * eps * max(1.0, np.abs(x_ref).max()), n


@pytest.mark.parametrize("medium_mode", [0, 2])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batch_lds_substitution_against_oracle(dtype, medium_mode):
    """Every blocked system within 512 rows: the whole forward and back
    substitution of each runs in one workgroup with its right-hand sides in
    LDS (8 columns at a time: m = 9..20 take several chunks).  With the grouped
    steps (medium_mode 2) the forward sweep rides on the factor's diagonal and
    panel launches and the substitution launch runs only the back sweep."""
    from linfa_linalg_hip import device as dev
    rng = np.random.default_rng(11)
    sizes = [129, 130, 191, 192, 193, 255, 256, 257, 400, 447, 448, 449, 511, 512] + \
        list(rng.integers(129, 513, 30))
    rhs = [0, 1, 2, 3, 5, 7, 8, 9, 12, 16, 17, 20, 1, 8] + [int(x) for x in rng.integers(1, 9, 30)]
    mats, bs = make_batch(sizes, rhs, dtype, 17)
    plan = dev.BatchPlan(sizes, rhs, prec="f64" if dtype == np.float64 else "f32", small_limit=128,
                         medium_mode=medium_mode)
    fac, xs, st = run(plan, mats, bs, dtype)
    assert np.all(st == -1)
    eps = np.finfo(dtype).eps
    for am, bm, f, x in zip(mats, bs, fac, xs):
        n = am.shape[0]
        x_ref = orc.solvec_inplace(np.array(am), np.array(bm))
        if bm.size:
            assert np.abs(x - x_ref).max() <= 1e3 * eps * max(1.0, np.abs(x_ref).max()), (n, bm.shape)


def test_batch_failure_is_isolated_and_replay():
    from linfa_linalg_hip import device as dev
    sizes = [50, 300, 90, 400, 20]
    rhs = [2, 3, 1, 4, 0]
    mats, bs = make_batch(sizes, rhs, np.float64, 3)
    mats[1][177, 177] = -1.0   # blocked path fails at row 177
    mats[2][40, 40] = -1.0     # LDS path fails at row 40
    plan = dev.BatchPlan(sizes, rhs)
    for _ in range(2):  # replaying the captured graph gives the same result
        fac, xs, st = run(plan, mats, bs, np.float64)
        assert list(st) == [-1, 177, 40, -1, -1]
        for i in (0, 3):
            x_ref = orc.solvec(np.array(mats[i]), bs[i])
            assert np.abs(xs[i] - x_ref).max() <= 1e-12 * max(1, np.abs(x_ref).max())


def test_batch_autotune_keeps_inputs():
    import torch
    from linfa_linalg_hip import device as dev
    rng = np.random.default_rng(4)
    sizes = [int(x) for x in np.exp(rng.uniform(np.log(8), np.log(512), 300))]
    rhs = [int(x) for x in rng.integers(1, 9, len(sizes))]
    mats, bs = make_batch(sizes, rhs, np.float64, 5)
    plan = dev.BatchPlan(sizes, rhs)
    a, b = pack(plan, mats, bs, np.float64)
    ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    st = torch.zeros(len(sizes), dtype=torch.int32, device="cuda")
    best = plan.autotune(ta, tb, st)
    assert best > 0
    assert np.array_equal(ta.cpu().numpy(), a) and np.array_equal(tb.cpu().numpy(), b)
    fac, xs, st2 = run(plan, mats, bs, np.float64)
    assert np.all(st2 == -1)
    for am, bm, x in zip(mats, bs, xs):
        x_ref = orc.solvec(np.array(am), bm)
        assert np.abs(x - x_ref).max() <= 1e-11 * max(1, np.abs(x_ref).max())


def test_batch_f32_refined_against_oracle():
    """LLH_PREC_F32_REFINED: f64 buffers, f32 factors, f64 refinement, one graph
    launch.  The f64 solution matches the oracle's f64 solvec to 1e-12; A is
    left unchanged; a replay gives the same result."""
    import torch
    from linfa_linalg_hip import device as dev
    rng = np.random.default_rng(2)
    sizes = SIZES + list(rng.integers(1, 400, 40))
    rhs = [int(x) for x in rng.integers(0, 9, len(sizes))]
    mats, bs = make_batch(sizes, rhs, np.float64, 9)
    plan = dev.BatchPlan(sizes, rhs, prec="f32ir", ir_iterations=2)
    a, b = pack(plan, mats, bs, np.float64)
    ta = torch.from_numpy(a).cuda()
    outs = []
    for _ in range(2):
        tb = torch.from_numpy(b).cuda()
        st = torch.zeros(len(mats), dtype=torch.int32, device="cuda")
        plan.solve(ta, tb, st)
        assert np.all(st.cpu().numpy() == -1)
        outs.append(tb.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(ta.cpu().numpy(), a)  # A is not overwritten in this mode
    _, xs = unpack(plan, a, outs[0], mats, bs)
    for am, bm, x in zip(mats, bs, xs):
        if bm.size == 0:
            continue
        x_ref = orc.solvec(np.array(am), bm)
        assert np.abs(x - x_ref).max() <= 1e-12 * max(1.0, np.abs(x_ref).max()), am.shape[0]


def test_batch_f32_refined_failure_is_isolated():
    import torch
    from linfa_linalg_hip import device as dev
    sizes = [50, 300, 90, 400, 20]
    rhs = [2, 3, 1, 4, 1]
    mats, bs = make_batch(sizes, rhs, np.float64, 3)
    mats[1][177, 177] = -1.0
    mats[2][40, 40] = -1.0
    plan = dev.BatchPlan(sizes, rhs, prec="f32ir")
    a, b = pack(plan, mats, bs, np.float64)
    ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    st = torch.zeros(len(sizes), dtype=torch.int32, device="cuda")
    plan.solve(ta, tb, st)
    assert list(st.cpu().numpy()) == [-1, 177, 40, -1, -1]
    _, xs = unpack(plan, a, tb.cpu().numpy(), mats, bs)
    for i in (1, 2):  # a failed factor leaves B as the caller passed it, bit for bit
        assert np.array_equal(xs[i], bs[i]), i
    for i in (0, 3, 4):
        x_ref = orc.solvec(np.array(mats[i]), bs[i])
        assert np.abs(xs[i] - x_ref).max() <= 1e-12 * max(1, np.abs(x_ref).max())


@pytest.mark.parametrize("prec", ["f64", "f32", "f32ir"])
def test_batch_medium_forms_bitwise_equal(prec):
    """The medium bucket's three forms (medium_mode 0: one workgroup per system
    through its factor and substitutions; 1: the factor so, then one
    substitution launch; 2: grouped launches, one per blocked step) run the same
    device code in the same order per system, so factors, solutions and status
    agree bit for bit -- with a failing system among them, and on a replay."""
    import torch
    from linfa_linalg_hip import device as dev
    rng = np.random.default_rng(21)
    sizes = [33, 64, 65, 127, 128, 129, 200, 256, 257, 300, 383, 448, 511, 512, 7, 20] + \
        [int(x) for x in rng.integers(33, 513, 40)]
    rhs = [0, 1, 8, 9, 16, 17, 3, 2, 1, 5, 8, 4, 20, 1, 2, 3] + [int(x) for x in rng.integers(0, 9, 40)]
    dtype = np.float32 if prec == "f32" else np.float64
    mats, bs = make_batch(sizes, rhs, dtype, 23)
    mats[9][211, 211] = -1.0  # n = 300 fails at row 211 (third block)
    outs = []
    for mode in (0, 1, 2):
        plan = dev.BatchPlan(sizes, rhs, prec=prec, small_limit=32, medium_mode=mode)
        assert plan.config["medium_mode"] == mode
        a, b = pack(plan, mats, bs, dtype)
        ta = torch.from_numpy(a).cuda()
        res = []
        for _ in range(2):
            ta.copy_(torch.from_numpy(a))
            tb = torch.from_numpy(b).cuda()
            st = torch.zeros(len(sizes), dtype=torch.int32, device="cuda")
            plan.solve(ta, tb, st)
            res.append((ta.cpu().numpy(), tb.cpu().numpy(), st.cpu().numpy()))
        assert all(np.array_equal(x, y) for x, y in zip(res[0], res[1])), mode  # replay
        outs.append(res[0])
        plan.close()
    expect = [-1] * len(sizes)
    expect[9] = 211
    for mode, (a, b, st) in enumerate(outs):
        assert list(st) == expect, mode
        assert np.array_equal(a, outs[2][0]) and np.array_equal(b, outs[2][1]), mode
    # and the solutions are the oracle's to the blocked path's accuracy
    _, xs = unpack(plan, outs[0][0], outs[0][1], mats, bs)
    eps = np.finfo(dtype).eps if prec != "f32ir" else 1e-15
    for i, (am, bm, x) in enumerate(zip(mats, bs, xs)):
        if i == 9 or bm.size == 0:
            continue
        x_ref = orc.solvec(np.array(am, dtype=np.float64), np.array(bm, dtype=np.float64))
        assert np.abs(x - x_ref).max() <= 1e3 * eps * max(1.0, np.abs(x_ref).max()), (i, sizes[i])

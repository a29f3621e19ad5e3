"""Allocation-free device solvec (llh_dev_solvec_noalloc_d), called through the
C ABI and checked against llh_dev_solvec_d on the same inputs: the same kernels
in the same order, so the factor and the solution must be bit-identical.  Calls
run back to back with shrinking sizes (the static scratch keeps the previous
call's contents), through both solve paths (one launch for m <= 64, blocked
launches above), and the bound on n is enforced before anything runs."""
import ctypes

import numpy as np
import pytest

import linfa_linalg_hip as llh
from linfa_linalg_hip import _abi
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def spd(n, seed):
    rng = np.random.default_rng(seed)
    b = rng.uniform(-1, 1, (n, n))
    return b @ b.T / n + np.eye(n)


def tt(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def call(name, a, b):
    from linfa_linalg_hip import device as dev
    c = dev.context()
    err = _abi.LlhError()
    st = getattr(c._lib, name)(c.handle, a.shape[0], ctypes.c_void_p(a.data_ptr()), a.stride(0), b.shape[1],
                               ctypes.c_void_p(b.data_ptr()), b.stride(0), ctypes.byref(err))
    return st, err


def test_noalloc_equals_dev_solvec_on_consecutive_calls():
    import torch
    for n, m, seed in ((3000, 16, 1), (1500, 80, 2), (700, 3, 3), (100, 5, 4)):
        a = spd(n, seed)
        b = np.random.default_rng(seed + 100).standard_normal((n, m))
        ta1, tb1 = tt(a), tt(b)
        st1, _ = call("llh_dev_solvec_d", ta1, tb1)
        ta2, tb2 = tt(a), tt(b)
        st2, e2 = call("llh_dev_solvec_noalloc_d", ta2, tb2)
        torch.cuda.synchronize()
        assert st1 == 0 and st2 == 0, (n, st1, st2, e2.message)
        assert torch.equal(ta1, ta2), n
        assert torch.equal(tb1, tb2), n
        x_ref = orc.solvec(np.array(a), b)
        x = tb2.cpu().numpy()
        assert np.linalg.norm(x - x_ref) <= 1e-12 * np.linalg.norm(x_ref), n


def test_noalloc_not_pd_and_bound():
    import torch
    n = 900
    a = spd(n, 7)
    a[400, 400] = -1.0
    b = np.random.default_rng(8).standard_normal((n, 2))
    tb = tt(b)
    st, err = call("llh_dev_solvec_noalloc_d", tt(a), tb)
    assert st == 2 and err.column <= 400, (st, err.column)
    assert np.array_equal(tb.cpu().numpy(), b)
    # above the static scratch: rejected before any work (the pointers are never read)
    big = llh.LLH_NOALLOC_MAX_N + 1 if hasattr(llh, "LLH_NOALLOC_MAX_N") else 16385
    small = torch.zeros((1, 1), dtype=torch.float64, device="cuda")
    from linfa_linalg_hip import device as dev
    c = dev.context()
    e = _abi.LlhError()
    st = c._lib.llh_dev_solvec_noalloc_d(c.handle, big, ctypes.c_void_p(small.data_ptr()), big, 1,
                                         ctypes.c_void_p(small.data_ptr()), 1, ctypes.byref(e))
    assert st == 16 and b"exceeds" in e.message, (st, e.message)

"""The multi-threaded oracle and the seeded workload generator (test
infrastructure, oracle/llh_oracle_par.c), checked on the CPU:

- the parallel solvec equals the serial oracle (itself pinned by the
  reference's known answers, tests/test_oracle.py) bit for bit, factor and
  solution, and fails at the same row;
- the generator's SplitMix64 matches the published first outputs of the
  generator seeded with 0; its logarithm is accurate to ~1 ulp; its normal and
  uniform draws have the right moments and distribution;
- the SPD matrix is exactly its definition (sequential sums in k order) on a
  brute-force case, and the committed full-size fixtures agree with the
  generator on their sampled entries and with their recorded checksums.
"""
import hashlib
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as orc
from oracle import par

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "full_size")


def test_splitmix64_published_outputs():
    # SplitMix64 (Steele, Lea, Flood 2014; Vigna's reference code) from state 0
    assert par.splitmix64(0, 0) == 0xE220A8397B1DCDAF
    assert par.splitmix64(0, 1) == 0x6E789E6AA1B965F4
    assert par.splitmix64(0, 2) == 0x06C45D188009454F


def test_log_accuracy():
    rng = np.random.default_rng(0)
    xs = np.concatenate([np.exp(rng.uniform(-700, 0, 20000)), rng.uniform(0.5, 1.0, 20000),
                         1.0 - rng.uniform(0, 1e-6, 2000)])
    worst = 0.0
    for x in xs:
        ref = math.log(x)
        if ref != 0.0:
            worst = max(worst, abs(par.log(x) - ref) / abs(ref))
    assert worst <= 5e-16, worst
    assert par.log(0.5) == -0.6931471805599453 and par.log(0.25) == -2 * 0.6931471805599453


def test_normal_and_uniform_draws():
    from scipy import stats
    z = par.gen_matrix(par.NORMAL, 12345, 3, 400, 500).ravel()
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1.0) < 0.01
    d = stats.kstest(z, "norm").statistic
    assert d < 1.63 / math.sqrt(z.size), d  # 1% level
    u = par.gen_matrix(par.UNIFORM, 12345, 3, 400, 500).ravel()
    assert u.min() >= -1.0 and u.max() < 1.0
    assert abs(u.mean()) < 0.01 and abs(u.var() - 1.0 / 3.0) < 0.01
    assert stats.kstest(u, "uniform", args=(-1, 2)).statistic < 1.63 / math.sqrt(u.size)
    # another stream or seed is another draw; element (i, j) is value(i * cols + j)
    assert not np.array_equal(par.gen_matrix(par.NORMAL, 12345, 4, 3, 3), par.gen_matrix(par.NORMAL, 12345, 3, 3, 3))
    rows = par.gen_rows(par.NORMAL, 12345, 3, 500, 17, 5)
    assert np.array_equal(rows, z.reshape(400, 500)[17:22])


def test_spd_is_its_definition():
    n, lam, seed = 7, 0.25, 99
    b = par.gen_matrix(par.NORMAL, seed, 0, n, n)
    want = np.empty((n, n))
    for i in range(n):
        for j in range(n):
            s = 0.0
            for k in range(n):  # k in order, each product and sum rounded on its own
                s = s + b[i, k] * b[j, k]
            want[i, j] = s / n + (lam if i == j else 0.0)
    assert np.array_equal(par.gen_spd(par.NORMAL, seed, n, lam), want)
    a = par.gen_spd(par.UNIFORM, 5, 300, 1.0)
    assert np.array_equal(a, a.T)
    assert np.linalg.eigvalsh(a).min() > 0.9


@pytest.mark.parametrize("n", [1, 2, 7, 63, 64, 65, 129, 200, 333])
@pytest.mark.parametrize("m", [1, 3, 16])
def test_parallel_solvec_equals_serial(n, m):
    a = par.gen_spd(par.NORMAL, 1000 + n, n, 0.5)
    a[np.triu_indices(n, 1)] = np.nan  # the upper triangle is never read
    b = par.gen_matrix(par.NORMAL, 1000 + n, 1, n, m)
    a1, a2 = np.array(a), np.array(a)
    x1 = orc.solvec(a1, b)
    x2 = par.solvec(a2, b)
    assert np.array_equal(x1, x2)
    assert np.array_equal(a1, a2, equal_nan=True)


def test_parallel_factor_fails_at_the_same_row():
    n = 300
    a = par.gen_spd(par.NORMAL, 7, n, 0.5)
    a[211, 211] = -1.0
    with pytest.raises(orc.OracleError) as e1:
        orc.cholesky_inplace_dirty(np.array(a))
    with pytest.raises(par.ParOracleError) as e2:
        par.cholesky_dirty_inplace(np.array(a))
    assert e1.value.column == e2.value.row == 211


def test_full_size_fixtures_match_the_generator():
    with open(os.path.join(FIX, "meta.json")) as f:
        meta = json.load(f)
    assert "c2_lam0.1" in meta
    for key, cfg in meta.items():
        x = np.load(os.path.join(FIX, cfg["x_ref"]))
        assert hashlib.sha256(x.tobytes()).hexdigest() == cfg["x_ref_sha256"], key
        assert x.shape == (cfg["n"], cfg["columns_kept"]), key
        n = cfg["n"]
        # a few sampled entries of A, recomputed from two generated rows of B each
        for i, j, hx in cfg["a_samples"][:6]:
            bi = par.gen_rows(par.NORMAL, cfg["seed"], 0, n, i, 1)[0]
            bj = par.gen_rows(par.NORMAL, cfg["seed"], 0, n, j, 1)[0]
            s = 0.0
            for k in range(n):
                s = s + bi[k] * bj[k]
            v = s / n
            if i == j:
                v = v + cfg["lam"]
            assert v == float.fromhex(hx), (key, i, j)

#!/usr/bin/env python3
"""Writes the full-size parity fixtures: the multi-threaded oracle's solution
(bit-identical to the reference's operation order, cholesky.rs:136-144 with
triangular.rs:95-144) of the seeded C2 and C3 workloads, plus sampled entries
of each input so the device generator can be checked at full size.

    python tests/golden/make_full_size_fixtures.py c2      # n = 16384, 64 RHS (~3 min, 8 threads)
    python tests/golden/make_full_size_fixtures.py c3      # n = 32768, 16 RHS, 4 lambdas (~40 min)

Workloads (llh_dev_gen_spd_d / oracle.par.gen_spd): A = B B^T / n + lam I with B
n x n ~ N(0, 1) drawn with (SEED, stream 0); right-hand sides N(0, 1) drawn
with (SEED, stream 1), n x m.  C3 keeps the first 4 of its 16 solution columns
(each column's solve is independent of the others).
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import par  # noqa: E402

SEED = 20241202
OUT = os.path.join(HERE, "full_size")
C2 = {"n": 16384, "m": 64, "keep": 64, "lams": [0.1]}
C3 = {"n": 32768, "m": 16, "keep": 4, "lams": [1.0, 1e-2, 1e-4, 1e-6]}


def samples(a, n, count=256, seed=0):
    rng = np.random.default_rng(seed)
    i = rng.integers(0, n, count)
    j = rng.integers(0, n, count)
    idx = list(zip(i.tolist(), j.tolist())) + [(0, 0), (n - 1, n - 1), (n - 1, 0), (0, n - 1)]
    return [[int(r), int(c), float(a[r, c]).hex()] for r, c in idx]


def run(name, cfg):
    os.makedirs(OUT, exist_ok=True)
    meta_path = os.path.join(OUT, "meta.json")
    meta = json.load(open(meta_path)) if os.path.exists(meta_path) else {}
    n, m = cfg["n"], cfg["m"]
    t0 = time.time()
    s = par.gen_spd(par.NORMAL, SEED, n, 0.0)  # B B^T / n; lam is one more rounded add on the diagonal
    print(f"{name}: generated B B^T / n in {time.time() - t0:.1f} s", flush=True)
    rhs = par.gen_matrix(par.NORMAL, SEED, 1, n, m)
    for lam in cfg["lams"]:
        a = s.copy()
        d = np.arange(n)
        a[d, d] = a[d, d] + lam
        samp = samples(a, n)
        t0 = time.time()
        x = par.solvec(a, rhs)[:, :cfg["keep"]].copy()
        print(f"{name} lam={lam:g}: solved in {time.time() - t0:.1f} s", flush=True)
        fname = f"{name}_lam{lam:g}_x_ref.npy"
        np.save(os.path.join(OUT, fname), x)
        meta[f"{name}_lam{lam:g}"] = {
            "n": n, "m": m, "columns_kept": cfg["keep"], "lam": lam, "seed": SEED, "dist": "normal",
            "rhs_stream": 1, "x_ref": fname, "x_ref_sha256": hashlib.sha256(x.tobytes()).hexdigest(),
            "x_ref_norm": float(np.linalg.norm(x)), "a_samples": samp,
            "oracle": "oracle/llh_oracle_par.c orc_par_solvec_d (bit-identical to the serial restatement)"}
        json.dump(meta, open(meta_path, "w"), indent=1)
        del a


if __name__ == "__main__":
    for which in sys.argv[1:] or ["c2"]:
        run(which, {"c2": C2, "c3": C3}[which])

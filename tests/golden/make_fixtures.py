"""Writes the golden fixtures under tests/golden/.

Run here, where the reference checkout is present:

    python tests/golden/make_fixtures.py /root/reference

1. ``reference_kats.json``: the literal known-answer data held in the
   reference's own tests (inputs and expected outputs only), parsed from the
   Rust array literals:
     - cholesky.rs:210-211 (decompose), :220/:230/:240 (bad_matrix),
       :249-254 (solvec), :263 (invc), :270-275 (corner cases);
     - tests/cholesky.rs:89-90 (cholesky_f32);
     - triangular.rs:233-245 (square), :250-257 (solve_triangular 2x2),
       :274 (diag_zero), :281/:289/:292 (solve_error);
     - tests/triangular.rs:51-177 (known_failure, 9x9 upper, 9x10 solution).
2. ``seeded_oracle.npz``: seeded SPD systems (n in {1,2,3,5,8,10,16,64,128},
   f64 and f32, 3 RHS) with the oracle's dirty factor and solution, for GPU
   parity checks that do not need the oracle library at run time.
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def parse_array_literal(text, start_pat):
    """Parse the `array![[...], ...]` literal that follows `start_pat`."""
    i = text.index(start_pat)
    j = text.index("array![", i) + len("array!")
    depth, k = 0, j
    while True:
        c = text[k]
        if c == "[":
            depth += 1
        elif c == "]":
            depth -= 1
            if depth == 0:
                break
        k += 1
    body = text[j:k + 1]
    body = re.sub(r"(\d)f32", r"\1", body)
    body = re.sub(r",\s*\]", "]", body)
    return json.loads(body)


def reference_kats(ref):
    with open(os.path.join(ref, "tests", "triangular.rs")) as f:
        tri_tests = f.read()
    kats = {
        "source": "linfa-linalg 0.2.1 unit and integration tests (data only)",
        "spd3": [[25.0, 15.0, -5.0], [15.0, 18.0, 0.0], [-5.0, 0.0, 11.0]],
        "spd3_lower": [[5.0, 0.0, 0.0], [3.0, 3.0, 0.0], [-1.0, 1.0, 3.0]],
        "solvec_x": [[10.0, -3.0, 2.2, 4.0], [0.0, 2.4, -0.9, 1.1], [5.5, 7.6, 8.1, 10.0]],
        "not_square_2x3": [[1.0, 2.0, 3.0], [3.0, 4.0, 5.0]],
        "non_pd": [[1.0, 2.0], [2.0, 1.0]],
        "zeros2": [[0.0, 0.0], [0.0, 0.0]],
        "square3": [[1, 2, 3], [4, 5, 6], [7, 8, 9]],
        "square3_upper": [[1, 2, 3], [0, 5, 6], [0, 0, 9]],
        "square3_lower": [[1, 0, 0], [4, 5, 0], [7, 8, 9]],
        "tri_col_3x2": [[1, 2], [3, 5], [6, 8]],
        "tri_lower2": [[1.0, 0.0], [3.0, 4.0]],
        "tri_upper2": [[4.4, 2.1], [0.0, 4.3]],
        "tri_expected2x3": [[2.2, 3.1, 2.2], [1.0, 0.0, 5.7]],
        "diag_zero": [[0.0, 3.0], [2.0, 0.0]],
        "solve_error_square": [[1.1, 2.2], [3.3, 2.1]],
        "solve_error_b": [[2.2, 3.3]],
        "solve_error_nonsquare": [[1.2, 3.3]],
        "known_failure_a": parse_array_literal(tri_tests, "fn known_failure"),
        "known_failure_x": parse_array_literal(tri_tests, "let x = array!["),
    }
    # Sanity: the literals above must appear verbatim in the reference tests.
    with open(os.path.join(ref, "src", "cholesky.rs")) as f:
        chol = f.read()
    assert "array![[25., 15., -5.], [15., 18., 0.], [-5., 0., 11.]]" in chol
    assert "[10., -3., 2.2, 4.]" in chol
    with open(os.path.join(ref, "src", "triangular.rs")) as f:
        tri = f.read()
    assert "array![[4.4, 2.1], [0.0, 4.3]]" in tri
    assert "array![[2.2, 3.1, 2.2], [1.0, 0.0, 5.7]]" in tri
    assert len(kats["known_failure_a"]) == 9 and len(kats["known_failure_a"][0]) == 9
    assert len(kats["known_failure_x"]) == 9 and len(kats["known_failure_x"][0]) == 10
    return kats


def spd(n, seed, dtype):
    """A = B^T B + I with B uniform in [-100, 100], as tests/cholesky.rs:9-19."""
    rng = np.random.default_rng(seed)
    b = rng.uniform(-100.0, 100.0, size=(n, n))
    a = b.T @ b
    a[np.diag_indices(n)] += 1.0
    return a.astype(dtype)


def seeded_oracle():
    from oracle import oracle as orc
    out = {}
    for dtype, tag in ((np.float64, "d"), (np.float32, "s")):
        for n in (1, 2, 3, 5, 8, 10, 16, 64, 128):
            a = spd(n, 1000 + n, dtype)
            rng = np.random.default_rng(2000 + n)
            b = rng.standard_normal((n, 3)).astype(dtype)
            fac = np.array(a, copy=True)
            x = orc.solvec_inplace(fac, np.array(b, copy=True))
            out[f"{tag}{n}_a"] = a
            out[f"{tag}{n}_b"] = b
            out[f"{tag}{n}_dirty"] = fac
            out[f"{tag}{n}_x"] = x
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    kats = reference_kats(ref)
    with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
        json.dump(kats, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "seeded_oracle.npz"), **seeded_oracle())
    print("wrote reference_kats.json and seeded_oracle.npz")


if __name__ == "__main__":
    main()

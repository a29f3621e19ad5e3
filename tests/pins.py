"""Known-answer and closed-form checks that pin a Cholesky/triangular
implementation to linfa-linalg 0.2.1's behaviour.

``pin_failures(impl)`` runs every check against any object exposing the
reference's entry points (the CPU oracle, a planted-mistake copy of it, or the
HIP library's host mirror) and returns the names of the checks that failed.
The data come from the reference's own tests (tests/golden/reference_kats.json)
or from constructions whose exact answer the mathematics fixes.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
KATS = json.load(open(os.path.join(HERE, "golden", "reference_kats.json")))
EPS64 = np.finfo(np.float64).eps
UPPER, LOWER = 0, 1


def arr(name, dtype=np.float64):
    return np.array(KATS[name], dtype=dtype)


def seq_matmul(a, b):
    """a @ b with each entry summed left to right, one rounding per op."""
    out = np.zeros((a.shape[0], b.shape[1]), dtype=a.dtype)
    for i in range(a.shape[0]):
        for j in range(b.shape[1]):
            s = a.dtype.type(0)
            for k in range(a.shape[1]):
                s = s + a[i, k] * b[k, j]
            out[i, j] = s
    return out


def integer_factor(n, rng):
    """Integer lower-triangular L0 with positive diagonal; A = L0 L0^T is exact in
    f64 and its Cholesky factor is exactly L0 (every quotient and square root in
    the algorithm is exact), so any dropped term, sign or index error shows."""
    l0 = np.tril(rng.integers(-4, 5, size=(n, n))).astype(np.float64)
    l0[np.diag_indices(n)] = rng.integers(1, 4, size=n)
    return l0, l0 @ l0.T


def _err_code(fn, *args):
    try:
        fn(*args)
    except Exception as e:  # noqa: BLE001 - any implementation's error type
        return getattr(e, "code", "raised"), e
    return 0, None


def pin_failures(impl, status_codes=(1, 2, 3)):
    """Return the names of failed checks (empty list == pinned)."""
    NOT_SQUARE, NOT_PD, WRONG_ROWS = status_codes
    bad = []

    def check(name, cond):
        if not cond:
            bad.append(name)

    # cholesky.rs:208-216 decompose (exact for this matrix)
    try:
        chol = impl.cholesky(arr("spd3"))
        check("decompose", np.array_equal(chol, arr("spd3_lower")))
        dirty = impl.cholesky_dirty(arr("spd3"))
        exp = arr("spd3_lower") + np.triu(arr("spd3"), 1)
        check("decompose_dirty_keeps_upper", np.array_equal(dirty, exp))
        chol32 = impl.cholesky(arr("spd3", np.float32))  # tests/cholesky.rs:87-95
        check("decompose_f32", np.array_equal(chol32, arr("spd3_lower", np.float32)))
    except Exception:  # noqa: BLE001
        bad.append("decompose_raised")

    # cholesky.rs:218-245 bad_matrix
    code, e = _err_code(impl.cholesky, arr("not_square_2x3"))
    check("not_square", code == NOT_SQUARE and (e.a, e.b) == (2, 3))
    code, _ = _err_code(impl.solvec, arr("not_square_2x3"), np.zeros((2, 3)))
    check("solvec_not_square", code == NOT_SQUARE)
    code, e = _err_code(impl.cholesky, arr("non_pd"))
    check("non_pd", code == NOT_PD and e.column == 1)
    code, _ = _err_code(impl.solvec, arr("non_pd"), np.zeros((2, 3)))
    check("solvec_non_pd", code == NOT_PD)
    code, e = _err_code(impl.cholesky, arr("zeros2"))
    check("zeros_non_pd", code == NOT_PD and e.column == 0)

    # cholesky.rs:247-259 solvec with 4 RHS, 1e-7
    try:
        a = arr("spd3")
        x = arr("solvec_x")
        b = seq_matmul(a, x)
        out = impl.solvec(np.array(a), b)
        check("solvec_kat", np.abs(out - x).max() <= 1e-7)
        a2 = np.array(a)
        impl.solvec(a2, b)
        check("solvec_leaves_dirty_factor", np.array_equal(a2, impl.cholesky_dirty(a)))
    except Exception:  # noqa: BLE001
        bad.append("solvec_raised")

    # cholesky.rs:261-266 invc: abs_diff_eq with the default epsilon (f64::EPSILON)
    try:
        inv = impl.invc(arr("spd3"))
        check("invc_eps", np.abs(seq_matmul(arr("spd3"), inv) - np.eye(3)).max() <= EPS64)
        check("invc_symmetric_within_eps", np.abs(inv - inv.T).max() <= 2 * EPS64)
    except Exception:  # noqa: BLE001
        bad.append("invc_raised")

    # cholesky.rs:268-277 corner cases
    try:
        check("empty", impl.cholesky(np.zeros((0, 0))).shape == (0, 0))
        check("empty_invc", impl.invc(np.zeros((0, 0))).shape == (0, 0))
        check("one", np.array_equal(impl.cholesky(np.ones((1, 1))), np.ones((1, 1))))
        check("one_invc", np.array_equal(impl.invc(np.ones((1, 1))), np.ones((1, 1))))
    except Exception:  # noqa: BLE001
        bad.append("corner_raised")

    # triangular.rs:231-246 square (integer-valued data)
    sq = arr("square3")
    check("into_upper", np.array_equal(impl.triangular_inplace(np.array(sq), UPPER),
                                       arr("square3_upper")))
    check("into_lower", np.array_equal(impl.triangular_inplace(np.array(sq), LOWER),
                                       arr("square3_lower")))
    check("is_triangular", (not impl.is_triangular(sq, LOWER)) and
          impl.is_triangular(arr("square3_upper"), UPPER) and
          not impl.is_triangular(arr("square3_upper"), LOWER) and
          impl.is_triangular(arr("square3_lower"), LOWER) and
          impl.is_triangular(np.zeros((0, 0)), LOWER) and
          not impl.is_triangular(arr("tri_col_3x2"), UPPER))
    code, e = _err_code(impl.triangular_inplace, arr("tri_col_3x2"), UPPER)
    check("into_triangular_not_square", code == NOT_SQUARE and (e.a, e.b) == (3, 2))

    # triangular.rs:248-262 2x2 solves, 1e-7
    try:
        exp = arr("tri_expected2x3")
        lo, up = arr("tri_lower2"), arr("tri_upper2")
        x = impl.solve_triangular(lo, seq_matmul(lo, exp), LOWER)
        check("solve_lower_2x2", np.abs(x - exp).max() <= 1e-7)
        x = impl.solve_triangular(up, seq_matmul(up, exp), UPPER)
        check("solve_upper_2x2", np.abs(x - exp).max() <= 1e-7)
    except Exception:  # noqa: BLE001
        bad.append("solve_2x2_raised")

    # triangular.rs:264-299 corner cases and errors
    try:
        check("solve_empty", impl.solve_triangular(np.zeros((0, 0)), np.zeros((0, 0)), UPPER).shape
              == (0, 0))
        check("solve_one", np.array_equal(impl.solve_triangular(np.ones((1, 1)), np.ones((1, 1)),
                                                                UPPER), np.ones((1, 1))))
        impl.solve_triangular(arr("diag_zero"), np.zeros((2, 2)), LOWER)
    except Exception:  # noqa: BLE001
        bad.append("solve_corner_raised")
    code, _ = _err_code(impl.solve_triangular, arr("solve_error_nonsquare"),
                        arr("solve_error_nonsquare"), LOWER)
    check("solve_not_square", code == NOT_SQUARE)
    code, e = _err_code(impl.solve_triangular, arr("solve_error_square"), arr("solve_error_b"),
                        UPPER)
    check("solve_wrong_rows", code == WRONG_ROWS and (e.a, e.b) == (2, 1))

    # tests/triangular.rs:49-180 known_failure: 9x9 upper, 10 RHS, 1e-4
    try:
        a, x = arr("known_failure_a"), arr("known_failure_x")
        b = seq_matmul(a, x)
        out = impl.solve_triangular(a, b, UPPER)
        check("known_failure", np.abs(out - x).max() <= 1e-4)
        out = impl.solve_triangular(np.ascontiguousarray(a.T), seq_matmul(a.T, x), LOWER)
        check("known_failure_transposed", np.abs(out - x).max() <= 1e-4)
    except Exception:  # noqa: BLE001
        bad.append("known_failure_raised")

    # solve_triangular_system with diag_fn = |diag[i]| (QR, qr.rs:145-150, 172-177)
    if hasattr(impl, "solve_triangular_system"):
        rng_q = np.random.default_rng(11)
        try:
            a = arr("known_failure_a")
            xk = arr("known_failure_x")
            own = np.ascontiguousarray(np.diag(a))
            check("system_own_diag_is_solve",
                  np.array_equal(impl.solve_triangular_system(a, seq_matmul(a, xk), UPPER, own, False),
                                 impl.solve_triangular(a, seq_matmul(a, xk), UPPER)))
            for n in (1, 3, 6, 9):
                r = np.triu(rng_q.integers(-3, 4, size=(n, n)).astype(np.float64), 1)
                d = rng_q.choice([-1.0, 1.0], size=n)
                xs = rng_q.integers(-5, 6, size=(n, 3)).astype(np.float64)
                eff = r + np.eye(n)  # pivots |d_i| = 1
                poisoned = r + np.diag(np.full(n, np.nan))  # a's diagonal must not be read
                b = eff @ xs
                check(f"system_upper_abs_{n}",
                      np.array_equal(impl.solve_triangular_system(poisoned, np.array(b), UPPER, d, True), xs))
                # QR's solve_tr_into: R^T (a transposed view) as Lower
                bt = eff.T @ xs
                check(f"system_lower_transposed_{n}",
                      np.array_equal(impl.solve_triangular_system(poisoned.T, np.array(bt), LOWER, d, True), xs))
                signed = r + np.diag(d)
                check(f"system_signed_{n}",
                      np.array_equal(impl.solve_triangular_system(poisoned, signed @ xs, UPPER, d, False), xs))
        except Exception:  # noqa: BLE001
            bad.append("system_raised")
        code, e = _err_code(impl.solve_triangular_system, arr("solve_error_square"), arr("solve_error_b"),
                            UPPER, np.ones(2), True)
        check("system_wrong_rows", code == WRONG_ROWS and (e.a, e.b) == (2, 1))
        code, _ = _err_code(impl.solve_triangular_system, arr("solve_error_nonsquare"),
                            arr("solve_error_nonsquare"), LOWER, np.ones(3), True)
        check("system_not_square", code == NOT_SQUARE)

    # Exact-integer constructions (closed form: the factor is L0, the solution X).
    rng = np.random.default_rng(7)
    for n in (2, 3, 5, 8, 10):
        l0, a = integer_factor(n, rng)
        try:
            check(f"integer_cholesky_{n}", np.array_equal(impl.cholesky(a), l0))
            xs = rng.integers(-6, 7, size=(n, 4)).astype(np.float64)
            unit = np.array(l0)
            unit[np.diag_indices(n)] = rng.choice([-1.0, 1.0], size=n)
            check(f"integer_solve_lower_{n}",
                  np.array_equal(impl.solve_triangular(unit, unit @ xs, LOWER), xs))
            ut = np.ascontiguousarray(unit.T)
            check(f"integer_solve_upper_{n}",
                  np.array_equal(impl.solve_triangular(ut, ut @ xs, UPPER), xs))
            check(f"integer_solvec_{n}",
                  np.abs(impl.solvec(np.array(a), a @ xs) - xs).max() <= 1e-9 * n)
        except Exception:  # noqa: BLE001
            bad.append(f"integer_{n}_raised")
    return bad

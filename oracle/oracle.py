"""ctypes loader for the CPU oracle (``oracle/llh_oracle.c``).

TEST INFRASTRUCTURE ONLY.  Tests, the smoke check and the benchmark's
CPU-baseline leg use this module as the checker; the product library never
imports it.  Every method mirrors one reference entry point (cited in the C
source) over numpy arrays of any strides, including negative ones.

``Oracle(mutant=k)`` loads a copy with planted mistake ``k`` (built by the
Makefile); the oracle's tests use those to show their checks have teeth.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

OK, NOT_SQUARE, NOT_POSITIVE_DEFINITE, WRONG_ROWS = 0, 1, 2, 3
UPPER, LOWER = 0, 1
MUTANTS = tuple(range(1, 10))


class OrcErr(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int32), ("a", ctypes.c_int64),
                ("b", ctypes.c_int64), ("column", ctypes.c_int64)]


class OracleError(Exception):
    def __init__(self, err):
        self.code, self.a, self.b, self.column = err.code, err.a, err.b, err.column
        super().__init__(f"oracle status {err.code} ({err.a}, {err.b}, col {err.column})")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _sfx(a):
    if a.dtype == np.float64:
        return "d"
    if a.dtype == np.float32:
        return "s"
    raise TypeError(f"oracle supports float32/float64, not {a.dtype}")


def _view(a):
    """(rows, cols, ptr to element (0,0), row stride, col stride) in elements."""
    assert a.ndim == 2
    isz = a.itemsize
    return (a.shape[0], a.shape[1], a.ctypes.data, a.strides[0] // isz, a.strides[1] // isz)


def _check(st, err):
    if st:
        raise OracleError(err)


class Oracle:
    def __init__(self, mutant=0):
        name = "liborc.so" if mutant == 0 else f"liborc_mut{mutant}.so"
        path = os.path.join(_HERE, "_build", name)
        if not os.path.exists(path):
            build()
        self.lib = ctypes.CDLL(path)
        i64, ptr, cint = ctypes.c_int64, ctypes.c_void_p, ctypes.c_int
        eptr = ctypes.POINTER(OrcErr)
        for sfx in ("d", "s"):
            f = lambda n: getattr(self.lib, f"orc_{n}_{sfx}")  # noqa: E731
            f("cholesky_inplace_dirty").argtypes = [i64, i64, ptr, i64, i64, eptr]
            f("cholesky_inplace").argtypes = [i64, i64, ptr, i64, i64, eptr]
            f("triangular_inplace").argtypes = [i64, i64, ptr, i64, i64, cint, eptr]
            f("is_triangular").argtypes = [i64, i64, ptr, i64, i64, cint]
            f("solve_triangular_inplace").argtypes = [
                i64, i64, ptr, i64, i64, i64, i64, ptr, i64, i64, cint, eptr]
            f("solve_triangular_system").argtypes = [
                i64, i64, ptr, i64, i64, ptr, i64, cint, i64, i64, ptr, i64, i64, cint, eptr]
            f("solvec_inplace").argtypes = [i64, i64, ptr, i64, i64, i64, i64, ptr, i64, i64, eptr]
            f("invc_inplace").argtypes = [i64, i64, ptr, i64, i64, ptr, eptr]

    def _fn(self, name, a):
        return getattr(self.lib, f"orc_{name}_{_sfx(a)}")

    def cholesky_inplace_dirty(self, a):
        e = OrcErr()
        _check(self._fn("cholesky_inplace_dirty", a)(*_view(a), ctypes.byref(e)), e)
        return a

    def cholesky_inplace(self, a):
        e = OrcErr()
        _check(self._fn("cholesky_inplace", a)(*_view(a), ctypes.byref(e)), e)
        return a

    def triangular_inplace(self, a, uplo):
        e = OrcErr()
        _check(self._fn("triangular_inplace", a)(*_view(a), uplo, ctypes.byref(e)), e)
        return a

    def is_triangular(self, a, uplo):
        return bool(self._fn("is_triangular", a)(*_view(a), uplo))

    def solve_triangular_inplace(self, a, b, uplo):
        assert a.dtype == b.dtype
        e = OrcErr()
        _check(self._fn("solve_triangular_inplace", a)(*_view(a), *_view(b), uplo,
                                                        ctypes.byref(e)), e)
        return b

    def solve_triangular_system(self, a, b, uplo, diag, abs_diag=True):
        """solve_triangular_system with diag_fn(i) = |diag[i]| (or diag[i]), in place in b."""
        assert a.dtype == b.dtype == diag.dtype and diag.ndim == 1
        e = OrcErr()
        ds = diag.strides[0] // diag.itemsize
        _check(self._fn("solve_triangular_system", a)(*_view(a), diag.ctypes.data, ds, int(bool(abs_diag)),
                                                       *_view(b), uplo, ctypes.byref(e)), e)
        return b

    def solvec_inplace(self, a, b):
        assert a.dtype == b.dtype
        e = OrcErr()
        _check(self._fn("solvec_inplace", a)(*_view(a), *_view(b), ctypes.byref(e)), e)
        return b

    def invc_inplace(self, a):
        e = OrcErr()
        out = np.zeros((a.shape[0], a.shape[0]), dtype=a.dtype)
        _check(self._fn("invc_inplace", a)(*_view(a), out.ctypes.data, ctypes.byref(e)), e)
        return out

    # Non-mutating forms, built from the in-place ones as the reference does
    # (cholesky.rs:106-114, 160-162, 196-198; triangular.rs:184-186).
    def cholesky(self, a):
        return self.cholesky_inplace(np.array(a, copy=True))

    def cholesky_dirty(self, a):
        return self.cholesky_inplace_dirty(np.array(a, copy=True))

    def solvec(self, a, b):
        """Like SolveC::solvec, this overwrites `a` with the dirty factor."""
        return self.solvec_inplace(a, np.array(b, copy=True))

    def solve_triangular(self, a, b, uplo):
        return self.solve_triangular_inplace(a, np.array(b, copy=True), uplo)

    def invc(self, a):
        return self.invc_inplace(np.array(a, copy=True))


_default = None


def default():
    global _default
    if _default is None:
        _default = Oracle()
    return _default


def __getattr__(name):
    # module-level access: oracle.cholesky(...) etc. use the faithful oracle
    return getattr(default(), name)

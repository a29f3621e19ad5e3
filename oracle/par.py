"""ctypes loader for the multi-threaded oracle and the seeded workload
generator (``oracle/llh_oracle_par.c``).

TEST INFRASTRUCTURE ONLY: the checker for the full-size configurations and the
CPU restatement of the synthetic inputs; the product library never imports it.

- ``solvec(a, b)``: bit-identical to the serial oracle's ``solvec`` (the
  reference's operation order, cholesky.rs:136-144 with triangular.rs:95-144)
  for C-contiguous float64 arrays, on every host core.
- ``gen_matrix`` / ``gen_spd``: the seeded inputs (SplitMix64 counters, polar
  normals), which the product's device generator reproduces bit for bit.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
NORMAL, UNIFORM = 0, 1
_lib = None


class ParOracleError(Exception):
    def __init__(self, status, row):
        self.status, self.row = status, row
        super().__init__(f"oracle status {status} (row {row})")


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "_build", "liborc_par.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        L = ctypes.CDLL(path)
        u64, i64, dbl, ptr, cint = ctypes.c_uint64, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p, ctypes.c_int
        L.orc_splitmix64.argtypes = [u64, u64]
        L.orc_splitmix64.restype = u64
        L.orc_gen_key.argtypes = [u64, u64]
        L.orc_gen_key.restype = u64
        L.orc_log.argtypes = [dbl]
        L.orc_log.restype = dbl
        L.orc_gen_value.argtypes = [cint, u64, u64]
        L.orc_gen_value.restype = dbl
        L.orc_gen_matrix_d.argtypes = [cint, u64, u64, i64, i64, ptr, i64]
        L.orc_gen_matrix_d.restype = None
        L.orc_gen_rows_d.argtypes = [cint, u64, u64, i64, i64, i64, ptr]
        L.orc_gen_rows_d.restype = None
        L.orc_gen_spd_d.argtypes = [cint, u64, i64, dbl, ptr, i64, ptr]
        L.orc_gen_spd_d.restype = None
        L.orc_par_cholesky_dirty_d.argtypes = [i64, ptr, i64, ctypes.POINTER(i64)]
        L.orc_par_solve_lower_upper_d.argtypes = [i64, ptr, i64, i64, ptr, i64]
        L.orc_par_solve_lower_upper_d.restype = None
        L.orc_par_solvec_d.argtypes = [i64, ptr, i64, i64, ptr, i64, ctypes.POINTER(i64)]
        _lib = L
    return _lib


def splitmix64(state, c):
    return lib().orc_splitmix64(state, c)


def log(s):
    return lib().orc_log(float(s))


def gen_matrix(dist, seed, stream, rows, cols):
    x = np.empty((rows, cols), dtype=np.float64)
    lib().orc_gen_matrix_d(dist, seed, stream, rows, cols, x.ctypes.data, cols)
    return x


def gen_rows(dist, seed, stream, cols, row0, nrows):
    """Rows row0 .. row0 + nrows of a draw with `cols` columns."""
    x = np.empty((nrows, cols), dtype=np.float64)
    lib().orc_gen_rows_d(dist, seed, stream, cols, row0, nrows, x.ctypes.data)
    return x


def gen_spd(dist, seed, n, lam):
    """A = B B^T / n + lam I, B n x n drawn with (seed, stream 0)."""
    a = np.empty((n, n), dtype=np.float64)
    w = np.empty((n, n), dtype=np.float64)
    lib().orc_gen_spd_d(dist, seed, n, float(lam), a.ctypes.data, n, w.ctypes.data)
    return a


def _c64(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous and a.flags.writeable
    return a


def cholesky_dirty_inplace(a):
    _c64(a)
    fail = ctypes.c_int64(-1)
    st = lib().orc_par_cholesky_dirty_d(a.shape[0], a.ctypes.data, a.shape[1], ctypes.byref(fail))
    if st:
        raise ParOracleError(st, fail.value)
    return a


def solvec_inplace(a, b):
    """A := dirty factor, b := A^-1 b (b n x m C-contiguous)."""
    _c64(a)
    _c64(b)
    assert b.shape[0] == a.shape[0]
    fail = ctypes.c_int64(-1)
    st = lib().orc_par_solvec_d(a.shape[0], a.ctypes.data, a.shape[1], b.shape[1], b.ctypes.data, b.shape[1],
                                ctypes.byref(fail))
    if st:
        raise ParOracleError(st, fail.value)
    return b


def solvec(a, b):
    """Like SolveC::solvec: overwrites `a` with the dirty factor, returns x."""
    return solvec_inplace(a, np.array(b, dtype=np.float64, order="C", copy=True))

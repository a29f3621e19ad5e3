"""linfa_linalg_hip: linfa-linalg's Cholesky / triangular-solve path on an MI355X.

A host-side mirror of the reference's trait surface (src/cholesky.rs,
src/triangular.rs) over numpy arrays, calling the C ABI of include/llh.h in
``libllh.so`` through ctypes.  Names, argument meaning, in-place/"dirty"
semantics and error behaviour follow the reference:

    CholeskyInplace  cholesky_inplace_dirty, cholesky_into_dirty,
                     cholesky_inplace, cholesky_into          cholesky.rs:16-44
    Cholesky         cholesky_dirty, cholesky                 cholesky.rs:86-115
    SolveCInplace    solvec_inplace, solvec_into              cholesky.rs:118-145
    SolveC           solvec                                   cholesky.rs:148-163
    InverseCInplace  invc_inplace                             cholesky.rs:166-183
    InverseC         invc                                     cholesky.rs:186-199
    IntoTriangular   triangular_inplace, into_triangular      triangular.rs:16-54
    Triangular       is_triangular                            triangular.rs:57-91
    SolveTriangular* solve_triangular_inplace / _into / solve_triangular
                                                              triangular.rs:147-187

Errors are raised as LinalgError subclasses (lib.rs:33-60).  The numpy arrays
may have any strides, negative ones included; the in-place forms write through
them.  All arithmetic runs on the GPU: there is no CPU fallback, and importing
this module without the built library raises.
"""
import ctypes
import enum
import os
import threading

import numpy as np

from . import _abi

__all__ = [
    "UPLO", "LinalgError", "NotSquare", "NotPositiveDefinite", "WrongRows", "HipError",
    "cholesky_inplace_dirty", "cholesky_into_dirty", "cholesky_inplace", "cholesky_into",
    "cholesky_dirty", "cholesky", "solvec_inplace", "solvec_into", "solvec", "invc_inplace",
    "invc", "triangular_inplace", "into_triangular", "is_triangular", "solve_triangular_inplace",
    "solve_triangular_into", "solve_triangular", "solve_triangular_system", "Context", "library_path",
]


class UPLO(enum.IntEnum):
    """triangular.rs:10-13"""
    Upper = 0
    Lower = 1


class LinalgError(Exception):
    """lib.rs:33-60.  ``code`` is the llh_status; ``a``/``b`` carry the
    NotSquare (rows, cols) or WrongRows (expected, actual) payload; ``column`` is
    the failing row of NotPositiveDefinite (-1 otherwise)."""

    def __init__(self, code, a=0, b=0, column=-1, message=""):
        self.code, self.a, self.b, self.column = code, a, b, column
        super().__init__(message or _abi.status_name(code))


class NotSquare(LinalgError):
    @property
    def rows(self):
        return self.a

    @property
    def cols(self):
        return self.b


class NotPositiveDefinite(LinalgError):
    pass


class WrongRows(LinalgError):
    @property
    def expected(self):
        return self.a

    @property
    def actual(self):
        return self.b


class HipError(LinalgError):
    pass


_ERRORS = {1: NotSquare, 2: NotPositiveDefinite, 3: WrongRows}


def _raise(st, err):
    if st == 0:
        return
    cls = _ERRORS.get(st, HipError)
    raise cls(st, err.a, err.b, err.column, err.message.decode(errors="replace"))


def library_path():
    return _abi.LIB_PATH


class Context:
    """One llh_context (one HIP stream) on a device."""

    def __init__(self, device=-1):
        self._lib = _abi.lib()
        self._ptr = ctypes.c_void_p()
        err = _abi.LlhError()
        _raise(self._lib.llh_context_create(device, ctypes.byref(self._ptr), ctypes.byref(err)), err)

    @property
    def handle(self):
        return self._ptr

    def close(self):
        if self._ptr:
            self._lib.llh_context_destroy(self._ptr)
            self._ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


_tls = threading.local()


def default_context():
    ctx = getattr(_tls, "ctx", None)
    if ctx is None:
        ctx = Context()
        _tls.ctx = ctx
    return ctx


def _sfx(a):
    if a.dtype == np.float64:
        return "d"
    if a.dtype == np.float32:
        return "s"
    raise TypeError(f"linfa_linalg_hip supports float32 and float64 (NdFloat), not {a.dtype}")


def _view(a):
    """(rows, cols, pointer to element (0, 0), row stride, col stride)."""
    if a.ndim != 2:
        raise ValueError("expected a 2-D array")
    isz = a.itemsize
    if a.strides[0] % isz or a.strides[1] % isz:
        raise ValueError("strides must be whole elements")
    return (a.shape[0], a.shape[1], ctypes.c_void_p(a.ctypes.data), a.strides[0] // isz,
            a.strides[1] // isz)


def _writable(a):
    if not isinstance(a, np.ndarray):
        raise TypeError("in-place forms need a numpy array")
    if not a.flags.writeable:
        raise ValueError("array is read-only")
    return a


def _call(name, a, *args):
    ctx = default_context()
    err = _abi.LlhError()
    fn = getattr(ctx._lib, f"llh_{name}_{_sfx(a)}")
    _raise(fn(ctx.handle, *args, ctypes.byref(err)), err)


# --- CholeskyInplace (cholesky.rs:16-83) --------------------------------------

def cholesky_inplace_dirty(a):
    _call("cholesky_inplace_dirty", _writable(a), *_view(a))
    return a


def cholesky_into_dirty(a):
    return cholesky_inplace_dirty(a)


def cholesky_inplace(a):
    _call("cholesky_inplace", _writable(a), *_view(a))
    return a


def cholesky_into(a):
    return cholesky_inplace(a)


# --- Cholesky (cholesky.rs:86-115): to_owned, then the in-place form -----------

def cholesky_dirty(a):
    return cholesky_into_dirty(np.array(a, copy=True))


def cholesky(a):
    return cholesky_into(np.array(a, copy=True))


# --- SolveCInplace / SolveC (cholesky.rs:118-163) ------------------------------

def solvec_inplace(a, b):
    """Solves a x = b in place in b; a is left holding the dirty factor."""
    if a.dtype != b.dtype:
        raise TypeError("a and b must have the same dtype")
    _call("solvec_inplace", _writable(a), *_view(a), *_view(_writable(b)))
    return b


def solvec_into(a, b):
    return solvec_inplace(a, b)


def solvec(a, b):
    return solvec_into(a, np.array(b, copy=True))


# --- InverseCInplace / InverseC (cholesky.rs:166-199) --------------------------

def invc_inplace(a):
    n = a.shape[0] if a.ndim == 2 else 0
    out = np.zeros((n, n), dtype=a.dtype)
    rows, cols, ptr, rs, cs = _view(_writable(a))
    isz = out.itemsize
    _call("invc_inplace", a, rows, cols, ptr, rs, cs, ctypes.c_void_p(out.ctypes.data),
          out.strides[0] // isz, out.strides[1] // isz)
    return out


def invc(a):
    return invc_inplace(np.array(a, copy=True))


# --- IntoTriangular / Triangular (triangular.rs:16-91) -------------------------

def triangular_inplace(a, uplo):
    _call("triangular_inplace", _writable(a), *_view(a), int(uplo))
    return a


def into_triangular(a, uplo):
    return triangular_inplace(a, uplo)


def is_triangular(a, uplo):
    a = np.asarray(a)
    out = ctypes.c_int(0)
    _call("is_triangular", a, *_view(a), int(uplo), ctypes.byref(out))
    return bool(out.value)


# --- SolveTriangularInplace / SolveTriangular (triangular.rs:147-187) ----------

def solve_triangular_inplace(a, b, uplo):
    a = np.asarray(a)
    if a.dtype != b.dtype:
        raise TypeError("a and b must have the same dtype")
    _call("solve_triangular_inplace", a, *_view(a), *_view(_writable(b)), int(uplo))
    return b


def solve_triangular_into(a, b, uplo):
    return solve_triangular_inplace(a, b, uplo)


def solve_triangular(a, b, uplo):
    return solve_triangular_into(a, np.array(b, copy=True), uplo)


def solve_triangular_system(a, b, uplo, diag, abs_diag=True):
    """solve_triangular_system(a, &mut b, uplo, diag_fn) (triangular.rs:133-144)
    with diag_fn(i) = |diag[i]| (abs_diag, as the QR solves pass it, qr.rs:145-150)
    or diag[i]; b is solved in place and returned.  a's diagonal is not read."""
    a = np.asarray(a)
    diag = np.asarray(diag)
    if a.dtype != b.dtype or diag.dtype != a.dtype:
        raise TypeError("a, b and diag must have the same dtype")
    if diag.ndim != 1:
        raise ValueError("diag must be 1-D")
    ds = diag.strides[0] // diag.itemsize
    _call("solve_triangular_system", a, *_view(a), ctypes.c_void_p(diag.ctypes.data), ds, int(bool(abs_diag)),
          *_view(_writable(b)), int(uplo))
    return b


from . import device  # noqa: E402,F401  (device-resident forms over torch tensors)

if os.environ.get("LLH_EAGER_CONTEXT") == "1":
    default_context()

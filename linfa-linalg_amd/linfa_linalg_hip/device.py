"""Device-resident forms (llh_dev_* in include/llh.h) over torch CUDA tensors.

torch only provides device memory and the stream here: the context is bound to
torch's current stream, so these calls are ordered with the caller's torch work.
Tensors must be 2-D, row-major contiguous rows (stride(1) == 1) of float64 or
float32.
"""
import ctypes

from . import _abi, Context, _raise

_ctx = {}


def context():
    import torch
    dev = torch.cuda.current_device()
    stream = torch.cuda.current_stream(dev).cuda_stream
    key = (dev, stream)
    if key not in _ctx:
        c = Context(dev)
        err = _abi.LlhError()
        _raise(c._lib.llh_context_set_stream(c.handle, ctypes.c_void_p(stream), ctypes.byref(err)), err)
        _ctx[key] = c
    return _ctx[key]


def _sfx(t):
    import torch
    if t.dtype == torch.float64:
        return "d"
    if t.dtype == torch.float32:
        return "s"
    raise TypeError(f"unsupported dtype {t.dtype}")


def _check(t):
    if not t.is_cuda or t.dim() != 2 or t.stride(1) != 1:
        raise ValueError("expected a 2-D CUDA tensor with unit column stride")


def _call(name, t, *args):
    c = context()
    err = _abi.LlhError()
    _raise(getattr(c._lib, f"llh_dev_{name}_{_sfx(t)}")(c.handle, *args, ctypes.byref(err)), err)


def potrf(a, zero_upper=False):
    """In-place lower Cholesky of a square tensor (dirty unless zero_upper)."""
    _check(a)
    n = a.shape[0]
    if a.shape[1] != n:
        raise ValueError("square matrix expected")
    _call("potrf", a, n, ctypes.c_void_p(a.data_ptr()), a.stride(0), int(bool(zero_upper)))
    return a


def solve_triangular(a, b, uplo, trans=False):
    """Solves op(a) x = b in place in b; uplo names the triangle of op(a)."""
    _check(a)
    _check(b)
    n, m = b.shape
    _call("solve_triangular", a, int(uplo), int(bool(trans)), n, ctypes.c_void_p(a.data_ptr()),
          a.stride(0), m, ctypes.c_void_p(b.data_ptr()), b.stride(0))
    return b


def solvec(a, b):
    """a := dirty factor, b := a^-1 b (cholesky.rs:136-144)."""
    _check(a)
    _check(b)
    n, m = b.shape
    _call("solvec", a, n, ctypes.c_void_p(a.data_ptr()), a.stride(0), m,
          ctypes.c_void_p(b.data_ptr()), b.stride(0))
    return b


def set_profiling(enable):
    c = context()
    err = _abi.LlhError()
    _raise(c._lib.llh_context_set_profiling(c.handle, int(bool(enable)), ctypes.byref(err)), err)


def kernel_stats():
    """{class name: (launches, total_ms, flops, bytes)} of the profiled launches."""
    c = context()
    out = {}
    for name, cls in _abi.KERNEL_CLASSES.items():
        st = _abi.KernelStats()
        err = _abi.LlhError()
        _raise(c._lib.llh_context_kernel_stats(c.handle, cls, ctypes.byref(st), ctypes.byref(err)), err)
        out[name] = (st.launches, st.total_ms, st.flops, st.bytes)
    return out

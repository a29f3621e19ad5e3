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


def set_solve_spin_limit(limit):
    """Bound on the one-launch solve's flag waits for this stream's context
    (negative: every wait gives up at once, to test the error path; 0: the
    default)."""
    c = context()
    err = _abi.LlhError()
    _raise(c._lib.llh_context_set_solve_spin_limit(c.handle, int(limit), ctypes.byref(err)), err)


GEN_NORMAL, GEN_UNIFORM = 0, 1


def gen_matrix(dist, seed, stream, rows, cols):
    """A new rows x cols float64 CUDA tensor holding the seeded draw
    (llh_dev_gen_matrix_d; dist GEN_NORMAL or GEN_UNIFORM)."""
    import torch
    x = torch.empty((rows, cols), dtype=torch.float64, device="cuda")
    c = context()
    err = _abi.LlhError()
    _raise(c._lib.llh_dev_gen_matrix_d(c.handle, dist, seed, stream, rows, cols, ctypes.c_void_p(x.data_ptr()),
                                       cols, ctypes.byref(err)), err)
    return x


def gen_spd(dist, seed, n, lam, out=None):
    """A = B B^T / n + lam I with B the (seed, stream 0) draw (llh_dev_gen_spd_d),
    into `out` (n x n float64 CUDA tensor) or a new tensor."""
    import torch
    a = out if out is not None else torch.empty((n, n), dtype=torch.float64, device="cuda")
    _check(a)
    w = torch.empty((n, n), dtype=torch.float64, device="cuda")
    c = context()
    err = _abi.LlhError()
    _raise(c._lib.llh_dev_gen_spd_d(c.handle, dist, seed, n, float(lam), ctypes.c_void_p(a.data_ptr()), a.stride(0),
                                    ctypes.c_void_p(w.data_ptr()), ctypes.byref(err)), err)
    del w
    return a


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


def invc(a, out):
    """invc_inplace on device tensors: a := its dirty factor, out := a^-1
    (both triangles)."""
    _check(a)
    _check(out)
    n = a.shape[0]
    if a.shape[1] != n or tuple(out.shape) != (n, n):
        raise ValueError("square matrices of one size expected")
    _call("invc", a, n, ctypes.c_void_p(a.data_ptr()), a.stride(0), ctypes.c_void_p(out.data_ptr()), out.stride(0))
    return out


def orthonormalize(v):
    """LOBPCG's orthonormalize (lobpcg/algorithm.rs:75-97): returns (u, r) with
    r = cholesky(V^T V) (lower) and u = V r^-T; u is v, overwritten in place."""
    import torch
    _check(v)
    n, k = v.shape
    r = torch.zeros(k, k, dtype=v.dtype, device=v.device)
    _call("orthonormalize", v, n, k, ctypes.c_void_p(v.data_ptr()), v.stride(0), ctypes.c_void_p(r.data_ptr()),
          r.stride(0))
    return v, r


def apply_constraints(v, cholesky_yy, y):
    """LOBPCG's apply_constraints (lobpcg/algorithm.rs:60-73): v -= y (L_yy^-1 (y^T v))
    in place; cholesky_yy is the lower Cholesky factor of y^T y."""
    for t in (v, cholesky_yy, y):
        _check(t)
    n, k = v.shape
    ky = y.shape[1]
    if y.shape[0] != n or tuple(cholesky_yy.shape) != (ky, ky):
        raise ValueError("shape mismatch")
    _call("apply_constraints", v, n, k, ctypes.c_void_p(v.data_ptr()), v.stride(0), ky,
          ctypes.c_void_p(y.data_ptr()), y.stride(0), ctypes.c_void_p(cholesky_yy.data_ptr()),
          cholesky_yy.stride(0))
    return v


def set_profiling(enable, classes=None):
    """Live kernel timing on or off; classes (names of _abi.KERNEL_CLASSES)
    restricts the recorded launches to those classes."""
    c = context()
    err = _abi.LlhError()
    if enable and classes is not None:
        mask = 0
        for name in classes:
            mask |= 1 << _abi.KERNEL_CLASSES[name]
        _raise(c._lib.llh_context_set_profiling_classes(c.handle, mask, ctypes.byref(err)), err)
        return
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


def _info_dict(info):
    return {"iterations": info.iterations, "converged": bool(info.converged),
            "fell_back": bool(info.fell_back),
            "precision_used": {v: k for k, v in _abi.PREC.items()}[info.precision_used],
            "residual": info.residual, "anorm": info.anorm}


def _opts(max_iter, fallback):
    return _abi.RefineOpts(int(max_iter), _abi.FALLBACK_F64 if fallback else 0)


def solve_mixed(a, b, x, prec="f32", max_iter=30, fallback=True):
    """x = a^-1 b for an f64 SPD a (lower triangle read), factored in `prec`
    with f64 iterative refinement (llh_dev_solve_mixed).  prec "f32": f32
    trailing updates; "f16": f16 trailing updates (f32 accumulation) of the
    diagonally scaled a, for well-conditioned a; "f64": no refinement.  Returns
    the refinement report."""
    for t in (a, b, x):
        _check(t)
    n, m = b.shape
    c = context()
    err = _abi.LlhError()
    info = _abi.RefineInfo()
    opts = _opts(max_iter, fallback)
    st = c._lib.llh_dev_solve_mixed(c.handle, _abi.PREC[prec], n, ctypes.c_void_p(a.data_ptr()), a.stride(0), m,
                                    ctypes.c_void_p(b.data_ptr()), b.stride(0), ctypes.c_void_p(x.data_ptr()),
                                    x.stride(0), ctypes.byref(opts), ctypes.byref(info), ctypes.byref(err))
    _raise(st, err)
    return _info_dict(info)


class Factor:
    """A resident factor of an f64 SPD matrix (llh_factor_*): repeated
    right-hand sides reuse it.  `a` must stay alive and unchanged."""

    def __init__(self, a, prec="f32", fallback=True):
        _check(a)
        self._a = a
        self._c = context()
        self._ptr = ctypes.c_void_p()
        err = _abi.LlhError()
        st = self._c._lib.llh_factor_create(self._c.handle, _abi.PREC[prec], a.shape[0],
                                            ctypes.c_void_p(a.data_ptr()), a.stride(0),
                                            _abi.FALLBACK_F64 if fallback else 0, ctypes.byref(self._ptr),
                                            ctypes.byref(err))
        _raise(st, err)
        self.fallback = fallback

    @property
    def precision(self):
        return {v: k for k, v in _abi.PREC.items()}[self._c._lib.llh_factor_precision(self._ptr)]

    @property
    def device_bytes(self):
        return self._c._lib.llh_factor_device_bytes(self._ptr)

    def solve(self, b, x, max_iter=30):
        _check(b)
        _check(x)
        err = _abi.LlhError()
        info = _abi.RefineInfo()
        opts = _opts(max_iter, self.fallback)
        st = self._c._lib.llh_factor_solve(self._c.handle, self._ptr, b.shape[1], ctypes.c_void_p(b.data_ptr()),
                                           b.stride(0), ctypes.c_void_p(x.data_ptr()), x.stride(0),
                                           ctypes.byref(opts), ctypes.byref(info), ctypes.byref(err))
        _raise(st, err)
        return _info_dict(info)

    def close(self):
        if self._ptr:
            self._c._lib.llh_factor_destroy(self._ptr)
            self._ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


class BatchPlan:
    """A plan for many independent SPD systems of uneven sizes (llh_batch_*).

    Matrix i lives in a packed buffer at a_off[i] (row-major, leading dimension
    lda[i]); its right-hand sides at b_off[i] (row-major n_i x m_i).  solve()
    factors every A_i in place and overwrites every B_i with its solution in one
    hipGraph launch; status[i] is -1 or the failing row.  prec "f32ir": float64
    buffers, each A_i factored in f32 from a copy (A_i left unchanged) and B_i
    refined ir_iterations times in f64, still one launch."""

    def __init__(self, n, m, prec="f64", small_limit=None, small_waves=None, ir_iterations=None,
                 medium_mode=None):
        import numpy as np
        self._c = context()
        self.n = np.ascontiguousarray(n, dtype=np.int32)
        self.m = np.ascontiguousarray(m, dtype=np.int32)
        self.prec = prec
        cfg = None
        if any(v is not None for v in (small_limit, small_waves, ir_iterations, medium_mode)):
            cfg = _abi.BatchConfig(int(small_limit if small_limit is not None else (128 if prec == "f64" else 176)),
                                   int(small_waves if small_waves is not None else 4),
                                   int(ir_iterations or 0), int(medium_mode or 0))
        self._ptr = ctypes.c_void_p()
        err = _abi.LlhError()
        i32p = ctypes.POINTER(ctypes.c_int32)
        st = self._c._lib.llh_batch_plan_create(
            self._c.handle, _abi.PREC[prec], len(self.n), self.n.ctypes.data_as(i32p),
            self.m.ctypes.data_as(i32p), ctypes.byref(cfg) if cfg is not None else None,
            ctypes.byref(self._ptr), ctypes.byref(err))
        _raise(st, err)
        cnt = len(self.n)
        self.a_off = np.zeros(cnt, dtype=np.uint64)
        self.lda = np.zeros(cnt, dtype=np.int32)
        self.b_off = np.zeros(cnt, dtype=np.uint64)
        err = _abi.LlhError()
        _raise(self._c._lib.llh_batch_layout(self._ptr, self.a_off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                             self.lda.ctypes.data_as(i32p),
                                             self.b_off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                             ctypes.byref(err)), err)

    @property
    def a_elems(self):
        return self._c._lib.llh_batch_a_elems(self._ptr)

    @property
    def b_elems(self):
        return self._c._lib.llh_batch_b_elems(self._ptr)

    @property
    def config(self):
        c = self._c._lib.llh_batch_get_config(self._ptr)
        return {"small_limit": c.small_limit, "small_waves": c.small_waves, "ir_iterations": c.ir_iterations,
                "medium_mode": c.medium_mode}

    def solve(self, a, b, status):
        err = _abi.LlhError()
        st = self._c._lib.llh_batch_solve(self._c.handle, self._ptr, ctypes.c_void_p(a.data_ptr()),
                                          ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(status.data_ptr()),
                                          ctypes.byref(err))
        _raise(st, err)

    def autotune(self, a, b, status):
        err = _abi.LlhError()
        best = ctypes.c_double(0.0)
        st = self._c._lib.llh_batch_autotune(self._c.handle, self._ptr, ctypes.c_void_p(a.data_ptr()),
                                             ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(status.data_ptr()),
                                             ctypes.byref(best), ctypes.byref(err))
        _raise(st, err)
        return best.value

    def close(self):
        if self._ptr:
            self._c._lib.llh_batch_plan_destroy(self._ptr)
            self._ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


class Server:
    """Continuous-batching solve server (llh_server_*).

    submit() takes a host (numpy) or device (torch) f64 matrix a (n x n,
    C-contiguous; None when the factor under matrix_id is resident), host b
    (n x m) and host x (n x m, written); the arrays are kept alive until the
    request is collected with wait() or poll()."""

    def __init__(self, budget_bytes=0, streams=4, max_inflight=64, small_limit=128, ir_iterations=3):
        from . import default_context
        self._c = default_context()
        self._ptr = ctypes.c_void_p()
        cfg = _abi.ServerConfig(int(budget_bytes), int(streams), int(max_inflight), int(small_limit),
                                int(ir_iterations))
        err = _abi.LlhError()
        _raise(self._c._lib.llh_server_create(self._c.handle, ctypes.byref(cfg), ctypes.byref(self._ptr),
                                              ctypes.byref(err)), err)
        self._keep = {}

    def submit(self, matrix_id, a, b, x, precision="f64"):
        import numpy as np
        n, m = b.shape
        on_dev = a is not None and not isinstance(a, np.ndarray)
        if a is not None:
            if on_dev:
                _check(a)
                a_ptr = a.data_ptr()
            else:
                assert a.dtype == np.float64 and a.flags.c_contiguous and a.shape == (n, n)
                a_ptr = a.ctypes.data
        else:
            a_ptr = None
        assert b.dtype == np.float64 and b.flags.c_contiguous
        assert x.dtype == np.float64 and x.flags.c_contiguous and x.shape == (n, m) and x.flags.writeable
        req = _abi.Request(int(matrix_id), n, m, _abi.PREC[precision], int(on_dev), a_ptr, b.ctypes.data,
                           x.ctypes.data)
        t = ctypes.c_uint64(0)
        err = _abi.LlhError()
        _raise(self._c._lib.llh_server_submit(self._ptr, ctypes.byref(req), ctypes.byref(t), ctypes.byref(err)), err)
        self._keep[t.value] = (a, b, x)
        return t.value

    @staticmethod
    def _res(r):
        return {"status": r.status, "cache_hit": bool(r.cache_hit), "batched": bool(r.batched),
                "precision_used": {v: k for k, v in _abi.PREC.items()}.get(r.precision_used),
                "failed_row": r.failed_row, "latency_s": r.done_s - r.submit_s,
                "queue_s": r.admit_s - r.submit_s if r.admit_s > 0 else None, "residual": r.residual,
                "done_s": r.done_s, "submit_s": r.submit_s}

    def poll(self, ticket):
        r = _abi.Result()
        if self._c._lib.llh_server_poll(self._ptr, ticket, ctypes.byref(r)):
            self._keep.pop(ticket, None)
            return self._res(r)
        return None

    def wait(self, ticket):
        r = _abi.Result()
        err = _abi.LlhError()
        st = self._c._lib.llh_server_wait(self._ptr, ticket, ctypes.byref(r), ctypes.byref(err))
        self._keep.pop(ticket, None)
        out = self._res(r)
        out["error"] = None if st == 0 else _abi.status_name(st)
        return out

    def stats(self):
        st = _abi.ServerStats()
        self._c._lib.llh_server_get_stats(self._ptr, ctypes.byref(st))
        return {k: getattr(st, k) for k, _ in _abi.ServerStats._fields_}

    def evict_all(self):
        self._c._lib.llh_server_evict_all(self._ptr)

    def inject(self, kind, value):
        """Host-forced failures for tests (llh_server_inject): kind is
        "staging", "event" or "solve_spin"."""
        err = _abi.LlhError()
        _raise(self._c._lib.llh_server_inject(self._ptr, _abi.SERVER_INJECT[kind], int(value),
                                              ctypes.byref(err)), err)

    def close(self):
        if self._ptr:
            self._c._lib.llh_server_destroy(self._ptr)
            self._ptr = ctypes.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

"""ctypes bindings of include/llh.h (the C ABI of libllh.so)."""
import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libllh.so")
HEADER = os.path.join(os.path.dirname(os.path.dirname(HERE)), "include", "llh.h")

STATUS = {0: "ok", 1: "NotSquare", 2: "NotPositiveDefinite", 3: "WrongRows",
          16: "InvalidArgument", 17: "HipError", 18: "OutOfMemory", 19: "NoConvergence",
          20: "Rejected"}


def status_name(code):
    return STATUS.get(code, f"status {code}")


class LlhError(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int32), ("hip_error", ctypes.c_int32), ("a", ctypes.c_uint64),
                ("b", ctypes.c_uint64), ("column", ctypes.c_int64), ("message", ctypes.c_char * 192)]


_lib = None

_sz, _pd, _vp, _int = ctypes.c_size_t, ctypes.c_ssize_t, ctypes.c_void_p, ctypes.c_int
_err = ctypes.POINTER(LlhError)


class KernelStats(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_uint64), ("total_ms", ctypes.c_double),
                ("flops", ctypes.c_double), ("bytes", ctypes.c_double)]


class RefineOpts(ctypes.Structure):
    _fields_ = [("max_iter", ctypes.c_int32), ("flags", ctypes.c_int32)]


class RefineInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("fell_back", ctypes.c_int32), ("precision_used", ctypes.c_int32),
                ("residual", ctypes.c_double), ("anorm", ctypes.c_double)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("small_limit", ctypes.c_int32), ("small_waves", ctypes.c_int32),
                ("ir_iterations", ctypes.c_int32), ("medium_mode", ctypes.c_int32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("streams", ctypes.c_int32),
                ("max_inflight", ctypes.c_int32), ("small_limit", ctypes.c_int32),
                ("ir_iterations", ctypes.c_int32)]


class Request(ctypes.Structure):
    _fields_ = [("matrix_id", ctypes.c_uint64), ("n", ctypes.c_int32), ("m", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("a_on_device", ctypes.c_int32), ("a", ctypes.c_void_p),
                ("b", ctypes.c_void_p), ("x", ctypes.c_void_p)]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("cache_hit", ctypes.c_int32), ("batched", ctypes.c_int32),
                ("precision_used", ctypes.c_int32), ("failed_row", ctypes.c_int64),
                ("submit_s", ctypes.c_double), ("admit_s", ctypes.c_double), ("done_s", ctypes.c_double),
                ("residual", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("submitted", "completed", "batched", "cache_hits", "rejected",
                                               "steps", "bytes_in_use", "bytes_peak", "resident_factors")]


PREC = {"f64": 0, "f32": 1, "f32ir": 2, "f16": 3}
SERVER_INJECT = {"staging": 1, "event": 2, "solve_spin": 3}
FALLBACK_F64 = 1

KERNEL_CLASSES = {"trailing_syrk": 0, "panel_update": 1, "diag_factor": 2, "panel_trsm": 3,
                  "solve_gemm": 4, "solve_diag": 5}


def _declare(lib):
    lib.llh_context_create.argtypes = [_int, ctypes.POINTER(ctypes.c_void_p), _err]
    lib.llh_context_destroy.argtypes = [_vp]
    lib.llh_context_destroy.restype = None
    lib.llh_context_set_stream.argtypes = [_vp, _vp, _err]
    lib.llh_context_stream.argtypes = [_vp]
    lib.llh_context_stream.restype = _vp
    lib.llh_status_string.argtypes = [ctypes.c_int32]
    lib.llh_status_string.restype = ctypes.c_char_p
    lib.llh_context_set_profiling.argtypes = [_vp, _int, _err]
    lib.llh_context_set_profiling_classes.argtypes = [_vp, ctypes.c_uint32, _err]
    lib.llh_context_kernel_stats.argtypes = [_vp, _int, ctypes.POINTER(KernelStats), _err]
    lib.llh_context_set_solve_spin_limit.argtypes = [_vp, ctypes.c_int64, _err]
    lib.llh_factor_create.argtypes = [_vp, _int, _sz, _vp, _sz, ctypes.c_int32, ctypes.POINTER(_vp), _err]
    lib.llh_factor_solve.argtypes = [_vp, _vp, _sz, _vp, _sz, _vp, _sz, ctypes.POINTER(RefineOpts),
                                     ctypes.POINTER(RefineInfo), _err]
    lib.llh_factor_precision.argtypes = [_vp]
    lib.llh_factor_device_bytes.argtypes = [_vp]
    lib.llh_factor_device_bytes.restype = _sz
    lib.llh_factor_destroy.argtypes = [_vp]
    lib.llh_factor_destroy.restype = None
    lib.llh_dev_solve_mixed.argtypes = [_vp, _int, _sz, _vp, _sz, _sz, _vp, _sz, _vp, _sz,
                                        ctypes.POINTER(RefineOpts), ctypes.POINTER(RefineInfo), _err]
    i32p, u64p = ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_uint64)
    lib.llh_batch_plan_create.argtypes = [_vp, _int, _sz, i32p, i32p, ctypes.POINTER(BatchConfig),
                                          ctypes.POINTER(_vp), _err]
    lib.llh_batch_a_elems.argtypes = [_vp]
    lib.llh_batch_a_elems.restype = _sz
    lib.llh_batch_b_elems.argtypes = [_vp]
    lib.llh_batch_b_elems.restype = _sz
    lib.llh_batch_layout.argtypes = [_vp, u64p, i32p, u64p, _err]
    lib.llh_batch_get_config.argtypes = [_vp]
    lib.llh_batch_get_config.restype = BatchConfig
    lib.llh_batch_solve.argtypes = [_vp, _vp, _vp, _vp, _vp, _err]
    lib.llh_batch_autotune.argtypes = [_vp, _vp, _vp, _vp, _vp, ctypes.POINTER(ctypes.c_double), _err]
    lib.llh_batch_plan_destroy.argtypes = [_vp]
    lib.llh_batch_plan_destroy.restype = None
    lib.llh_server_create.argtypes = [_vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(_vp), _err]
    lib.llh_server_submit.argtypes = [_vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64), _err]
    lib.llh_server_poll.argtypes = [_vp, ctypes.c_uint64, ctypes.POINTER(Result)]
    lib.llh_server_wait.argtypes = [_vp, ctypes.c_uint64, ctypes.POINTER(Result), _err]
    lib.llh_server_get_stats.argtypes = [_vp, ctypes.POINTER(ServerStats)]
    lib.llh_server_get_stats.restype = None
    lib.llh_server_inject.argtypes = [_vp, _int, ctypes.c_int64, _err]
    lib.llh_server_evict_all.argtypes = [_vp]
    lib.llh_server_evict_all.restype = None
    lib.llh_server_destroy.argtypes = [_vp]
    lib.llh_server_destroy.restype = None
    for s in ("d", "s"):
        f = lambda n: getattr(lib, f"llh_{n}_{s}")  # noqa: E731
        f("cholesky_inplace_dirty").argtypes = [_vp, _sz, _sz, _vp, _pd, _pd, _err]
        f("cholesky_inplace").argtypes = [_vp, _sz, _sz, _vp, _pd, _pd, _err]
        f("triangular_inplace").argtypes = [_vp, _sz, _sz, _vp, _pd, _pd, _int, _err]
        f("is_triangular").argtypes = [_vp, _sz, _sz, _vp, _pd, _pd, _int, ctypes.POINTER(_int), _err]
        f("solve_triangular_inplace").argtypes = [_vp, _sz, _sz, _vp, _pd, _pd, _sz, _sz, _vp, _pd, _pd,
                                                  _int, _err]
        f("solve_triangular_system").argtypes = [_vp, _sz, _sz, _vp, _pd, _pd, _vp, _pd, _int, _sz, _sz, _vp,
                                                 _pd, _pd, _int, _err]
        f("solvec_inplace").argtypes = [_vp, _sz, _sz, _vp, _pd, _pd, _sz, _sz, _vp, _pd, _pd, _err]
        f("invc_inplace").argtypes = [_vp, _sz, _sz, _vp, _pd, _pd, _vp, _pd, _pd, _err]
        f("dev_potrf").argtypes = [_vp, _sz, _vp, _sz, _int, _err]
        f("dev_solve_triangular").argtypes = [_vp, _int, _int, _sz, _vp, _sz, _sz, _vp, _sz, _err]
        f("dev_solvec").argtypes = [_vp, _sz, _vp, _sz, _sz, _vp, _sz, _err]
        f("dev_invc").argtypes = [_vp, _sz, _vp, _sz, _vp, _sz, _err]
        f("dev_orthonormalize").argtypes = [_vp, _sz, _sz, _vp, _sz, _vp, _sz, _err]
        f("dev_apply_constraints").argtypes = [_vp, _sz, _sz, _vp, _sz, _sz, _vp, _sz, _vp, _sz, _err]
    lib.llh_dev_gen_matrix_d.argtypes = [_vp, _int, ctypes.c_uint64, ctypes.c_uint64, _sz, _sz, _vp, _sz, _err]
    lib.llh_dev_gen_spd_d.argtypes = [_vp, _int, ctypes.c_uint64, _sz, ctypes.c_double, _vp, _sz, _vp, _err]
    lib.llh_dev_solvec_noalloc_d.argtypes = [_vp, _sz, _vp, _sz, _sz, _vp, _sz, _err]


def _share_torch_runtime():
    """torch wheels carry their own HIP runtime under the same SONAME
    (libamdhip64.so.7).  Loading torch first makes libllh.so bind to that one,
    so a process never holds two HIP runtimes; without torch the system
    runtime is used."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Loads libllh.so; raises if it has not been built (no fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C linfa-linalg_amd` "
                               "or __graft_entry__.build()")
        _share_torch_runtime()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
        try:
            from . import _abi_ext
            _abi_ext.declare(_lib)
        except ImportError:
            pass
    return _lib


def declared_symbols(header=HEADER):
    """Every function name a header declares, macro families expanded by the C
    preprocessor."""
    import subprocess
    out = subprocess.run(["gcc", "-E", "-P", "-x", "c", header], check=True, capture_output=True,
                         text=True).stdout
    return sorted(set(re.findall(r"\b(llh_[a-z0-9_]+)\s*\(", out)))

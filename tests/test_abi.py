"""CPU-side checks of the boundary: the C-ABI library loads and exports every
symbol include/llh.h declares, and the host mirror exposes the reference's
names.  No compute call is made here (there is no GPU in this environment)."""
import ctypes
import os

import pytest

from linfa_linalg_hip import _abi

HEADERS = sorted(os.path.join(os.path.dirname(_abi.HEADER), f)
                 for f in os.listdir(os.path.dirname(_abi.HEADER)) if f.endswith(".h"))


def test_library_built():
    assert os.path.exists(_abi.LIB_PATH), "run __graft_entry__.build() first"


@pytest.mark.parametrize("header", HEADERS, ids=os.path.basename)
def test_exports_every_declared_symbol(header):
    lib = ctypes.CDLL(_abi.LIB_PATH)
    names = _abi.declared_symbols(header)
    assert names, header
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_status_strings_without_gpu():
    lib = _abi.lib()
    assert lib.llh_status_string(2) == b"NotPositiveDefinite"
    assert lib.llh_status_string(3) == b"WrongRows"


def test_mirror_surface():
    import linfa_linalg_hip as llh
    for name in ["cholesky_inplace_dirty", "cholesky_into_dirty", "cholesky_inplace", "cholesky_into",
                 "cholesky_dirty", "cholesky", "solvec_inplace", "solvec_into", "solvec",
                 "invc_inplace", "invc", "triangular_inplace", "into_triangular", "is_triangular",
                 "solve_triangular_inplace", "solve_triangular_into", "solve_triangular"]:
        assert callable(getattr(llh, name)), name
    assert int(llh.UPLO.Upper) == 0 and int(llh.UPLO.Lower) == 1
    assert issubclass(llh.NotSquare, llh.LinalgError)

"""CPU checks of the built artefacts: the contact library loads and exports every
entry point include/atac.h declares (no GPU call is made)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "atac.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(atac_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "atac_batch_eval" in syms and "atac_batch_vjp" in syms and len(syms) >= 20


def test_library_exports_every_declared_symbol():
    from atac import _lib
    assert os.path.exists(_lib.LIB_PATH), "run __graft_entry__.build() first"
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert set(_lib.EXPORTED) == set(declared_symbols())


def test_library_reports_version_and_errors_without_gpu():
    from atac import _lib
    L = _lib.lib()
    assert L.atac_version() >= 100
    # a null handle is rejected before any device work
    rc = L.atac_batch_eval(None, 0, None, None)
    assert rc == _lib.ATAC_ERR_INVALID
    assert b"null batch" in L.atac_last_error()


def test_grid_size_limits_refused_before_allocation():
    """Grids beyond the detection's 32-bit cell offsets (2^27 nodes; the padded
    cell-bound table then stays below 2^30 entries) are refused before any value
    is read or allocated."""
    from atac import _lib
    L = _lib.lib()
    one = (ctypes.c_float * 1)(0.0)
    box0 = (ctypes.c_double * 3)(0.0, 0.0, 0.0)
    box1 = (ctypes.c_double * 3)(1.0, 1.0, 1.0)
    d = (ctypes.c_double * 3)(1e-3, 1e-3, 1e-3)
    out = ctypes.c_void_p()
    for dims, msg in (((1024, 1024, 128), b"2^27"), ((1, 1 << 27, 1), b"2^27"), ((2, 4096, 16384), b"2^27")):
        rc = L.atac_grid_create(one, *dims, box0, box1, d, ctypes.byref(out))
        assert rc == _lib.ATAC_ERR_INVALID, dims
        assert msg in L.atac_last_error(), (dims, L.atac_last_error())
        assert not out.value


def test_package_imports_without_loading_library():
    import importlib
    import atac
    importlib.reload(atac)
    assert atac.EVAL_DERIV == 1 and atac.EVAL_CHECKPOINT == 4


def test_noalloc_bvh_force_refuses_sizes_beyond_its_scratch():
    """atac_bvh_force_noalloc checks the point count against its fixed scratch
    before anything else (no device call is made)."""
    from atac import _lib
    L = _lib.lib()
    rc = L.atac_bvh_force_noalloc(None, None, None, _lib.BVH_NOALLOC_MAX_POINTS + 1, 0, None, None, None)
    assert rc == _lib.ATAC_ERR_INVALID
    assert b"fixed scratch" in L.atac_last_error()
    rc = L.atac_bvh_force_noalloc(None, None, None, 16, 0, None, None, None)
    assert rc == _lib.ATAC_ERR_INVALID and b"bad arguments" in L.atac_last_error()

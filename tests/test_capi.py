"""The C ABI library loads (no GPU needed) and exports every entry point
declared in include/nrf.h; the Python binding declares each of them."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "nrf.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(nrf_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_step():
    names = declared()
    for n in ("nrf_tsdf_integrate", "nrf_mc_extract", "nrf_voxel_knn", "nrf_warp_apply_twist",
              "nrf_solver_solve", "nrf_solver_solve_gn", "nrf_pipeline_step"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from cudafusion_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libnrf.so not built")
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared()) <= set(_lib.exported_symbols())


def test_status_strings_without_device():
    from cudafusion_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libnrf.so not built")
    lib = _lib.load()
    assert lib.nrf_status_string(0) == b"ok"
    assert lib.nrf_status_string(6).startswith(b"persistent kernel")
    assert b"gfx950" in lib.nrf_build_info()
    # argument validation happens on the host, before any device call
    vp = _lib.volume_params((0, 4, 4), 0.01, 0.03, (0, 0, 0))
    assert lib.nrf_tsdf_reset(C.byref(vp), C.c_void_p(16), None) == 1
    assert lib.nrf_mc_workspace_size(C.byref(vp)) == 0

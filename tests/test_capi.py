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


def test_struct_layouts_match_the_header(tmp_path):
    """Every ctypes record of the binding has the size and field offsets the
    C compiler gives the header's struct."""
    import shutil
    import subprocess

    from cudafusion_amd import _lib as L
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    pairs = {"nrf_volume_params": L.VolumeParams, "nrf_camera": L.Camera, "nrf_solver_params": L.SolverParams,
             "nrf_gn_params": L.GNParams, "nrf_pcg_result": L.PcgResult, "nrf_solver_totals": L.SolverTotals,
             "nrf_pipeline_config": L.PipelineConfig, "nrf_batch_frame": L.BatchFrame,
             "nrf_assoc_params": L.AssocParams, "nrf_track_params": L.TrackParams,
             "nrf_track_stats": L.TrackStats,
             "nrf_batch_tuning": L.BatchTuning}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "nrf.h"', "int main(void) {"]
    for cname, py in pairs.items():
        lines.append(f'  printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                               text=True).stdout.splitlines())
    for cname, py in pairs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, f"{cname}.{f}"


def test_host_checks_without_device():
    """Entry points that reject their arguments on the host, before any
    device call; the learned preconditioner's parameter count equals the
    oracle's definition."""
    from cudafusion_amd import _lib
    from oracle import pyoracle as po
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libnrf.so not built")
    lib = _lib.load()
    assert lib.nrf_lp_param_count() == po.lp_param_count() == 46025
    assert not lib.nrf_batch_create(None, 0, None, 0)
    assert not lib.nrf_batch_create(None, 40, None, 0)
    assert not lib.nrf_lp_create(2, 100, 1)  # no such variant
    assert not lib.nrf_lp_create(0, 0, 1)
    assert lib.nrf_batch_step(None, None, None) == 1
    assert lib.nrf_lp_adam(None, 1e-3, 1, None) == 1
    sp = _lib.solver_params(5, 30, 0.0, 10.0, 1.0, 1e-2, 0, 0, 0, 0, 4)  # pcg_pipelined 4 is not a form
    assert not lib.nrf_solver_create(C.byref(sp), 100)
    # the PCG synchronisation tags hold 1 + 2 * iterations in 16 bits, for the
    # main PCG and the refinement PCG alike
    sp = _lib.solver_params(5, 32768, pcg_pipelined=2)
    assert not lib.nrf_solver_create(C.byref(sp), 100)
    sp = _lib.solver_params(5, 30, pcg_pipelined=2, refine_iterations=32768)
    assert not lib.nrf_solver_create(C.byref(sp), 100)
    # the plain form (the reference's structure) exists in FP32 without
    # refinement or kept pairs only
    for kw in (dict(jtj_precision=1), dict(refine_jacobi=1), dict(psd_dropped_pairs=1), dict(plain_form=2)):
        sp = _lib.solver_params(5, 30, **dict(dict(plain_form=1), **kw))
        assert not lib.nrf_solver_create(C.byref(sp), 100), kw
    # the workspace-free marching cubes refuses a volume beyond its scratch
    # (dims[1] * dims[2] rows) before any device call
    vp = _lib.volume_params((8, 600, 600), 0.01, 0.03, (0, 0, 0))
    counts = (C.c_int32 * 2)()
    assert lib.nrf_mc_extract_noalloc(C.byref(vp), C.c_void_p(256), None, None, None, 0, C.addressof(counts),
                                      None) == 2
    assert b"262144" in lib.nrf_last_error()

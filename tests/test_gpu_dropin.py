"""Runs the C++ program that drives the drop-in class API (TSDFVolume,
WarpField, GaussNewtonSolver) the way the reference's own tests do."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_dropin")


def test_dropin_class_api(gpu):
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/test_dropin is not built (run __graft_entry__.build())")
    r = subprocess.run([EXE], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "FAIL" not in r.stdout and r.stdout.count("PASS") >= 7

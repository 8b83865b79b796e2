"""The C ABI library: it loads without a GPU, exports every function that
include/mppi_hip.h declares, and its host-side validation works; and the port
of the reference's 46 driver tests (tests/cpp/mppi_driver_tests.cpp) passes —
its 21 device-free tests here, all 46 on the GPU."""
import ctypes
import os
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "mppi_hip.h"
CPP_TESTS = ROOT / "tests" / "cpp" / "mppi_driver_tests"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(mppi_hip_\w+)\s*\(", text)))


def test_header_declares_the_backend_entry_points():
    names = declared_functions()
    for required in ("mppi_hip_create", "mppi_hip_destroy", "mppi_hip_link_plugin",
                     "mppi_hip_bind_params", "mppi_hip_step", "mppi_hip_simulate",
                     "mppi_hip_simulate_batch", "mppi_hip_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from mppi_metal_amd import _native
    lib = _native.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    bound = {name for name, _, _ in _native.SIGNATURES}
    assert set(declared_functions()) == bound


def test_create_validates_config_without_a_device():
    from mppi_metal_amd import _native
    lib = _native.load()
    sig = (ctypes.c_float * 2)(1.0, 1.0)
    lo = (ctypes.c_float * 2)(-1.0, -1.0)
    hi = (ctypes.c_float * 2)(1.0, 1.0)
    ctx = ctypes.c_void_p()
    cfg = _native.MppiHipConfig(4, 2, 10, 0, 1.0, 0, sig, lo, hi, 0, -1)
    assert lib.mppi_hip_create(ctypes.byref(cfg), ctypes.byref(ctx)) == -1
    assert b"sample_count" in lib.mppi_hip_last_error(None)
    cfg = _native.MppiHipConfig(4, 2, 10, 64, 1.0, 1, (ctypes.c_float * 4)(1, 2, 2, 1), lo, hi, 0, -1)
    assert lib.mppi_hip_create(ctypes.byref(cfg), ctypes.byref(ctx)) == -1
    assert b"positive definite" in lib.mppi_hip_last_error(None)
    assert lib.mppi_hip_step(None, 0, None, None, None, 0, None, 0, None, None) == -1
    assert lib.mppi_hip_version().startswith(b"mppi_hip")


def test_reference_driver_tests_device_free():
    out = subprocess.run([str(CPP_TESTS), "--cpu-only"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "0 failed" in out.stdout


@pytest.mark.gpu
def test_reference_driver_tests_all(require_gpu):
    out = subprocess.run([str(CPP_TESTS)], capture_output=True, text=True, timeout=300)
    print(out.stdout[-3000:])
    assert out.returncode == 0, out.stdout[-4000:] + out.stderr[-4000:]
    assert "0 failed" in out.stdout

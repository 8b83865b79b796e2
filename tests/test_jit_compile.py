"""hipRTC compilation of the library kernels with a plugin, on the CPU
(mppi_hip_compile_check: same sources and options as linking, no device).
These catch JIT-only compile errors before a GPU run.

torch is imported first, as in every Python process that drives the library:
torch loads its own code-object compiler (libamd_comgr), which hipRTC then
uses, and its clang differs from the system ROCm's in places (the type of a
__builtin_shufflevector result, for one).  test_compiles_without_torch runs
the same checks in a process without torch, the C++ callers' case, where the
system ROCm's compiler is used."""
import ctypes
import subprocess
import sys
from pathlib import Path

import pytest
import torch  # noqa: F401  (loads torch's libamd_comgr before the library)

from mppi_metal_amd import _native
from mppi_metal_amd.build import BUNDLED_PLUGINS
from mppi_metal_amd.controller import PLUGIN_DIR


def _check(src, dyn, stage, term, D, C, cov=0, sets=0, mode=0):
    lib = _native.load()
    log = ctypes.create_string_buffer(1 << 16)
    rc = lib.mppi_hip_compile_check(src, str(PLUGIN_DIR).encode(), dyn.encode(), stage.encode(), term.encode(),
                                    D, C, cov, mode, sets, log, len(log))
    return rc, log.value.decode(errors="replace")


@pytest.mark.parametrize("plugin", [p for p in BUNDLED_PLUGINS if p[0] in ("unicycle.hip", "bicycle.hip", "quadrotor.hip")],
                         ids=lambda p: p[0])
def test_bundled_plugin_compiles_with_surrogate_kernels(plugin):
    name, dyn, stage, term, D, C = plugin
    rc, log = _check((PLUGIN_DIR / name).read_bytes(), dyn, stage, term, D, C, sets=2)
    assert rc == 0, log


@pytest.mark.parametrize("mode", [1, 2], ids=["fp16", "mixed"])
def test_quadrotor_compiles_in_fast_modes(mode):
    name, dyn, stage, term, D, C = [p for p in BUNDLED_PLUGINS if p[0] == "quadrotor.hip"][0]
    rc, log = _check((PLUGIN_DIR / name).read_bytes(), dyn, stage, term, D, C, cov=1, sets=1, mode=mode)
    assert rc == 0, log


def test_compile_error_is_reported_with_the_log():
    src = b'#include "mppi_core.h"\nMPPI_CALLABLE void dyn(float* x, const float* u, const uint8_t* m, mppi_rng r) { x[0] += undefined_name; }\n'
    rc, log = _check(src, "dyn", "", "", 2, 1)
    assert rc == -4                                  # MPPI_HIP_E_COMPILE
    assert "undefined_name" in log


def test_rejects_missing_arguments():
    rc, log = _check(b"", "", "", "", 0, 0)
    assert rc == -1 and "required" in log


_NO_TORCH = r"""
import ctypes, sys
sys.path.insert(0, sys.argv[1])
from mppi_metal_amd import _native
from mppi_metal_amd.build import BUNDLED_PLUGINS
from mppi_metal_amd.controller import PLUGIN_DIR
assert "torch" not in sys.modules
lib = _native.load()
for name, dyn, stage, term, D, C in BUNDLED_PLUGINS:
    if name not in ("bicycle.hip", "quadrotor.hip"):
        continue
    for cov, sets, mode in ((0, 2, 0), (1, 1, 2)) if name == "quadrotor.hip" else ((0, 2, 0),):
        log = ctypes.create_string_buffer(1 << 16)
        rc = lib.mppi_hip_compile_check((PLUGIN_DIR / name).read_bytes(), str(PLUGIN_DIR).encode(), dyn.encode(),
                                        stage.encode(), term.encode(), D, C, cov, mode, sets, log, len(log))
        assert rc == 0, (name, cov, sets, mode, log.value.decode(errors="replace"))
assert "torch" not in sys.modules
print("ok")
"""


def test_compiles_without_torch():
    root = Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, "-c", _NO_TORCH, str(root)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout[-2000:] + r.stderr[-4000:]

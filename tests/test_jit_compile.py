"""hipRTC compilation of the library kernels with a plugin, on the CPU
(mppi_hip_compile_check: same sources and options as linking, no device).
These catch JIT-only compile errors before a GPU run."""
import ctypes

import pytest

from mppi_metal_amd import _native
from mppi_metal_amd.build import BUNDLED_PLUGINS
from mppi_metal_amd.controller import PLUGIN_DIR


def _check(src, dyn, stage, term, D, C, cov=0, sets=0, mode=0):
    lib = _native.load()
    log = ctypes.create_string_buffer(1 << 16)
    rc = lib.mppi_hip_compile_check(src, str(PLUGIN_DIR).encode(), dyn.encode(), stage.encode(), term.encode(),
                                    D, C, cov, mode, sets, log, len(log))
    return rc, log.value.decode(errors="replace")


@pytest.mark.parametrize("plugin", [p for p in BUNDLED_PLUGINS if p[0] in ("unicycle.hip", "bicycle.hip")],
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

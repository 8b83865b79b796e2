"""mppi_hip_calibrate: the measured issue ceilings that the benchmark's
roofline figures are stated against lie between what one wave64
instruction per four cycles per SIMD and the published peaks allow."""
import ctypes

import pytest

from mppi_metal_amd import _native


@pytest.mark.gpu
def test_calibrate_reports_plausible_ceilings(require_gpu):
    lib = _native.load()
    v, m, x = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    assert lib.mppi_hip_calibrate(0, ctypes.byref(v), ctypes.byref(m), ctypes.byref(x)) == 0
    print(f"[calibrate] valu={v.value / 1e12:.2f} T lane-op/s mfma={m.value / 1e12:.1f} TFLOP/s "
          f"mixed/sum={x.value:.3f}", flush=True)
    # 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz = 39.3 T; the published packed
    # peak 78.65 T lane-ops and 2.5 PFLOP/s (at 2.4 GHz, so with a little slack
    # for the clock) bound them from above
    assert 10e12 < v.value <= 78.65e12
    assert 300e12 < m.value <= 2.6e15
    # a stream of both kinds takes between the longer part and the sum of both
    assert 0.45 < x.value < 1.1
    assert _native.STATUS_NAMES[lib.mppi_hip_calibrate(0, None, ctypes.byref(m), None)] == "E_INVALID"

"""Philox generators: Random123 known-answer vectors.

The oracle's own Philox (oracle/oracle_philox.h), the library header's
Philox4x32-10 (include/mppi_hip/mppi_philox.h, compiled for the host) and the
scene generator's Python Philox must all reproduce the published vectors.
The reference's generator is Philox2x32-10 (shaders/philox.h:19-36); its
vectors pin the 2x32 restatement.
"""
import numpy as np
import pytest

import oracle_lib as ol
from mppi_metal_amd import scenes

KAT_4x32 = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]
KAT_2x32 = [
    ((0, 0), 0, (0xFF1DAE59, 0x6CD10DF2)),
    ((0xFFFFFFFF, 0xFFFFFFFF), 0xFFFFFFFF, (0x2C3F628B, 0xAB4FD7AD)),
    ((0x243F6A88, 0x85A308D3), 0x13198A2E, (0xDD7CE038, 0xF62A4C12)),
    # The reference's first draw at seed 42, counter (step 0, index 0).
    ((0, 0), 42, (0x2531BEF0, 0xFC3BC596)),
]


@pytest.mark.parametrize("ctr,key,want", KAT_4x32)
def test_philox4x32_kat(ctr, key, want):
    assert ol.philox4x32(ctr, key) == list(want)
    assert ol.philox4x32(ctr, key, library=True) == list(want)
    assert ol.philox4x32(ctr, key, library="split") == list(want)
    assert scenes.philox4x32_10(ctr, key) == list(want)


@pytest.mark.parametrize("ctr,key,want", KAT_2x32)
def test_philox2x32_kat(ctr, key, want):
    assert ol.philox2x32(ctr, key) == list(want)


def test_library_philox_matches_oracle_on_random_counters():
    rng = np.random.default_rng(1)
    for _ in range(2000):
        c = [int(v) for v in rng.integers(0, 2**32, 4, dtype=np.uint64)]
        k = [int(v) for v in rng.integers(0, 2**32, 2, dtype=np.uint64)]
        assert ol.philox4x32(c, k, library=True) == ol.philox4x32(c, k)
        assert ol.philox4x32(c, k, library="split") == ol.philox4x32(c, k)


def test_uniform_conversion_edges():
    # (2j + 1) * 2^-24 with j = x >> 9: exact, symmetric, never 0 or 1.
    assert scenes.u01(0) == 2.0 ** -24
    assert scenes.u01(0xFFFFFFFF) == 1.0 - 2.0 ** -24
    assert scenes.u01(0x80000000) == 0.5 + 2.0 ** -24


def test_normals_distribution():
    # 40k normals from consecutive blocks: mean ~0, var ~1, no tail beyond the
    # largest Box-Muller radius sqrt(-2 ln 2^-24) = 5.77.
    z = np.concatenate([ol.normals4(b, s, 0, 0, 0, 42) for b in range(100) for s in range(100)])
    assert abs(z.mean()) < 0.03
    assert abs(z.var() - 1.0) < 0.03
    assert np.abs(z).max() < 5.771


def test_normals_fp32_vs_fp64():
    # Same integer stream, FP32 Box-Muller (library elementary functions)
    # against FP64 libm: error is stated and bounded here.
    err = 0.0
    for b in range(200):
        z32 = ol.normals4(b, 7, 3, 0, 5, 2**40 + 17).astype(np.float64)
        z64 = ol.normals4(b, 7, 3, 0, 5, 2**40 + 17, fp64=True)
        err = max(err, float(np.max(np.abs(z32 - z64))))
    assert err < 4e-6, err


def test_binary16_rounding_matches_numpy():
    """The FP16 mode's noise rounding: the library's bit-level
    mppi_round_f16 and the oracle's scaling restatement both equal numpy's
    float32 -> float16 conversion (IEEE round to nearest even), including
    ties, half subnormals and overflow."""
    import ctypes
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.standard_normal(200000).astype(np.float32) * 3,
        rng.uniform(-1e-4, 1e-4, 20000).astype(np.float32),          # half subnormals
        np.float32(2.0) ** rng.integers(-30, 17, 5000).astype(np.float32) * rng.uniform(1, 2, 5000).astype(np.float32),
        # exact ties between two binary16 neighbours
        (np.arange(1, 2049, dtype=np.float32) + np.float32(0.5)) * np.float32(2.0 ** -10),
        np.array([0.0, -0.0, 65504.0, 65519.0, 65520.0, -70000.0, 5.96e-8, 2.98e-8, 2.99e-8], np.float32),
    ]).astype(np.float32)
    want = x.astype(np.float16).astype(np.float32)
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    for fn in ("oracle_round_f16", "oracle_library_round_f16"):
        got = np.empty_like(x)
        getattr(ol.lib(), fn)(fp(x), fp(got), ctypes.c_size_t(x.size))
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), fn

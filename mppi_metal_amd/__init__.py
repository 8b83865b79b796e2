"""mppi_metal_amd — MPPI control on AMD Instinct MI355X (gfx950).

The drop-in for the reference's `mppi_metal::MPPIDriver` is the C++ facade
(include/mppi_metal/mppi_driver.hpp) in libmppi_hip.so, over the C ABI
(include/mppi_hip.h).  This Python package binds the same C ABI
(`HipMPPI`) and holds the bundled model plugins (plugins/*.hip), the scene
generator and the build script.
"""
from .controller import HipBatch, HipMPPI, HipServer, MppiHipError, PLUGIN_DIR  # noqa: F401
from . import scenes  # noqa: F401

__all__ = ["HipBatch", "HipMPPI", "HipServer", "MppiHipError", "PLUGIN_DIR", "scenes"]

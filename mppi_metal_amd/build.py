"""Build the native parts of mppi_metal_amd in-tree.

Steps (each compiles every time it is run):
  1. embed the JIT sources (library kernels + device headers) into
     csrc/_jit_sources.inc as raw string literals;
  2. compile the kernels ahead of time for gfx950 once per bundled plugin
     (hipcc --cuda-device-only), so a syntax or type error in the kernel
     source fails the build here rather than at hipRTC time on the GPU box;
  3. link libmppi_hip.so (C ABI + the C++ MPPIDriver facade) with hipcc;
  4. build the C++ port of the reference's driver tests (tests/cpp);
  5. build the CPU oracle (oracle/liboracle.so) — test infrastructure.

Usage: python -m mppi_metal_amd.build   (or build() from __graft_entry__).
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INC = ROOT / "include"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are kept out of all device code.
NO_PACKED_FP32 = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

# (plugin file, dynamics, stage cost, terminal cost, state dim, control dim)
BUNDLED_PLUGINS = [
    ("double_integrator.hip", "mppi_dynamics", "mppi_stage_cost", "mppi_terminal_cost", 4, 2),
    ("unicycle.hip", "mppi_dynamics", "unicycle_stage_cost_obstacles", "unicycle_terminal_cost", 4, 2),
    ("bicycle.hip", "bicycle_dynamics", "bicycle_pursuit_cost", "bicycle_terminal_cost", 6, 2),
    ("quadrotor.hip", "quad_dynamics", "quad_stage_cost", "quad_terminal_cost", 12, 4),
]

EMBEDDED = [
    ("kKernelSource", CSRC / "mppi_kernels.hip"),
    ("kLaunchHeader", CSRC / "mppi_launch.h"),
    ("kCoreHeader", INC / "mppi_hip" / "mppi_core.h"),
    ("kMathHeader", INC / "mppi_hip" / "mppi_math.h"),
    ("kPhiloxHeader", INC / "mppi_hip" / "mppi_philox.h"),
]


def _run(cmd, cwd=None):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True, cwd=cwd)


def embed_sources() -> Path:
    out = CSRC / "_jit_sources.inc"
    parts = ["// Generated by mppi_metal_amd/build.py: JIT sources as string literals.\n"]
    for name, path in EMBEDDED:
        text = path.read_text()
        if ")MPPISRC\"" in text:
            raise RuntimeError(f"{path} contains the raw-string delimiter")
        # Split into chunks: some compilers cap a single string literal's length.
        chunks = [text[i:i + 8000] for i in range(0, len(text), 8000)] or [""]
        body = "\n".join(f'R"MPPISRC({c})MPPISRC"' for c in chunks)
        parts.append(f"static const char {name}[] =\n{body};\n")
    new = "\n".join(parts)
    if not out.exists() or out.read_text() != new:
        out.write_text(new)
    return out


def check_kernels():
    """Ahead-of-time device compile of the kernels with each bundled plugin."""
    tmp = ROOT / "_build" / "kernel_check"
    tmp.mkdir(parents=True, exist_ok=True)
    for plugin, dyn, stage, term, sdim, cdim in BUNDLED_PLUGINS:
        pdir = tmp / plugin.replace(".hip", "")
        pdir.mkdir(exist_ok=True)
        (pdir / "mppi_user_plugin.hip").write_text((PKG / "plugins" / plugin).read_text())
        for cov in (0, 1):
            _run([HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-ffp-contract=off",
                  *NO_PACKED_FP32, "--cuda-device-only", "-c",
                  f"-I{pdir}", f"-I{INC / 'mppi_hip'}", f"-I{CSRC}",
                  f"-DMPPI_SDIM={sdim}", f"-DMPPI_CDIM={cdim}", f"-DMPPI_DYNAMICS={dyn}",
                  f"-DMPPI_STAGE_COST={stage}", f"-DMPPI_TERMINAL_COST={term}",
                  f"-DMPPI_COV_FULL={cov}",
                  CSRC / "mppi_kernels.hip", "-o", pdir / f"kernels_cov{cov}.o"])


def build_library() -> Path:
    out = PKG / "libmppi_hip.so"
    srcs = [CSRC / "hip_backend.cpp", CSRC / "batch_backend.cpp", CSRC / "mppi_driver.cpp"]
    _run([HIPCC, f"--offload-arch={ARCH}", "-O2", "-std=c++17", "-fPIC", "-shared",
          *NO_PACKED_FP32, "-Wall", f"-I{INC}", f"-I{INC / 'mppi_hip'}", f"-I{CSRC}",
          *srcs, "-lhiprtc", "-o", out])
    return out


def build_cpp_tests() -> Path:
    out = ROOT / "tests" / "cpp" / "mppi_driver_tests"
    _run(["g++", "-O1", "-std=c++17", f"-I{INC}", ROOT / "tests" / "cpp" / "mppi_driver_tests.cpp",
          f"-L{PKG}", "-lmppi_hip", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/../../mppi_metal_amd",
          "-o", out])
    return out


def build_oracle() -> Path:
    _run(["make", "-s", "-C", ROOT / "oracle", "-B"])
    return ROOT / "oracle" / "liboracle.so"


def build_all():
    embed_sources()
    check_kernels()
    lib = build_library()
    build_cpp_tests()
    build_oracle()
    return lib


if __name__ == "__main__":
    build_all()
    sys.exit(0)

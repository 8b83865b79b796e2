"""Build the native parts of mppi_metal_amd in-tree.

Steps (each compiles every time it is run):
  1. embed the JIT sources (library kernels + device headers) into
     csrc/_jit_sources.inc as raw string literals;
  2. compile the kernels ahead of time for gfx950 once per bundled plugin
     (hipcc --cuda-device-only), so a syntax or type error in the kernel
     source fails the build here rather than at hipRTC time on the GPU box;
  3. link libmppi_hip.so (C ABI + the C++ MPPIDriver facade) with hipcc;
  4. build the C++ port of the reference's driver tests (tests/cpp);
  5. build the example programs (examples/);
  6. build the CPU oracle (oracle/liboracle.so) — test infrastructure.

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
    ("point_mass3.hip", "pm3_dynamics", "pm3_stage_cost", "pm3_terminal_cost", 6, 3),
]

EMBEDDED = [
    ("kKernelSource", CSRC / "mppi_kernels.hip"),
    ("kLaunchHeader", CSRC / "mppi_launch.h"),
    ("kCoreHeader", INC / "mppi_hip" / "mppi_core.h"),
    ("kMathHeader", INC / "mppi_hip" / "mppi_math.h"),
    ("kPhiloxHeader", INC / "mppi_hip" / "mppi_philox.h"),
    ("kMlpHeader", CSRC / "mppi_mlp.h"),
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


def check_kernels(lib_path: Path):
    """Compile the library kernels with each bundled plugin through hipRTC,
    exactly as linking does at run time (mppi_hip_compile_check: no device
    needed), in every kernel set: plain, full covariance with the batch
    kernels, and the surrogate kernels where state + control <= 16."""
    import ctypes
    lib = ctypes.CDLL(str(lib_path))
    fn = lib.mppi_hip_compile_check
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_char_p] * 5 + [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint32,
                                           ctypes.c_uint32, ctypes.c_char_p, ctypes.c_size_t]
    log = ctypes.create_string_buffer(1 << 16)
    for plugin, dyn, stage, term, sdim, cdim in BUNDLED_PLUGINS:
        src = (PKG / "plugins" / plugin).read_bytes()
        # (covariance_full, numeric_mode, kernel sets)
        variants = [(0, 0, 0), (1, 0, 1), (1, 2, 1), (0, 1, 0)] + ([(0, 0, 2)] if sdim + cdim <= 16 else [])
        for cov, mode, sets in variants:
            rc = fn(src, str(PKG / "plugins").encode(), dyn.encode(), stage.encode(), term.encode(),
                    sdim, cdim, cov, mode, sets, log, len(log))
            if rc != 0:
                raise RuntimeError(f"{plugin} (covariance_full={cov}, numeric_mode={mode}, kernel_sets={sets}) "
                                   f"failed to compile:\n{log.value.decode(errors='replace')}")


def build_library() -> Path:
    out = PKG / "libmppi_hip.so"
    srcs = [CSRC / "hip_backend.cpp", CSRC / "batch_backend.cpp", CSRC / "server_backend.cpp",
            CSRC / "surrogate.hip", CSRC / "noalloc.hip", CSRC / "calibrate.hip", CSRC / "mppi_driver.cpp"]
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


def build_examples() -> None:
    """examples/unicycle_closed_loop, and, where the reference's sources are
    present, the reference's unicycle examples compiled unchanged against this
    library (examples/_ref/, outside the repository's history)."""
    _run(["make", "-s", "-C", ROOT / "examples", "-B", "all"])
    if Path("/root/reference/examples").is_dir():
        _run(["make", "-s", "-C", ROOT / "examples", "-B", "reference"])


def build_oracle() -> Path:
    _run(["make", "-s", "-C", ROOT / "oracle", "-B"])
    return ROOT / "oracle" / "liboracle.so"


def build_all():
    embed_sources()
    lib = build_library()
    check_kernels(lib)
    build_cpp_tests()
    build_examples()
    build_oracle()
    return lib


if __name__ == "__main__":
    build_all()
    sys.exit(0)

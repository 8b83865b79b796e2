"""Shape of the persistent batch kernels in the gfx950 assembly the compiler
makes of them (CPU only: hipcc cross-compiles).

A persistent kernel loops: thread 0 takes a ticket, a barrier publishes the
tile index, every thread runs the tile, a barrier closes the iteration.  When
the loop's exit depended on a value the compiler treats as varying (the index
read from LDS), it compiled the loop as two nested loops: the ticket fetch in
the outer one, the barriers and the tile in an inner one that thread 0 left
after a tile ending in thread-0-only writes, while the other lanes of its
wavefront re-entered the inner loop -- re-running the tile with a stale index
and, thread 0 being masked off, never leaving it (the launch never ended).
This test compiles the batch kernels with the bicycle plugin and checks, for
every persistent kernel, that the ticket fetch and the barrier after it sit
in the same loop, and that every loop holding a barrier is continued by a
scalar branch with the wavefront's lanes reconverged."""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HIPCC = "/opt/rocm/bin/hipcc"


def _assembly(tmp_path):
    (tmp_path / "mppi_user_plugin.hip").write_text((ROOT / "mppi_metal_amd" / "plugins" / "bicycle.hip").read_text())
    out = tmp_path / "k.s"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-S",
           "-DMPPI_SDIM=6", "-DMPPI_CDIM=2", "-DMPPI_DYNAMICS=bicycle_dynamics",
           "-DMPPI_STAGE_COST=bicycle_pursuit_cost", "-DMPPI_TERMINAL_COST=bicycle_terminal_cost",
           "-DMPPI_BATCH_KERNELS=1", f"-I{tmp_path}", f"-I{ROOT / 'include' / 'mppi_hip'}",
           f"-I{ROOT / 'mppi_metal_amd' / 'csrc'}", "-x", "hip",
           str(ROOT / "mppi_metal_amd" / "csrc" / "mppi_kernels.hip"), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True)
    return out.read_text().splitlines()


def _kernels(lines, pattern):
    starts = [(i, l.split(":")[0]) for i, l in enumerate(lines) if re.match(r"^mppi_\w+:", l)]
    for n, (i0, name) in enumerate(starts):
        if re.search(pattern, name):
            i1 = starts[n + 1][0] if n + 1 < len(starts) else len(lines)
            yield name, lines[i0:i1]


def _loops(body):
    """Innermost loop header of every line (None outside loops), from the
    compiler's block annotations."""
    inner, cur = [], None
    for j, l in enumerate(body):
        if l.startswith(".LBB"):
            ann = l + " " + (body[j + 1] if j + 1 < len(body) else "")
            if "Loop Header" in ann:
                cur = l.split(":")[0][1:]
            else:
                m = re.search(r"Header=(BB\w+)", l)
                cur = ("L" + m.group(1)) if m else None
        inner.append(cur)
    return inner


@pytest.mark.skipif(not shutil.which(HIPCC) and not Path(HIPCC).exists(), reason="hipcc not available")
def test_persistent_loops_are_uniform(tmp_path):
    lines = _assembly(tmp_path)
    seen = 0
    for name, body in _kernels(lines, r"_persist_b(64|128|256)$"):
        seen += 1
        inner = _loops(body)
        atom = next(j for j, l in enumerate(body) if "global_atomic_add_x2" in l)
        # the tile index is the first LDS read after the ticket fetch; in a
        # one-wavefront workgroup the compiler drops the barriers
        read = next(j for j in range(atom, len(body)) if "ds_read_b32" in body[j])
        assert inner[atom] is not None and inner[atom] == inner[read], \
            f"{name}: the ticket fetch ({inner[atom]}) and the tile index read ({inner[read]}) are in different loops"
        bars = [j for j in range(atom, len(body)) if "s_barrier" in body[j]]
        if bars:
            assert inner[atom] == inner[bars[0]], \
                f"{name}: the ticket fetch ({inner[atom]}) and the barrier after it ({inner[bars[0]]}) are in different loops"
        # every loop with a barrier: the block that jumps or falls back to its
        # header restores exec (s_or_b64 exec) and branches on SCC
        headers = {inner[j] for j, l in enumerate(body) if "s_barrier" in l}
        for h in headers:
            hj = next(j for j, l in enumerate(body) if l.startswith("." + h + ":"))
            latch_lines = [j for j in range(len(body)) if inner[j] == h and j < hj]
            latch_lines += [j for j in range(hj, len(body)) if re.search(r"branch\w*\s+\." + h + r"\s*$", body[j])]
            assert latch_lines or h == inner[atom], f"{name}: no latch found for {h}"
            if latch_lines and h != inner[atom]:
                tail = body[max(latch_lines) - 8: max(latch_lines) + 1]
                assert any("s_or_b64 exec, exec" in t for t in tail), f"{name}: {h} continues without reconverging"
                assert not any("s_cbranch_exec" in t for t in tail), f"{name}: {h} continues on an exec mask"
    assert seen == 9

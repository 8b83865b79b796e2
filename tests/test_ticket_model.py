"""CPU model of the persistent batch kernels' ticket protocol
(tests/cpp/ticket_model.cpp): host threads as workgroups, an atomic 64-bit
counter as the ticket counter.  Every tile of every launch runs exactly once,
each launch takes exactly n_tiles + grid tickets (so the next launch of the
same shape starts aligned without any reset), and from any counter value
every workgroup leaves."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_ticket_protocol_model(tmp_path):
    exe = tmp_path / "ticket_model"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", str(ROOT / "tests" / "cpp" / "ticket_model.cpp"),
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr

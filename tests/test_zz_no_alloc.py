"""mppi_hip_step_noalloc (the step without allocation) through the C ABI:
its results equal mppi_hip_step's bit for bit on the same inputs (the same
kernels run; only where their work buffers live differs), on two calls in a
row of which the second is smaller (a second context on the same device:
the shared scratch array holds the first call's contents), on both step
paths, on a caller's stream, and sizes past the scratch's limit are refused.
Device memory for inputs and outputs comes from torch (plumbing only)."""
import numpy as np
import pytest

from problems import unicycle

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to("cuda")


def _noalloc(hip, step, x0, u_nom, stream=0):
    import torch
    x_d, u_d = _dev(x0), _dev(u_nom)
    out = torch.empty(hip.H * hip.C, dtype=torch.float32, device="cuda")
    diag = torch.empty(3, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    hip.step_noalloc(step, x_d.data_ptr(), u_d.data_ptr(), out.data_ptr(), diag.data_ptr(), stream)
    if stream:
        torch.cuda.synchronize()
    return out.cpu().numpy().reshape(hip.H, hip.C), diag.cpu().numpy()


@pytest.mark.parametrize("path", [0, 1], ids=["auto", "two_kernel"])
def test_noalloc_equals_step(require_gpu, path):
    x0 = np.array([0.3, -0.2, 0.1, 0.5], np.float32)
    big, _, _ = unicycle(65536, 64, 16)
    small, _, _ = unicycle(1000, 20, 8)
    for hip, H in ((big, 64), (small, 20)):          # the second call is smaller
        hip.set_step_path(path)
        u_nom = (np.random.default_rng(H).standard_normal((H, 2)) * 0.3).astype(np.float32)
        u_ref, d_ref = hip.step(7, x0, u_nom)
        u_na, d_na = _noalloc(hip, 7, x0, u_nom)
        assert np.array_equal(u_na, u_ref) and np.array_equal(d_na, d_ref), f"H={H}"
    # again on the large context after the small one, on a caller's stream
    import torch
    s = torch.cuda.Stream()
    u_nom = (np.random.default_rng(1).standard_normal((64, 2)) * 0.3).astype(np.float32)
    u_ref, d_ref = big.step(9, x0, u_nom)
    u_na, d_na = _noalloc(big, 9, x0, u_nom, stream=s.cuda_stream)
    assert np.array_equal(u_na, u_ref) and np.array_equal(d_na, d_ref)


def test_noalloc_limits(require_gpu):
    from mppi_metal_amd import MppiHipError
    import torch
    hip, _, _ = unicycle(1 << 21, 64, 8)             # 8 MiB of costs alone
    x = torch.zeros(4, device="cuda")
    u = torch.zeros(64 * 2, device="cuda")
    with pytest.raises(MppiHipError, match="8 MiB"):
        hip.step_noalloc(0, x.data_ptr(), u.data_ptr(), u.data_ptr())
    with pytest.raises(MppiHipError, match="must not be null"):
        hip.step_noalloc(0, 0, u.data_ptr(), u.data_ptr())
    # the allocating entry point's read-back refers to its own last call
    small, _, _ = unicycle(512, 16, 8)
    small.step(0, np.zeros(4, np.float32), np.zeros((16, 2), np.float32))
    xs, us = torch.zeros(4, device="cuda"), torch.zeros(32, device="cuda")
    out = torch.empty(32, device="cuda")
    small.step_noalloc(0, xs.data_ptr(), us.data_ptr(), out.data_ptr())
    with pytest.raises(MppiHipError, match="no iteration"):
        small.read_costs()

"""icam_hip_query_device_noalloc against icam_hip_query_device, through the C ABI.

The allocation-free entry point takes its scratch from the library's static device
array; its answers must be the same as the entry point whose kernels it shares, on two
calls in a row (the second smaller, so stale scratch from the first would show), for a
part of closed-form faces and a part with a B-spline face, in both precisions.
"""
import numpy as np
import pytest

from implicitcam_amd import FP64, MIXED, Part, PartGpu, uniform_points

pytestmark = pytest.mark.gpu
SEED = 0x2545F4914F6CDD1D


def _run(g, fn, pts, prec, feat):
    import torch

    n = len(pts)
    x = torch.from_numpy(np.ascontiguousarray(pts)).cuda()
    d = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
    fl = torch.full((n,), 7, dtype=torch.int32, device="cuda")
    ft = torch.full((n,), -1, dtype=torch.int32, device="cuda") if feat else None
    torch.cuda.synchronize()
    fn(x.data_ptr(), n, d.data_ptr(), fl.data_ptr(), ft.data_ptr() if feat else None, prec, None)
    g.query_stats()  # waits for the device
    return d.cpu().numpy(), fl.cpu().numpy().view(np.uint32), (ft.cpu().numpy().view(np.uint32) if feat else None)


@pytest.mark.parametrize("name,need", [("c1", 12 * 1.6), ("c2", 12 * 0.5)])
@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_noalloc_matches_query_device(name, need, prec):
    part = Part(name)
    g = PartGpu(part, need)
    tol = 0.0 if prec == FP64 else 1e-6 * part.diagonal
    for call, (n, feat) in enumerate(((1 << 17, True), (5000, False))):
        pts = uniform_points(SEED + call, n, part.lo - 4, part.hi + 4)
        d0, f0, t0 = _run(g, g.query_device, pts, prec, feat)
        d1, f1, t1 = _run(g, g.query_device_noalloc, pts, prec, feat)
        assert (f1 == f0).all(), (name, call)
        assert (f1 == 0).all(), (name, call)
        assert np.abs(d1 - d0).max() <= tol, (name, call, np.abs(d1 - d0).max())
        assert ((d1 < 0) == (d0 < 0)).all(), (name, call)
        if feat:
            # the nearest feature: the same one, but for exact ties between features
            assert (t1 == t0).mean() >= 0.999, (name, call)


def test_noalloc_refuses_past_its_maximum():
    part = Part("c1")
    g = PartGpu(part, 12 * 1.6)
    with pytest.raises(RuntimeError, match="at most"):
        g.query_device_noalloc(0, (1 << 18) + 1, 0, 0, None, FP64, None)

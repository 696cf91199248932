"""icam_hip_cast (the viewer's ray cast over baked brick levels) on an MI355X against its
CPU restatement (oracle/cast.py, pinned in tests/test_cast.py) on the same maps and
atlases: one level of the box with a hole and of the NURBS-top block, and two nested
levels. Both run in single precision without contraction: hits and misses agree, t to
1e-5 of a cell, gradients to 1e-4 of their length, levels and slots exactly."""
import numpy as np
import pytest

from implicitcam_amd import FP64, Part, PartGpu, cast_device, level_for
from oracle.cast import cast
from test_cast import level_tuple, rays_at

pytestmark = pytest.mark.gpu


def gpu_cast(levels, maps, atlases, rays):
    import torch

    n = len(rays)
    d_rays = torch.from_numpy(np.ascontiguousarray(rays)).cuda()
    dm = [torch.from_numpy(m.view(np.int32)).cuda() for m in maps]
    da = [torch.from_numpy(a.view(np.int32)).cuda() for a in atlases]
    out = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    hit = torch.empty((n, 2), dtype=torch.int32, device="cuda")
    cast_device(levels, [x.data_ptr() for x in dm], [x.data_ptr() for x in da], d_rays.data_ptr(), n, out.data_ptr(), hit.data_ptr())
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    hh = hit.cpu().numpy().view(np.uint32)
    return o[:, 0], o[:, 1:], hh[:, 0], hh[:, 1]


def agree(a, b, h):
    t, g, k, s = a
    tr, gr, kr, sr = b
    assert (t >= 0).sum() > 100 and (t < 0).sum() > 50
    assert ((t >= 0) == (tr >= 0)).all()
    assert (k == kr).all() and (s == sr).all()
    both = t >= 0
    assert np.abs(t[both] - tr[both]).max() <= 1e-5 * h
    gl = np.linalg.norm(gr[both], axis=1)
    assert (np.linalg.norm(g[both] - gr[both], axis=1) <= 1e-4 * np.maximum(gl, 1e-6)).all()


@pytest.mark.parametrize("name,h", [("c1", 1.6), ("c2", 1.0)])
def test_cast_one_level(name, h):
    part = Part(name)
    g = PartGpu(part, 12 * h)
    lv = level_for(part, h)
    m, a, _ = g.bake(lv, FP64)
    rays = rays_at(part, 4096, 11, margin=0.2 * part.diagonal)
    agree(gpu_cast([lv], [m], [a], rays), cast([level_tuple(lv, m, a)], rays), h)
    g.close()


def test_cast_nested_levels():
    part = Part("c1")
    g = PartGpu(part, 12 * 3.2)
    coarse = level_for(part, 3.2)
    fine = level_for(part, 1.6)
    fine.lo[0] = 0.5 * (part.lo[0] + part.hi[0])
    fine.nb[0] = int(np.ceil((part.hi[0] + 1.6 - fine.lo[0]) / (8 * 1.6)))
    mf, af, _ = g.bake(fine, FP64)
    mc, ac, _ = g.bake(coarse, FP64)
    rays = rays_at(part, 4096, 13, margin=10.0)
    res = gpu_cast([fine, coarse], [mf, mc], [af, ac], rays)
    agree(res, cast([level_tuple(fine, mf, af), level_tuple(coarse, mc, ac)], rays), 1.6)
    assert (res[2][res[0] >= 0] == 0).sum() > 100 and (res[2][res[0] >= 0] == 1).sum() > 100
    g.close()

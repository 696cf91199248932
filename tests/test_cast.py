"""The ray cast over baked brick levels (the viewer's, crates/icam-sim/src/cache.wgsl:191-503):
its CPU restatement (oracle/cast.py), the checker of icam_hip_cast, pinned against the
part's exact signed distance (the FP64 oracle) -- on planes the cache's trilinear
interpolant is exact up to the i16 quantum, so rays meet the box's flat faces where the
planes are; everywhere a hit lies within the cell's interpolation error of the surface, a
ray that meets nothing never enters the solid, and a ray crossing nested levels is met in
the finest level holding the point."""
import numpy as np

from implicitcam_amd import Part, level_for
from oracle import Oracle
from oracle.cast import AIR, cast


def rays_at(part, n, seed, margin=0.0):
    """Rays from a sphere around the part aimed at points of its box."""
    rng = np.random.default_rng(seed)
    c = 0.5 * (part.lo + part.hi)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    o = c + 0.75 * part.diagonal * u
    tgt = part.lo - margin + rng.random((n, 3)) * (part.hi - part.lo + 2 * margin)
    d = tgt - o
    d /= np.linalg.norm(d, axis=1)[:, None]
    return np.concatenate([o, d], axis=1)


def level_tuple(lv, m, a):
    return (np.array(lv.lo[:]), lv.h, np.array(lv.nb[:]), m, a)


def test_cast_meets_planes_and_the_surface():
    part = Part("c1")
    o = Oracle(part)
    lv = level_for(part, 1.6)
    m, a, _ = o.bake(lv)
    rays = rays_at(part, 500, 3, margin=25.0)
    t, g, k, slot = cast([level_tuple(lv, m, a)], rays)
    hit = t >= 0
    assert 0.3 < hit.mean() < 0.9
    assert (k[hit] == 0).all() and (k[~hit] == 0xFFFFFFFF).all() and (slot[~hit] == AIR).all()
    p = rays[hit, :3] + t[hit, None].astype(np.float64) * rays[hit, 3:]
    d, _, _ = o.sdf(p)
    # within the interpolation error of the cells: the interpolant rounds a convex edge of
    # the box off (up to 0.28 h), and is exact to rounding on the faces' planes
    assert np.abs(d).max() <= 0.3 * lv.h
    assert np.median(np.abs(d)) <= 1e-4 * lv.h
    # the top face away from its edges (box sides, the hole's fillet): exact up to rounding
    c = 0.5 * (part.lo + part.hi)
    top = (np.abs(p[:, 2] - part.hi[2]) < 0.5 * lv.h) & (np.abs(p[:, 0] - c[0]) < 50 - 2 * lv.h) & (np.abs(p[:, 1] - c[1]) < 30 - 2 * lv.h) & \
          (np.hypot(p[:, 0] - c[0], p[:, 1] - c[1]) > 13 + 2 * lv.h)
    assert top.sum() > 10
    assert np.abs(p[top, 2] - part.hi[2]).max() <= 1e-3 * lv.h
    # the gradient there points up
    gt = g[hit][top]
    assert (gt[:, 2] > 0.99 * np.linalg.norm(gt, axis=1)).all()
    # rays that meet nothing never enter the solid
    miss = np.nonzero(~hit)[0][:60]
    for i in miss:
        s = np.linspace(0.0, 1.5 * part.diagonal, 800)
        q = rays[i, :3] + s[:, None] * rays[i, 3:]
        dq, _, _ = o.sdf(q)
        assert dq.min() > -0.3 * lv.h


def test_cast_reads_the_finest_level_holding_the_ray():
    """A coarse level over the whole part and a fine one over its +x half: hits in the fine
    level's box are reported by it, within its cells' error, the others by the coarse one."""
    part = Part("c1")
    o = Oracle(part)
    coarse = level_for(part, 3.2)
    fine = level_for(part, 1.6)
    fine.lo[0] = 0.5 * (part.lo[0] + part.hi[0])
    fine.nb[0] = int(np.ceil((part.hi[0] + 1.6 - fine.lo[0]) / (8 * 1.6)))
    mc, ac, _ = o.bake(coarse)
    mf, af, _ = o.bake(fine)
    rays = rays_at(part, 400, 5, margin=10.0)
    t, g, k, slot = cast([level_tuple(fine, mf, af), level_tuple(coarse, mc, ac)], rays)
    hit = t >= 0
    p = rays[hit, :3] + t[hit, None].astype(np.float64) * rays[hit, 3:]
    d, _, _ = o.sdf(p)
    hs = np.where(k[hit] == 0, fine.h, coarse.h)
    assert (np.abs(d) <= 0.3 * hs).all()
    fhi = np.array(fine.lo[:]) + 8 * fine.h * np.array(fine.nb[:])
    in_fine = (p >= np.array(fine.lo[:]) + 1e-3).all(1) & (p <= fhi - 1e-3).all(1)
    assert (k[hit][in_fine] == 0).all() and in_fine.sum() > 20
    assert (k[hit] == 1).sum() > 20

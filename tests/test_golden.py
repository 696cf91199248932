"""The golden vectors (tests/golden/oracle_vectors.npz, written by
tests/golden/make_golden.py from the FP64 oracle), pinned independently of the oracle,
and the oracle held to them.

- can and c1: every stored distance against the closed forms (the reference's capped
  cylinder, crates/icam-core/src/solid3.rs:2049-2057, and c1's per-face forms);
- bump box: the sign against the surface height, |d| against dense samples
  (solid3.rs:2145-2210);
- c2: |d| no larger than the nearest of dense surface samples;
- the c1 level: AIR / SOLID bricks by the closed form's sign at their centres, every
  stored node within one quantum of the closed form quantized as
  crates/icam-sim/src/partcache.rs:911-915 does (clamp h sqrt 3, 32767 steps), and the
  two-per-word packing low half first (partcache.rs:962-970).

The GPU side reads the same vectors in tests/test_gpu.py.
"""
import os

import numpy as np
import pytest
from scipy.spatial import cKDTree

from closed_forms import bump_samples, bump_top, c1_sdf, can_sdf
from implicitcam_amd.part import Part
from oracle import Oracle

G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_vectors.npz"))


def test_can_vectors_closed_form():
    assert np.abs(G["can_d"] - can_sdf(G["can_xyz"])).max() < 1e-9


def test_c1_vectors_closed_form():
    ref = c1_sdf(G["c1_xyz"])
    assert np.abs(G["c1_d"] - ref).max() < 1e-8
    assert ((G["c1_d"] < 0) == (ref < 0))[np.abs(ref) > 1e-8].all()


def test_bump_vectors_dense():
    pts, d = G["bump_xyz"], G["bump_d"]
    samples, h = bump_samples(300)
    dense, _ = cKDTree(samples).query(pts)
    inside = (pts[:, 0] > 0) & (pts[:, 0] < 10) & (pts[:, 1] > 0) & (pts[:, 1] < 10) & (pts[:, 2] > 0) & (pts[:, 2] < bump_top(pts[:, 0], pts[:, 1]))
    assert ((d < 0) == inside).all()
    assert (np.abs(d) <= dense + 1e-9).all() and (np.abs(d) >= dense - h).all()


def test_c2_vectors_dense(built):
    part = Part("c2")
    o = Oracle(part)
    t = np.linspace(0, 1, 181)
    top = np.array([o.nurbs_eval(0, a, b) for a in t for b in t])
    d = G["c2_d"]
    dense, _ = cKDTree(top).query(G["c2_xyz"])
    # the top is one face of several: no point is farther from the solid than from it
    assert (np.abs(d) <= dense + 1e-9).all()


def test_c1_level_rules():
    lo, h, nb = G["c1_level"][:3], G["c1_level"][3], G["c1_level"][4:].astype(int)
    m, atlas = G["c1_map"], G["c1_atlas"]
    ids = np.arange(len(m))
    idx = np.stack([ids % nb[0], (ids // nb[0]) % nb[1], ids // (nb[0] * nb[1])], 1)
    first = lo + 8 * h * idx
    dc = c1_sdf(first + 4 * h)
    assert (dc[m == 0xFFFFFFFF] > 0).all() and (dc[m == 0xFFFFFFFE] < 0).all()
    stored = np.flatnonzero((m != 0xFFFFFFFF) & (m != 0xFFFFFFFE))
    assert (m[stored] == np.arange(len(stored))).all() and len(atlas) == 365 * len(stored)
    n = np.arange(729)
    off = h * np.stack([n % 9, (n // 9) % 9, n // 81], 1)
    clamp = h * np.sqrt(3)
    for s, b in enumerate(stored):
        w = atlas[365 * s: 365 * (s + 1)]
        q = np.stack([w & 0xFFFF, w >> 16], 1).ravel()[:729].astype(np.uint16).view(np.int16).astype(int)
        ref = np.round(np.clip(c1_sdf(first[b] + off) / clamp, -1, 1) * 32767).astype(int)
        assert np.abs(q - ref).max() <= 1
        assert (w[-1] >> 16) == 0  # the 730th half of the last word is zero


@pytest.mark.parametrize("name", ["can", "c1", "bump", "c2"])
def test_oracle_reproduces_vectors(built, name):
    o = Oracle(Part(name))
    d, feat, _ = o.sdf(G[f"{name}_xyz"])
    assert (d == G[f"{name}_d"]).all()
    assert (feat == G[f"{name}_feat"]).all()


def test_oracle_reproduces_level(built):
    from implicitcam_amd import _native

    lv = _native.Level()
    lv.lo[:] = list(G["c1_level"][:3])
    lv.h = float(G["c1_level"][3])
    lv.nb[:] = [int(x) for x in G["c1_level"][4:]]
    m, a, _ = Oracle(Part("c1")).bake(lv)
    assert (m == G["c1_map"]).all() and (a == G["c1_atlas"]).all()

"""Writes tests/golden/oracle_vectors.npz: seeded query points with their signed
distances and nearest features, and one small brick level (map and atlas), from the
FP64 oracle.

The reference ships no vectors for this path (its fixtures are code: the capped
cylinder and the B-spline box of crates/icam-core/src/solid3.rs:2013-2143), so these
are the oracle's own, and tests/test_golden.py pins them independently: the can's
and c1's values against their closed forms, the bump box's against dense sampling,
the brick level against brick rules restated in numpy.

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from implicitcam_amd.part import Part, level_for, uniform_points  # noqa: E402
from oracle import Oracle  # noqa: E402

SEED = 0x2545F4914F6CDD1D
# (part, points, margin around the part's box in mm)
CASES = [("can", 512, 2.0), ("c1", 1024, 5.0), ("bump", 512, 3.0), ("c2", 1024, 3.0)]


def main():
    out = {}
    for name, n, margin in CASES:
        part = Part(name)
        o = Oracle(part)
        pts = uniform_points(SEED ^ len(name), n, part.lo - margin, part.hi + margin)
        d, feat, flags = o.sdf(pts, threads=1)
        assert not (flags & 1).any(), f"{name}: a capped search"
        out[f"{name}_xyz"] = pts
        out[f"{name}_d"] = d
        out[f"{name}_feat"] = feat
    part = Part("c1")
    lv = level_for(part, 3.2, 4)
    m, a, _ = Oracle(part).bake(lv, threads=1)
    out["c1_level"] = np.array([*lv.lo[:], lv.h, *lv.nb[:]])
    out["c1_map"] = m
    out["c1_atlas"] = a
    path = os.path.join(HERE, "oracle_vectors.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()

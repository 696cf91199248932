"""Synthetic parts (the evaluator's input, include/icam_part.h) and seeded inputs.

Parts are made by the native generator (csrc/gen/partgen.cpp):

- ``can``: the reference's solid cylinder r 3, z 0..5 (crates/icam-core/src/solid3.rs:2013-2040);
- ``bump``: its 10 x 10 box with a biquadratic B-spline top (solid3.rs:2094-2143);
- ``c1``: box 100 x 60 x 40 minus a through hole r 10 with a torus fillet r 3;
- ``c2``: 120 x 120 block whose top is a rational bicubic NURBS (16 x 16 net);
- ``turned:<type>:<grooves>``: solids of revolution (planes, cylinders, cones with
  their apex on the axis, spheres with their poles on it, ring tori);
- ``plate:<k>``: a plate with k holes, plain, filleted (torus) or countersunk (cone);
- ``nurbs:<n>``: a block with an n x n rational bicubic top;
- ``c4:<i>``: part i of the 64-part batched workload (face counts log-uniform 6-400).
"""
import ctypes

import numpy as np

from . import _native


class Part:
    """An owned synthetic part; `.desc` is the `icam_part_desc*` the libraries take."""

    def __init__(self, name, seed=0x2545F4914F6CDD1D):
        self.name = name
        self.seed = seed
        self._lib = _native.gen()
        self.desc = self._lib.icam_gen_part(name.encode(), seed)
        if not self.desc:
            raise ValueError(f"unknown part {name!r}")
        d = self.desc.contents
        self.lo = np.array(d.lo[:])
        self.hi = np.array(d.hi[:])
        self.n_faces, self.n_edges, self.n_verts = d.n_faces, d.n_edges, d.n_verts

    def describe(self):
        """What the part was made from (the generator's icam_gen_describe), as float64."""
        n = self._lib.icam_gen_describe(self.desc, None, 0)
        out = np.empty(max(n, 1))
        self._lib.icam_gen_describe(self.desc, out.ctypes.data_as(_native.c_double_p), n)
        return out[:n]

    @property
    def diagonal(self):
        return float(np.linalg.norm(self.hi - self.lo))

    @property
    def centre(self):
        return 0.5 * (self.lo + self.hi)

    def close(self):
        if self.desc:
            self._lib.icam_gen_free(self.desc)
            self.desc = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def uniform_points(seed, n, lo, hi):
    """n points uniform in the box [lo, hi] from the reference checks' xorshift64
    (crates/icam-check/src/exact_check.rs:46-53), float64 (n, 3)."""
    out = np.empty((n, 3), dtype=np.float64)
    lo = np.ascontiguousarray(lo, dtype=np.float64)
    hi = np.ascontiguousarray(hi, dtype=np.float64)
    _native.gen().icam_gen_uniform(seed, n, lo.ctypes.data_as(_native.c_double_p), hi.ctypes.data_as(_native.c_double_p),
                                   out.ctypes.data_as(_native.c_double_p))
    return out


def near_trim_points(part, seed, n, radius):
    """n points within `radius` (mm) of the trim curves of the part's planar faces,
    float64 (n, 3), from the reference checks' xorshift64."""
    out = np.empty((n, 3), dtype=np.float64)
    part._lib.icam_gen_near_trim(part.desc, seed, n, radius, out.ctypes.data_as(_native.c_double_p))
    return out


def near_surface_points(part, seed, n, radius):
    """n points within `radius` (mm) of the part's surface (a face drawn by area, a point
    of it inside its loops, an offset uniform in the ball), float64 (n, 3), from the
    reference checks' xorshift64."""
    out = np.empty((n, 3), dtype=np.float64)
    part._lib.icam_gen_near_surface(part.desc, seed, n, radius, out.ctypes.data_as(_native.c_double_p))
    return out


def level_for(part, h, max_bricks=None):
    """The brick level of cell side h over the part, placed as the reference's cache
    places a level (crates/icam-sweep/src/cache.rs:505-521): first node a cell below the
    box's low corner, enough bricks to pass its high corner, at most max_bricks a side."""
    lo = part.lo - h
    nb = []
    for k in range(3):
        full = (int(np.ceil((part.hi[k] - part.lo[k] + 2 * h) / h)) + 7) // 8
        nb.append(min(full, max_bricks) if max_bricks else full)
    lv = _native.Level()
    lv.lo[:] = list(lo)
    lv.h = h
    lv.nb[:] = nb
    return lv


def c4_level(part, grid):
    """The batched workload's level for a part: a cubic grid of `grid` nodes a side over
    the part's largest extent and a cell of margin each way (part-scaled spacing),
    grid / 8 bricks along the largest axis."""
    h = float(np.max(part.hi - part.lo) + 2.0) / (grid - 2)
    return level_for(part, h, grid // 8)


def c4_grid(index, seed=0x2545F4914F6CDD1D):
    """Part `index`'s grid in the batched workload: 32, 64, 128 or 256 nodes a side,
    log-uniform from the reference checks' xorshift64."""
    x = (seed ^ (0xD1B54A32D192ED03 * (index + 1))) & 0xFFFFFFFFFFFFFFFF or 1
    for _ in range(3):
        x ^= (x << 13) & 0xFFFFFFFFFFFFFFFF
        x ^= x >> 7
        x ^= (x << 17) & 0xFFFFFFFFFFFFFFFF
    return 32 << ((x >> 11) % 4)

"""Closed-form signed distances that pin the oracle independently of it.

- `can`: the reference's own capped-cylinder formula (crates/icam-core/src/solid3.rs:2049-2057).
- `c1`: the box with a hole and a torus fillet, as the minimum over its faces of each
  trimmed face's distance in closed form (a planar region's in-plane distance and
  height, a cylinder band's and a quarter tube's distance in the meridian half-plane),
  signed by a constructive-solid inside test. No feature trees, no trim loops.
"""
import numpy as np


def can_sdf(p):
    p = np.asarray(p, dtype=np.float64)
    rho = np.hypot(p[:, 0], p[:, 1])
    dr, dz = rho - 3.0, np.abs(p[:, 2] - 2.5) - 2.5
    inside = (dr <= 0) & (dz <= 0)
    return np.where(inside, np.maximum(dr, dz), np.hypot(np.maximum(dr, 0), np.maximum(dz, 0)))


def c1_sdf(p, ax=50.0, ay=30.0, az=20.0, rh=10.0, rf=3.0):
    p = np.asarray(p, dtype=np.float64)
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    rho = np.hypot(x, y)
    R, zc = rh + rf, az - rf

    def rect(u, v, hu, hv):
        return np.hypot(np.maximum(np.abs(u) - hu, 0), np.maximum(np.abs(v) - hv, 0))

    def rect_minus_disc(u, v, hu, hv, r):
        rr = np.hypot(u, v)
        inside = (np.abs(u) <= hu) & (np.abs(v) <= hv)
        return np.where(inside & (rr < r), r - rr, rect(u, v, hu, hv))

    ds = [
        np.hypot(rect_minus_disc(x, y, ax, ay, R), z - az),
        np.hypot(rect_minus_disc(x, y, ax, ay, rh), z + az),
        np.hypot(rect(y, z, ay, az), x - ax),
        np.hypot(rect(y, z, ay, az), x + ax),
        np.hypot(rect(x, z, ax, az), y - ay),
        np.hypot(rect(x, z, ax, az), y + ay),
        np.hypot(rho - rh, np.maximum(np.maximum(-az - z, z - zc), 0)),
    ]
    psi = np.arctan2(z - zc, rho - R)
    on_arc = (psi >= np.pi / 2) & (psi <= np.pi)
    tube = np.abs(np.hypot(rho - R, z - zc) - rf)
    ends = np.minimum(np.hypot(rho - R, z - az), np.hypot(rho - rh, z - zc))
    ds.append(np.where(on_arc, tube, ends))
    d = np.min(np.array(ds), axis=0)
    in_box = (np.abs(x) < ax) & (np.abs(y) < ay) & (np.abs(z) < az)
    removed = (rho < rh) | ((rho < R) & (z > zc) & (np.hypot(rho - R, z - zc) > rf))
    return np.where(in_box & ~removed, -d, d)


def bump_top(x, y):
    """The `bump_box` top z = S(x, y) (solid3.rs:2094-2101)."""
    zc = np.array([[5.0, 5.0, 5.0], [7.0, 9.0, 5.0], [5.0, 6.0, 5.0]])
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)

    def b(t):
        return np.array([(1 - t) ** 2, 2 * t * (1 - t), t * t])

    return np.einsum("i...,ij,j...->...", b(x / 10), zc, b(y / 10))


def bump_samples(n=300):
    """Dense samples of every face of `bump_box` (solid3.rs:2150-2160)."""
    a = np.arange(n + 1) / n * 10
    FX, FY = np.meshgrid(a, a, indexing="ij")
    s = [np.stack([FX, FY, bump_top(FX, FY)], -1).reshape(-1, 3), np.stack([FX, FY, 0 * FX], -1).reshape(-1, 3)]
    for px, py in [(FX, 0 * FX), (FX, 0 * FX + 10), (0 * FX, FX), (0 * FX + 10, FX)]:
        s.append(np.stack([px, py, bump_top(px, py) * FY / 10], -1).reshape(-1, 3))
    return np.concatenate(s), 10.0 / n


def reference_points(k, cx=0.0, cy=0.0, cz=2.5, r=6.0, rz=5.0):
    """The reference tests' deterministic points (solid3.rs:2060-2065, 2167)."""
    t = np.arange(k, dtype=np.float64)
    return np.stack([cx + r * np.sin(t * 0.37), cy + r * np.cos(t * 0.71), cz + rz * np.sin(t * 1.3)], 1)

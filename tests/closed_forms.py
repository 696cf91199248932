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


# ---------------------------------------------------------------- solids of revolution
#
# A profile is a closed counterclockwise loop in the (rho, z) half-plane of segments
# and circular arcs (rows of the generator's description: arc, r0, z0, r1, z1, rc, zc,
# rad, psi0, psi1, face). Revolved about an axis, its faces are the elements marked
# `face`; the nearest point of a surface of revolution to p lies in p's meridian
# half-plane, so the distance is the planar distance from (rho, z) to those elements,
# and the sign is the planar inside test of the whole loop.


def _elements(info, at):
    n = int(info[at])
    return info[at + 1: at + 1 + 11 * n].reshape(n, 11), at + 1 + 11 * n


def _seg_dist(r, z, a, b):
    ab = b - a
    t = np.clip(((r - a[0]) * ab[0] + (z - a[1]) * ab[1]) / max(ab @ ab, 1e-300), 0, 1)
    return np.hypot(r - a[0] - t * ab[0], z - a[1] - t * ab[1])


def _arc_dist(r, z, e):
    rc, zc, rad, p0, p1 = e[5], e[6], e[7], e[8], e[9]
    lo, hi = min(p0, p1), max(p0, p1)
    phi = np.arctan2(z - zc, r - rc)
    phi = lo + np.mod(phi - lo, 2 * np.pi)
    on = phi <= hi
    radial = np.abs(np.hypot(r - rc, z - zc) - rad)
    ends = np.minimum(np.hypot(r - e[1], z - e[2]), np.hypot(r - e[3], z - e[4]))
    return np.where(on, radial, ends)


def profile_dist(r, z, els):
    d = np.full(np.shape(r), np.inf)
    for e in els:
        if e[10] == 0:
            continue
        if e[0]:
            d = np.minimum(d, _arc_dist(r, z, e))
        else:
            d = np.minimum(d, _seg_dist(r, z, np.array([e[1], e[2]]), np.array([e[3], e[4]])))
    return d


def profile_inside(r, z, els):
    """Even-odd test against the loop with its arcs replaced by chords, flipped inside
    each circular segment between an arc and its chord (at most a half circle)."""
    inside = np.zeros(np.shape(r), dtype=bool)
    for e in els:
        a, b = (e[1], e[2]), (e[3], e[4])
        crosses = ((a[1] > z) != (b[1] > z))
        with np.errstate(divide="ignore", invalid="ignore"):
            x = a[0] + (z - a[1]) * (b[0] - a[0]) / (b[1] - a[1])
        inside ^= crosses & (r < x)
        if e[0]:
            rc, zc, rad = e[5], e[6], e[7]
            pm = 0.5 * (e[8] + e[9])
            mid = (rc + rad * np.cos(pm), zc + rad * np.sin(pm))
            side = lambda pr, pz: (b[0] - a[0]) * (pz - a[1]) - (b[1] - a[1]) * (pr - a[0])  # noqa: E731
            seg = (np.hypot(r - rc, z - zc) < rad) & (np.sign(side(r, z)) == np.sign(side(*mid)))
            inside ^= seg
    return inside


def turned_sdf(p, info):
    els, _ = _elements(info, 1)
    p = np.asarray(p, dtype=np.float64)
    r, z = np.hypot(p[:, 0], p[:, 1]), p[:, 2]
    d = profile_dist(r, z, els)
    return np.where(profile_inside(r, z, els), -d, d)


def plate_sdf(p, info):
    """A plate with holes: the least of its faces' distances (top and bottom planar
    regions with the holes' discs removed, the four sides, each hole's surfaces of
    revolution), negative inside the box and outside every hole's region."""
    W, D, T, k = info[1], info[2], info[3], int(info[4])
    p = np.asarray(p, dtype=np.float64)
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    ax, ay, az = W / 2, D / 2, T / 2
    at = 5
    holes = []
    for _ in range(k):
        hx, hy = info[at], info[at + 1]
        els, at = _elements(info, at + 2)
        holes.append((hx, hy, els))

    def rect(u, v, hu, hv):
        return np.hypot(np.maximum(np.abs(u) - hu, 0), np.maximum(np.abs(v) - hv, 0))

    def region(level):
        # in-plane distance from (x, y) to the rectangle minus the holes' discs where
        # the holes meet the plane z = level
        d = rect(x, y, ax, ay)
        for hx, hy, els in holes:
            rr = max(max(e[1], e[3]) for e in els if abs(e[2] - level) < 1e-12 and abs(e[4] - level) < 1e-12)
            rho = np.hypot(x - hx, y - hy)
            d = np.where(rho < rr, rr - rho, d)
        return d

    ds = [np.hypot(region(az), z - az), np.hypot(region(-az), z + az), np.hypot(rect(y, z, ay, az), x - ax),
          np.hypot(rect(y, z, ay, az), x + ax), np.hypot(rect(x, z, ax, az), y - ay), np.hypot(rect(x, z, ax, az), y + ay)]
    removed = np.zeros(len(p), dtype=bool)
    for hx, hy, els in holes:
        rho = np.hypot(x - hx, y - hy)
        ds.append(profile_dist(rho, z, els))
        removed |= profile_inside(rho, z, els)
    d = np.min(np.array(ds), axis=0)
    in_box = (np.abs(x) < ax) & (np.abs(y) < ay) & (np.abs(z) < az)
    return np.where(in_box & ~removed, -d, d)


def described_sdf(part, p):
    """The closed form of a generated turned part or plate (None for other parts)."""
    info = part.describe()
    if len(info) == 0:
        return None
    return turned_sdf(p, info) if info[0] == 1 else plate_sdf(p, info)


# ---------------------------------------------------------------- the trimmed plate (C3)
#
# Its surfaces are planes and vertical walls over closed outlines (curves of the
# plane), so every distance is a function of the planar distance to an outline and of
# heights: the nearest point of a wall over outline C between heights z0 and z1 is
# (c, clamp(z, z0, z1)) with c the outline's nearest point to (x, y). Outlines are
# projected exactly: nearest of dense samples, then Newton on that cubic piece and its
# neighbours.

_KN = np.array([0, 0, 0, 0, 0.2, 0.4, 0.6, 0.8, 1, 1, 1, 1.0])


def _bezier_pieces_of_segment(ctrl):
    """The 5 cubic Bezier pieces (4 x 2 each) of a clamped cubic B-spline with 8 control
    points and knots 0 0 0 0 .2 .4 .6 .8 1 1 1 1 (knot insertion to multiplicity 3)."""
    kn = list(_KN)
    P = [np.array(c, dtype=np.float64) for c in ctrl]
    for u in (0.2, 0.4, 0.6, 0.8):
        while kn.count(u) < 3:
            k = max(i for i in range(len(kn) - 1) if kn[i] <= u < kn[i + 1])
            s = kn.count(u)
            Q = P[: k - 3 + 1]
            for i in range(k - 3 + 1, k - s + 1):
                a = (u - kn[i]) / (kn[i + 3] - kn[i])
                Q.append(a * P[i] + (1 - a) * P[i - 1])
            Q += P[k - s:]
            kn.insert(k + 1, u)
            P = Q
    return [np.array(P[3 * m: 3 * m + 4]) for m in range(5)]


class Outline:
    """A closed outline of F five-span segments (control points from the description)."""

    def __init__(self, segs):
        self.pieces = np.array([pc for sg in segs for pc in _bezier_pieces_of_segment(sg)])  # (n, 4, 2)
        t = np.linspace(0, 1, 65)[:-1]
        B = np.stack([(1 - t) ** 3, 3 * t * (1 - t) ** 2, 3 * t * t * (1 - t), t ** 3], 1)
        self.samples = np.einsum("tk,nkd->ntd", B, self.pieces).reshape(-1, 2)
        self.sample_piece = np.repeat(np.arange(len(self.pieces)), len(t))
        self.sample_t = np.tile(t, len(self.pieces))
        from scipy.spatial import cKDTree

        self.tree = cKDTree(self.samples)

    def _eval(self, pc, t):
        s = 1 - t
        b = np.stack([s ** 3, 3 * t * s * s, 3 * t * t * s, t ** 3], -1)
        d = np.stack([-3 * s * s, 3 * s * s - 6 * t * s, 6 * t * s - 3 * t * t, 3 * t * t], -1)
        e = np.stack([6 * s, -12 * s + 6 * t, 6 * s - 12 * t, 6 * t], -1)
        return np.einsum("nk,nkd->nd", b, pc), np.einsum("nk,nkd->nd", d, pc), np.einsum("nk,nkd->nd", e, pc)

    def project(self, xy):
        """Distance to the outline, nearest point and whether xy is inside (left of the
        counterclockwise outline)."""
        _, k = self.tree.query(xy)
        best_d = np.full(len(xy), np.inf)
        best_q = np.zeros_like(xy)
        best_tan = np.zeros_like(xy)
        n = len(self.pieces)
        for off in (-1, 0, 1):
            pid = (self.sample_piece[k] + off) % n
            t = np.clip(self.sample_t[k] if off == 0 else np.full(len(xy), 0.0 if off == 1 else 1.0), 0, 1)
            pc = self.pieces[pid]
            for _ in range(30):
                q, d1, d2 = self._eval(pc, t)
                r = q - xy
                g = np.sum(r * d1, 1)
                h = np.sum(d1 * d1, 1) + np.sum(r * d2, 1)
                h = np.where(h > 0, h, np.sum(d1 * d1, 1))
                t = np.clip(t - g / h, 0, 1)
            q, d1, _ = self._eval(pc, t)
            d = np.hypot(*(q - xy).T)
            better = d < best_d
            best_d, best_q[better], best_tan[better] = np.where(better, d, best_d), q[better], d1[better]
        cross = best_tan[:, 0] * (xy[:, 1] - best_q[:, 1]) - best_tan[:, 1] * (xy[:, 0] - best_q[:, 0])
        return best_d, best_q, cross > 0


def c3_sdf(p, info):
    W, D, T, nl = info[1], info[2], info[3], int(info[4])
    ax, ay, az = W / 2, D / 2, T / 2
    p = np.asarray(p, dtype=np.float64)
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    at = 5
    loops = []
    for _ in range(nl):
        kind, z0, z1, F = info[at], info[at + 1], info[at + 2], int(info[at + 3])
        segs = info[at + 4: at + 4 + 16 * F].reshape(F, 8, 2)
        at += 4 + 16 * F
        loops.append((int(kind), z0, z1, Outline(segs)))

    def rect(u, v, hu, hv):
        return np.hypot(np.maximum(np.abs(u) - hu, 0), np.maximum(np.abs(v) - hv, 0))

    xy = p[:, :2]
    top_plane = rect(x, y, ax, ay)  # in-plane distance to the top face's region
    ds = [np.hypot(rect(x, y, ax, ay), z + az), np.hypot(rect(y, z, ay, az), x - ax), np.hypot(rect(y, z, ay, az), x + ax),
          np.hypot(rect(x, z, ax, az), y - ay), np.hypot(rect(x, z, ax, az), y + ay)]
    inside_box = (np.abs(x) < ax) & (np.abs(y) < ay) & (np.abs(z) < az)
    solid = inside_box.copy()
    for kind, z0, z1, ol in loops:
        d2, _, inl = ol.project(xy)
        top_plane = np.where(inl, d2, top_plane)
        dz = np.maximum(np.maximum(z0 - z, z - z1), 0)
        ds.append(np.hypot(d2, dz))  # the wall
        zc = z0 if kind == 0 else z1  # the floor or the boss's top
        ds.append(np.where(inl, np.abs(z - zc), np.hypot(d2, z - zc)))
        if kind == 0:
            solid &= ~(inl & (z > z0))
        else:
            solid |= inl & (z >= az) & (z < z1)
    ds.append(np.hypot(top_plane, z - az))
    d = np.min(np.array(ds), axis=0)
    return np.where(solid, -d, d)


def c3_near_trim_points(info, n, radius, seed):
    """n points within `radius` of the outlines' rims on the plate's top face (the trim
    curves of that face and of the walls' ends), seeded."""
    rng = np.random.default_rng(seed)
    W, D, T, nl = info[1], info[2], info[3], int(info[4])
    at = 5
    outs = []
    for _ in range(nl):
        F = int(info[at + 3])
        outs.append(Outline(info[at + 4: at + 4 + 16 * F].reshape(F, 8, 2)))
        at += 4 + 16 * F
    li = rng.integers(0, nl, n)
    out = np.empty((n, 3))
    for k in range(nl):
        m = li == k
        ol = outs[k]
        pid = rng.integers(0, len(ol.pieces), m.sum())
        t = rng.random(m.sum())
        q, _, _ = ol._eval(ol.pieces[pid], t)
        v = rng.normal(size=(m.sum(), 3))
        v *= (radius * rng.random(m.sum()) ** (1 / 3) / np.linalg.norm(v, axis=1))[:, None]
        out[m, :2] = q + v[:, :2]
        out[m, 2] = T / 2 + v[:, 2]
    return out


def crease_points(o, k, h, ws):
    """Points at distance h from the C0 knot line of a "crease:<k>" part's top (face 0),
    each on the bisector of the two sides' normals at a point of the line, so that the
    distance along the line is critical there: outside over a ridge (k = 0, 1, expected
    +h), inside under a valley (k = 2, 3, expected -h); the other side sees the faces.
    Normals from one-sided differences of the surface across the line (knot 0.5 along u
    for even k, along v for odd k), at line parameters ws. Returns (points, expected)."""
    held = k & 1

    def at(a, w):
        return np.array(o.nurbs_eval(0, a, w) if held == 0 else o.nurbs_eval(0, w, a))

    pts, exp = [], []
    e = 1e-6
    for w in ws:
        q = at(0.5, w)
        t = at(0.5, w + e) - at(0.5, w - e)
        a, b = q - at(0.5 - e, w), at(0.5 + e, w) - q
        na, nb = (np.cross(a, t), np.cross(b, t)) if held == 0 else (np.cross(t, a), np.cross(t, b))
        na, nb = na / np.linalg.norm(na), nb / np.linalg.norm(nb)
        if na[2] < 0:
            na, nb = -na, -nb
        m = (na + nb) / np.linalg.norm(na + nb)
        s = 1.0 if k < 2 else -1.0
        pts.append(q + s * h * m)
        exp.append(s * h)
    return np.array(pts), np.array(exp)


# ---------------------------------------------------------------- the oblique cut
#
# A cylinder of radius R on a disc at z = 0, cut by the plane z = L + t x: a convex
# solid, so an inside point's distance is its nearest face's (plane or cylinder); the
# cut is an ellipse edge.


def oblique_frame(t):
    c = 1.0 / np.sqrt(1.0 + t * t)
    return c, t * c


def oblique_inside(p, R, L, t):
    """(inside, depth) of the oblique cut: the depth is the nearest face's distance."""
    c, _ = oblique_frame(t)
    rho = np.hypot(p[:, 0], p[:, 1])
    cut = (L + t * p[:, 0] - p[:, 2]) * c
    inside = (p[:, 2] > 0) & (cut > 0) & (rho < R)
    return inside, np.minimum(np.minimum(p[:, 2], cut), R - rho)


def oblique_edge_points(R, L, t, h, thetas):
    """Points at distance h outside the ellipse edge, each on the bisector of the
    cylinder's and the cut plane's outward normals at a point of the edge (the edge's
    tangent is normal to both, and the solid is convex there): expected +h."""
    c, s = oblique_frame(t)
    pts = []
    for th in thetas:
        q = np.array([R * np.cos(th), R * np.sin(th), L + R * t * np.cos(th)])
        m = np.array([np.cos(th), np.sin(th), 0.0]) + np.array([-s, 0.0, c])
        pts.append(q + h * m / np.linalg.norm(m))
    return np.array(pts), np.full(len(thetas), h)


def oblique_boundary_samples(R, L, t, step):
    """Points of the oblique cut's boundary at most `step` apart."""
    c, _ = oblique_frame(t)
    n = int(np.ceil(2 * np.pi * R / step))
    out = []
    for th in np.linspace(0, 2 * np.pi, n, endpoint=False):
        x, y = R * np.cos(th), R * np.sin(th)
        ztop = L + t * x
        for z in np.linspace(0, ztop, int(np.ceil(ztop / step)) + 1):
            out.append([x, y, z])
    g = np.arange(-R, R + step, step)
    gx, gy = np.meshgrid(g, g)
    keep = np.hypot(gx, gy) <= R
    disc = np.stack([gx[keep], gy[keep]], 1)
    out.extend([[x, y, 0.0] for x, y in disc])
    # the cut: the disc seen along z, lifted on to the plane (spacing grows by 1 / c)
    g2 = np.arange(-R, R + step * c, step * c)
    gx, gy = np.meshgrid(g2, g2)
    keep = np.hypot(gx, gy) <= R
    out.extend([[x, y, L + t * x] for x, y in zip(gx[keep], gy[keep])])
    return np.array(out)

"""CPU restatement of the viewer's ray cast over baked brick levels -- TEST INFRASTRUCTURE.

The reference ray-casts the part's brick cache in its viewer shader
(crates/icam-sim/src/cache.wgsl: trace_ray :441-503 over the levels, walk :335-420 over
a level's bricks, march :270-318 over a brick's cells, cell_hit :191-268 and span_root
:251-268 for the trilinear interpolant's first zero along the ray in a cell, tri :173-177
and grad :179-184, value :146-155). This module restates that algorithm in single
precision with numpy float32 scalars, one ray at a time (pure Python loops: small cases
only), in the operation order of the device kernel (implicitcam_amd/csrc/cast.hip, which
is compiled without floating-point contraction), so the two agree to rounding. It is the
checker of `icam_hip_cast` in tests; nothing in the product imports it.

Rays: (n, 6) float64 (origin, unit direction). Levels: (lo, h, nb) with a map (uint32,
x fastest) and an atlas (uint32, 365 words a slot) as a bake returns them.
Returns (t, grad, level, slot) per ray: t = -1 when nothing is met.
"""
import numpy as np

F = np.float32
BIG = F(1e30)
AIR, SOLID, HOST, WORDS = 0xFFFFFFFF, 0xFFFFFFFE, 0xFFFFFFFD, 365
ONE, ZERO, HALF = F(1), F(0), F(0.5)


def mix(a, b, t):
    return a * (ONE - t) + b * t  # cache.wgsl uses WGSL mix: a (1 - t) + b t


def node_value(atlas, slot, n):  # cache.wgsl:146-155 (unpack2x16snorm)
    w = int(atlas[slot * WORDS + (n >> 1)])
    q = (w >> 16) if (n & 1) else (w & 0xFFFF)
    if q >= 0x8000:
        q -= 0x10000
    return max(F(q) / F(32767), F(-1))


def trilerp(f, qx, qy, qz):  # cache.wgsl:173-177
    lo = mix(mix(f[0], f[1], qx), mix(f[2], f[3], qx), qy)
    hi = mix(mix(f[4], f[5], qx), mix(f[6], f[7], qx), qy)
    return mix(lo, hi, qz)


def trigrad(f, qx, qy, qz):  # cache.wgsl:179-184
    return (mix(mix(f[1] - f[0], f[3] - f[2], qy), mix(f[5] - f[4], f[7] - f[6], qy), qz),
            mix(mix(f[2] - f[0], f[3] - f[1], qx), mix(f[6] - f[4], f[7] - f[5], qx), qz),
            mix(mix(f[4] - f[0], f[5] - f[1], qx), mix(f[6] - f[2], f[7] - f[3], qx), qy))


def poly3(c, u):
    return ((c[3] * u + c[2]) * u + c[1]) * u + c[0]


def first_zero(c, lo, hi):  # cache.wgsl:251-268
    if hi <= lo or poly3(c, hi) > ZERO:
        return F(-1)
    for _ in range(20):
        m = HALF * (lo + hi)
        if poly3(c, m) > ZERO:
            lo = m
        else:
            hi = m
    return HALF * (lo + hi)


def cell_zero(f, q0, dq):  # cache.wgsl:191-249
    s0 = trilerp(f, *q0)
    if s0 <= ZERO:
        return ZERO
    three, two3 = F(3), F(2) / F(3)
    s1 = trilerp(f, q0[0] + dq[0] / three, q0[1] + dq[1] / three, q0[2] + dq[2] / three)
    s2 = trilerp(f, q0[0] + dq[0] * two3, q0[1] + dq[1] * two3, q0[2] + dq[2] * two3)
    s3 = trilerp(f, q0[0] + dq[0], q0[1] + dq[1], q0[2] + dq[2])
    d1 = s1 - s0
    d2 = s2 - F(2) * s1 + s0
    d3 = s3 - F(3) * s2 + F(3) * s1 - s0
    c = (s0, F(3) * (d1 - HALF * d2 + d3 / three), F(9) * (HALF * d2 - HALF * d3), F(27) * d3 / F(6))
    r1 = r2 = HALF
    qa, qb = F(3) * c[3], F(2) * c[2]
    disc = qb * qb - F(4) * qa * c[1]
    if disc >= ZERO:
        q = F(-0.5) * (qb + (ONE if qb >= ZERO else F(-1)) * np.sqrt(disc))
        if abs(qa) > F(1e-20):
            r1 = q / qa
        if abs(q) > F(1e-20):
            r2 = c[1] / q
    x = min(max(min(r1, r2), ZERO), ONE)
    y = min(max(max(r1, r2), ZERO), ONE)
    if HALF < x:
        cuts = (HALF, x, y)
    elif HALF < y:
        cuts = (x, HALF, y)
    else:
        cuts = (x, y, HALF)
    for lo, hi in ((ZERO, cuts[0]), (cuts[0], cuts[1]), (cuts[1], cuts[2]), (cuts[2], ONE)):
        r = first_zero(c, lo, hi)
        if r >= ZERO:
            return r
    return F(-1)


def next_axis(tn):
    if tn[0] <= tn[1] and tn[0] <= tn[2]:
        return 0
    return 1 if tn[1] <= tn[2] else 2


def brick_hit(atlas, slot, base, h, o, d, inv, t_in, t_out):  # cache.wgsl:270-318
    cell, step, tn, ts = [0] * 3, [0] * 3, [ZERO] * 3, [ZERO] * 3
    for k in range(3):
        pk = o[k] + d[k] * t_in
        cell[k] = int(min(max(np.floor((pk - base[k]) / h), ZERO), F(7)))
        step[k] = 1 if d[k] > ZERO else (-1 if d[k] < ZERO else 0)
        tn[k] = (base[k] + (F(cell[k]) + (ONE if d[k] > ZERO else ZERO)) * h - o[k]) * inv[k]
        ts[k] = abs(h * inv[k])
    t = t_in
    for _ in range(32):
        tc = min(min(tn[0], tn[1]), tn[2])
        te = min(tc, t_out)
        n0 = cell[0] + 9 * (cell[1] + 9 * cell[2])
        f = [node_value(atlas, slot, n0 + (k & 1) + 9 * ((k >> 1) & 1) + 81 * (k >> 2)) for k in range(8)]
        if min(f) <= ZERO:
            q0, dq = [], []
            for k in range(3):
                corner = base[k] + F(cell[k]) * h
                q0.append((o[k] + d[k] * t - corner) / h)
                dq.append(d[k] * (te - t) / h)
            u = cell_zero(f, q0, dq)
            if u >= ZERO:
                return t + u * (te - t), trigrad(f, q0[0] + dq[0] * u, q0[1] + dq[1] * u, q0[2] + dq[2] * u)
        if tc >= t_out:
            break
        t = tc
        a = next_axis(tn)
        cell[a] += step[a]
        tn[a] = tn[a] + ts[a]
        if cell[a] < 0 or cell[a] > 7:
            break
    return F(-1), None


def level_hit(L, o, d, inv, t0, t1):  # cache.wgsl:335-420 (the part's cache)
    lo, h, nb, mp, atlas = L
    side = F(8) * h
    step = [1 if d[k] > ZERO else (-1 if d[k] < ZERO else 0) for k in range(3)]
    ts = [abs(side * inv[k]) for k in range(3)]
    cell, tn = [0] * 3, [ZERO] * 3
    t, fresh = t0, True
    for _ in range(4096):
        if fresh:
            for k in range(3):
                pk = o[k] + d[k] * (t + F(1e-4) * h)
                cell[k] = int(min(max(np.floor((pk - lo[k]) / side), ZERO), F(nb[k]) - ONE))
                tn[k] = (lo[k] + (F(cell[k]) + (ONE if d[k] > ZERO else ZERO)) * side - o[k]) * inv[k]
            fresh = False
        entry = int(mp[cell[0] + nb[0] * (cell[1] + nb[1] * cell[2])])
        t_out = min(min(min(tn[0], tn[1]), tn[2]), t1)
        if entry in (SOLID, HOST):
            return t, (-d[0], -d[1], -d[2]), entry
        if entry != AIR:
            base = [lo[k] + F(cell[k]) * side for k in range(3)]
            r, g = brick_hit(atlas, entry, base, h, o, d, inv, t, t_out)
            if r >= ZERO:
                return r, g, entry
        if t_out >= t1:
            break
        t = t_out
        a = next_axis(tn)
        cell[a] += step[a]
        tn[a] = tn[a] + ts[a]
        if cell[a] < 0 or cell[a] >= nb[a]:
            break
    return F(-1), None, AIR


def cast(levels, rays):
    """levels: [(lo (3,), h, nb (3,), map, atlas)], finest first; rays: (n, 6)."""
    L = [([F(v) for v in lo], F(h), [int(v) for v in nb], mp, at) for lo, h, nb, mp, at in levels]
    n = len(rays)
    t_out = np.full(n, -1.0, np.float32)
    g_out = np.zeros((n, 3), np.float32)
    lv_out = np.full(n, 0xFFFFFFFF, np.uint32)
    slot_out = np.full(n, AIR, np.uint32)
    for i in range(n):
        o = [F(v) for v in rays[i, :3]]
        d = [F(v) for v in rays[i, 3:]]
        inv = [(BIG if d[k] + F(1e-30) >= ZERO else -BIG) if abs(d[k]) < F(1e-12) else ONE / d[k] for k in range(3)]
        enter, leave = [], []
        t = BIG
        for lo, h, nb, _, _ in L:
            a, b = -BIG, BIG
            for k in range(3):
                hi = lo[k] + F(nb[k]) * F(8) * h
                x0, x1 = (lo[k] - o[k]) * inv[k], (hi - o[k]) * inv[k]
                a, b = max(a, min(x0, x1)), min(b, max(x0, x1))
            inside = a <= b and b >= ZERO
            enter.append(max(a, ZERO) if inside else BIG)
            leave.append(b if inside else -BIG)
            t = min(t, enter[-1])
        for _ in range(64):
            if not t < BIG:
                break
            past = t + max(abs(t), ONE) * F(1e-6)
            k = next((j for j in range(len(L)) if enter[j] <= past < leave[j]), len(L))
            if k == len(L):
                t = min([e for e in enter if e > past], default=BIG)
                continue
            t1 = leave[k]
            for j in range(k):
                if enter[j] > past:
                    t1 = min(t1, enter[j])
            r, g, s = level_hit(L[k], o, d, inv, t, t1)
            if r >= ZERO:
                t_out[i], g_out[i], lv_out[i], slot_out[i] = r, g, k, s
                break
            t = t1
    return t_out, g_out, lv_out, slot_out

"""The brick rules the device runs, checked on the CPU through their host build.

`bridge_nodes` (implicitcam_amd/csrc/brick_rules.hpp) gives a band brick's nodes left
without a value their bridged values. It is checked bit for bit against the rule
restated below from the reference (crates/icam-sim/src/partcache.rs:872-909) on
synthetic 9^3 node sets: every pass averages the neighbours known when the pass began
(a snapshot), brackets by the nodes known exactly or by sign (its `exact` array), and
sends the nodes nothing reaches to air.
"""
import ctypes
import math
import os

import numpy as np
import pytest

LIB = os.path.join(os.path.dirname(__file__), "..", "implicitcam_amd", "_lib", "libicam_rules.so")


@pytest.fixture(scope="module")
def rules():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} is not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB)
    lib.icam_rules_bridge.restype = ctypes.c_long
    lib.icam_rules_bridge.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
    return lib


def around(n):
    """partcache.rs:811-819: the 26 neighbours of node n with their distances in cells,
    x fastest, then y, then z."""
    i, j, k = n % 9, (n // 9) % 9, n // 81
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                if (di, dj, dk) == (0, 0, 0):
                    continue
                a, b, c = i + di, j + dj, k + dk
                if 0 <= a < 9 and 0 <= b < 9 and 0 <= c < 9:
                    yield a + 9 * b + 81 * c, math.sqrt(di * di + dj * dj + dk * dk)


def bridge_restated(f, todo, h, clamp):
    """partcache.rs:872-909 on f (NaN where unknown) and the open nodes `todo`."""
    f = list(f)
    exact = list(f)
    todo = list(todo)
    bridged = 0
    while todo:
        before = len(todo)
        snapshot = list(f)
        keep = []
        for n in todo:
            s, cnt, lo, hi = 0.0, 0, -clamp, clamp
            for m, dist in around(n):
                if not math.isnan(snapshot[m]):
                    s += snapshot[m]
                    cnt += 1
                if not math.isnan(exact[m]):
                    lo = max(lo, exact[m] - dist * h)
                    hi = min(hi, exact[m] + dist * h)
            if cnt == 0:
                keep.append(n)
                continue
            v = s / cnt
            f[n] = min(max(v, lo), hi) if lo <= hi else 0.5 * (lo + hi)
            bridged += 1
        todo = keep
        if len(todo) == before:
            for n in todo:
                f[n] = clamp
                bridged += 1
            break
    return f, bridged


def run_case(rules, val, st, h, clamp):
    f = [float(v) if s in (0, 3) else math.nan for v, s in zip(val, st)]
    todo = [n for n in range(729) if st[n] in (1, 2)]
    want, want_count = bridge_restated(f, todo, h, clamp)
    v = np.ascontiguousarray(val, dtype=np.float64).copy()
    s = np.ascontiguousarray(st, dtype=np.uint8).copy()
    got_count = rules.icam_rules_bridge(v.ctypes.data, s.ctypes.data, h, clamp)
    assert got_count == want_count
    assert (s[np.asarray(todo, dtype=np.int64)] == 4).all() if todo else True
    np.testing.assert_array_equal(v, np.asarray(want))


def test_bridge_random_mixes(rules):
    rng = np.random.default_rng(7)
    h = 0.5
    clamp = h * math.sqrt(3.0)
    for case in range(12):
        st = rng.choice([0, 1, 2, 3], size=729, p=[0.45, 0.2, 0.25, 0.1]).astype(np.uint8)
        val = rng.uniform(-clamp, clamp, 729)
        val[st == 3] = np.where(rng.random((st == 3).sum()) < 0.5, -clamp, clamp)
        run_case(rules, val, st, h, clamp)


def test_bridge_needs_several_passes(rules):
    # a slab of exact nodes at z = 0 and everything above it open: each pass reaches one
    # layer further, so later layers average nodes bridged in earlier passes only
    h = 0.25
    clamp = h * math.sqrt(3.0)
    st = np.full(729, 2, dtype=np.uint8)
    val = np.zeros(729)
    for n in range(81):
        st[n] = 0
        val[n] = -0.3 * clamp + 0.01 * (n % 9) - 0.02 * (n // 9)
    st[81 * 8 + 40] = 3  # a filled node in the top layer
    val[81 * 8 + 40] = clamp
    run_case(rules, val, st, h, clamp)


def test_bridge_nothing_known(rules):
    # no node known: every node goes to air
    h = 1.0
    clamp = math.sqrt(3.0)
    st = np.full(729, 1, dtype=np.uint8)
    run_case(rules, np.zeros(729), st, h, clamp)


def test_bridge_empty_bracket(rules):
    # exact neighbours of a node farther apart than their distances allow: the middle
    h = 0.1
    clamp = h * math.sqrt(3.0)
    st = np.zeros(729, dtype=np.uint8)
    val = np.zeros(729)
    c = 4 + 9 * 4 + 81 * 4
    st[c] = 2
    val[c - 1], val[c + 1] = -clamp, clamp
    run_case(rules, val, st, h, clamp)

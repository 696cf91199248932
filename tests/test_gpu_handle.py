"""Sharing a handle, and the failure paths of the C ABI, on an MI355X.

The reference's viewer builds two levels at once through one runner
(crates/icam-sim/src/play.rs:979-997), so calls on a handle may come from several
threads and from several streams; a call that returns before its work is done
(query_device) must not let the next call reuse its scratch early. A batch whose band
outgrows its room, and a search that runs out of rounds, must say so.
"""
import ctypes
import os
import threading

import numpy as np
import pytest

from implicitcam_amd import FP64, MIXED, Batch, Part, PartGpu, c4_level, level_for, uniform_points
from implicitcam_amd.exact import lib
from oracle import Oracle

pytestmark = pytest.mark.gpu
SEED = 0x2545F4914F6CDD1D
PRECS = [FP64, MIXED]


def tol(prec, part):
    return 1e-9 if prec == FP64 else 1e-6 * part.diagonal


@pytest.fixture(scope="module")
def c2():
    p = Part("c2")
    return p, Oracle(p), PartGpu(p, 12 * 0.5)


@pytest.mark.parametrize("prec", PRECS)
def test_query_device_on_a_side_stream_then_query(c2, prec):
    """query_device enqueued on a side stream returns at once; a query on the handle
    right after it must wait for it (they share the handle's scratch): both answers
    match the oracle."""
    import torch

    part, o, g = c2
    n = 1 << 16
    pts1 = uniform_points(SEED, n, part.lo - 3, part.hi + 3)
    pts2 = uniform_points(SEED + 1, 1 << 12, part.lo - 3, part.hi + 3)
    x = torch.from_numpy(pts1).cuda()
    d = torch.empty(n, dtype=torch.float64, device="cuda")
    fl = torch.empty(n, dtype=torch.int32, device="cuda")
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    g.query_device(x.data_ptr(), n, d.data_ptr(), fl.data_ptr(), None, prec, side.cuda_stream)
    d2, f2, _ = g.query(pts2, prec)
    side.synchronize()
    d1, f1 = d.cpu().numpy(), fl.cpu().numpy()
    for pts, dd, ff in ((pts1, d1, f1), (pts2, d2, f2)):
        dref, _, _ = o.sdf(pts)
        assert (ff == 0).all()
        assert np.abs(dd - dref).max() <= tol(prec, part)
        assert ((dd < 0) == (dref < 0)).all()


def test_two_threads_bake_through_one_handle():
    """Two threads bake two levels of the box with a hole through one handle at once:
    each level's map and atlas equal the oracle's."""
    part = Part("c1")
    o = Oracle(part)
    g = PartGpu(part, 12 * 1.6)
    levels = [level_for(part, 1.6, 8), level_for(part, 2.4, 8)]
    want = [o.bake(lv)[:2] for lv in levels]
    got = [[] for _ in levels]
    errors = []

    def work(k):
        try:
            for rep in range(4):
                got[k].append(g.bake(levels[k], MIXED if rep % 2 else FP64)[:2])
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(e)

    ts = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    for k in range(2):
        assert len(got[k]) == 4
        for m, a in got[k]:
            assert (m == want[k][0]).all() and len(a) == len(want[k][1]) and (a == want[k][1]).all()


def test_batch_result_fails_when_a_band_outgrows_its_room(monkeypatch):
    """A part's band past its room (held to four bricks here) fails icam_hip_batch_result
    even when no statistics are asked for, before any result is handed out."""
    monkeypatch.setenv("ICAM_BATCH_ROOM", "4")
    part = Part("c2")
    lv = c4_level(part, 64)
    g = PartGpu(part, 12 * lv.h)
    b = Batch([g], [lv], MIXED, tune=False)
    b.run()
    b.sync()
    dm, da = ctypes.c_void_p(), ctypes.c_void_p()
    rc = lib().icam_hip_batch_result(b.b, 0, ctypes.byref(dm), ctypes.byref(da), None)
    assert rc != 0 and b"outgrew" in lib().icam_hip_last_error()
    assert dm.value is None and da.value is None
    with pytest.raises(RuntimeError, match="outgrew"):
        b.download(0)
    b.close()


def test_search_out_of_rounds_is_undecided(monkeypatch):
    """With the searches held to one round over parts, points whose parts under their
    answer were not all read come back undecided (flag 2, a lower bound), never as an
    exact answer; the points that do come back exact match the oracle."""
    monkeypatch.setenv("ICAM_SEARCH_ROUNDS", "1")
    part = Part("c2")
    g = PartGpu(part, 6.0)
    monkeypatch.delenv("ICAM_SEARCH_ROUNDS")
    o = Oracle(part)
    lo, hi = part.lo.copy(), part.hi.copy()
    lo[2] = hi[2] - 12.0  # about the NURBS top
    pts = uniform_points(SEED, 1 << 13, lo, hi + 2)
    dref, _, _ = o.sdf(pts)
    d, flags, _ = g.query(pts, FP64)
    assert (flags == 2).sum() > 0
    assert set(np.unique(flags)) <= {0, 2}
    ok = flags == 0
    assert np.abs(d[ok] - dref[ok]).max() <= 1e-9
    # an undecided value is a lower bound on the distance
    assert (d[~ok] <= np.abs(dref[~ok]) + 1e-9).all()


@pytest.mark.parametrize("lanes", [2, 4, 8])
@pytest.mark.parametrize("prec", PRECS)
def test_nodes_searched_several_lanes_a_node(monkeypatch, lanes, prec):
    """A brick's nodes searched two, four or eight lanes a node (the tuned knob; one by
    default on these parts; ICAM_NODE_LANES sets it for a handle): the NURBS-top block's
    coarse level and the trimmed plate's bake bit-identical to the oracle's."""
    monkeypatch.setenv("ICAM_NODE_LANES", str(lanes))
    for name, h in (("c2", 1.0), ("c3", 1.0)):
        part = Part(name)
        g = PartGpu(part, 12 * h)
        lv = level_for(part, h, 8)
        m_ref, a_ref, _ = Oracle(part).bake(lv)
        m, a, st = g.bake(lv, prec)
        assert st.undecided == 0
        assert (m == m_ref).all() and len(a) == len(a_ref) and (a == a_ref).all()
        g.close()


def test_tune_keeps_the_fastest_and_bakes_alike():
    """icam_hip_tune times the nodes' search at 1, 2, 4 and 8 lanes a node on a level, and in
    the mixed form the centres' decision in single and in double precision, and keeps the
    fastest: its report names every candidate and the choices, and the bakes after it stay
    bit-identical to the oracle's."""
    part = Part("c2")
    lv = level_for(part, 1.0, 8)
    g = PartGpu(part, 12 * lv.h)
    m_ref, a_ref, _ = Oracle(part).bake(lv)
    for prec in PRECS:
        rep = g.tune(lv, prec)
        for k in (1, 2, 4, 8):
            assert f"lanes a node {k}:" in rep
        assert "chosen" in rep
        assert ("centres chosen" in rep) == (prec == MIXED)
        m, a, _ = g.bake(lv, prec)
        assert (m == m_ref).all() and len(a) == len(a_ref) and (a == a_ref).all()
    g.close()

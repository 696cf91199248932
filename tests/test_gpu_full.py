"""Parity at the configured sizes (BASELINE's five workloads), on an MI355X, through the
C ABI, against the FP64 oracle and against size-independent properties.

- C2: the 256^3 level of the NURBS-top block at 0.5 mm (32^3 bricks), batched and
  single-level, map and atlas bit-identical to the oracle's in both precisions.
- C3: the trimmed plate's 2^24 queries (half within 0.05 mm of its trim curves): no
  point undecided, the two precisions' signs equal on every point and their distances
  within 1e-6 of the diagonal, both against the oracle on a 2^14 sample.
- C4: parts of the batched workload at their own grids (32^3 to 256^3 nodes), two of
  each grid size, in one tuned batch: maps and atlases bit-identical to the oracle's.
- C5: the large mixed solid (1,998 faces, 800 of them rational 32 x 32 nets): a 2^14
  sample (30 % uniform, 70 % within 1 mm of the surface) against the oracle, 2^22
  queries by properties, and a window of its 1024-node level against the oracle's bake.

The bar: sign and brick map bit-exact; distances within 1e-6 of the part's diagonal;
the double-precision atlas bit-identical; the mixed atlas bit-identical where the
search's single-precision winner is certain (C2, C4), else within one quantum on a
handful of nodes at near-ties between features (C5).
"""
import numpy as np
import pytest

from implicitcam_amd import FP64, MIXED, Batch, Part, PartGpu, c4_grid, c4_level, level_for, near_trim_points, uniform_points
from implicitcam_amd.part import near_surface_points
from oracle import Oracle

pytestmark = pytest.mark.gpu
SEED = 0x2545F4914F6CDD1D
PRECS = [FP64, MIXED]


def q16(a):
    return a.view(np.int16).astype(np.int64)


def assert_level(m, a, m_ref, a_ref, exact=True, what=""):
    assert (m == m_ref).all(), f"{what}: brick map differs"
    qa, qr = q16(a), q16(a_ref)
    assert len(qa) == len(qr), f"{what}: stored bricks differ"
    assert ((qa > 0) == (qr > 0)).all(), f"{what}: a node's sign differs"
    dq = np.abs(qa - qr)
    if exact:
        assert dq.max(initial=0) == 0, f"{what}: {int((dq > 0).sum())} nodes differ, max {dq.max()}"
    else:
        assert dq.max(initial=0) <= 1, f"{what}: max {dq.max()} quanta"
        assert (dq > 0).mean() < 1e-5, f"{what}: {int((dq > 0).sum())} nodes a quantum off"


@pytest.fixture(scope="module")
def c2_level():
    part = Part("c2")
    lv = level_for(part, 0.5, 32)
    lv.nb[:] = [32, 32, 32]
    g = PartGpu(part, 12 * 0.5)
    m_ref, a_ref, _ = Oracle(part).bake(lv)
    return part, lv, g, m_ref, a_ref


@pytest.mark.parametrize("prec", PRECS)
def test_c2_full_level_bit_exact(c2_level, prec):
    """256^3 nodes at 0.5 mm: the single-level entry point and the batched evaluator."""
    part, lv, g, m_ref, a_ref = c2_level
    m, a, st = g.bake(lv, prec)
    assert st.undecided == 0
    assert_level(m, a, m_ref, a_ref, True, "single")
    b = Batch([g], [lv], prec, tune=False)
    b.run()
    b.sync()
    m, a, st = b.download(0)
    b.close()
    assert_level(m, a, m_ref, a_ref, True, "batch")


def test_c3_full_queries():
    """2^24 trimmed-plate queries in both precisions: every point decided, the signs of
    the two equal, distances within the tolerance of each other; a 2^14 sample (half of
    it at the trim curves) against the oracle."""
    import torch

    part = Part("c3")
    g = PartGpu(part, 20.0)
    n = 1 << 24
    pts = np.concatenate([uniform_points(SEED, n // 2, part.lo - 2, part.hi + 2), near_trim_points(part, 0x9E3779B97F4A7C15, n // 2, 0.05)])
    dev = torch.device("cuda:0")
    d_pts = torch.from_numpy(pts).to(dev)
    out = {}
    for prec in PRECS:
        d = torch.empty(n, dtype=torch.float64, device=dev)
        f = torch.empty(n, dtype=torch.int32, device=dev)
        g.query_device(d_pts.data_ptr(), n, d.data_ptr(), f.data_ptr(), None, prec, None)
        torch.cuda.synchronize()
        out[prec] = (d.cpu().numpy(), f.cpu().numpy())
        assert (out[prec][1] == 0).all(), f"precision {prec}: {int((out[prec][1] != 0).sum())} points undecided"
    dd, dm = out[FP64][0], out[MIXED][0]
    assert ((dd < 0) == (dm < 0)).all()
    assert np.abs(dd - dm).max() <= 1e-6 * part.diagonal
    idx = np.concatenate([np.arange(0, n // 2, n >> 14), np.arange(n // 2, n, n >> 14)])
    dref, _, _ = Oracle(part).sdf(pts[idx])
    for prec in PRECS:
        d = out[prec][0][idx]
        assert ((d < 0) == (dref < 0)).all()
        assert np.abs(d - dref).max() <= (1e-9 if prec == FP64 else 1e-6 * part.diagonal)


def test_c4_configured_grids():
    """Two parts of each grid size of the batched workload (32^3, 64^3, 128^3, 256^3
    nodes) in one tuned mixed-precision batch, and the same in double precision: maps and
    atlases bit-identical to the oracle's."""
    by_grid = {}
    for i in range(64):
        by_grid.setdefault(c4_grid(i), []).append(i)
    idx = [i for gsz in sorted(by_grid) for i in by_grid[gsz][:2]]
    assert len({c4_grid(i) for i in idx}) == 4
    parts = [Part(f"c4:{i}") for i in idx]
    levels = [c4_level(p, c4_grid(i)) for p, i in zip(parts, idx)]
    gpus = [PartGpu(p, 12 * lv.h) for p, lv in zip(parts, levels)]
    refs = [Oracle(p).bake(lv) for p, lv in zip(parts, levels)]
    for prec in PRECS:
        b = Batch(gpus, levels, prec, tune=prec == MIXED)
        b.run()
        b.sync()
        for k, i in enumerate(idx):
            m, a, st = b.download(k)
            assert_level(m, a, refs[k][0], refs[k][1], True, f"c4:{i} grid {c4_grid(i)} precision {prec}")
        b.close()


@pytest.fixture(scope="module")
def c5():
    part = Part("c5")
    return part, Oracle(part), PartGpu(part, 12 * 0.4)


def c5_points(part, seed, n):
    """The workload's mix: 30 % uniform in the box grown by 5 mm, 70 % within 1 mm of
    the surface."""
    k = n * 3 // 10
    return np.concatenate([uniform_points(seed, k, part.lo - 5, part.hi + 5), near_surface_points(part, seed + 1, n - k, 1.0)])


@pytest.mark.parametrize("prec", PRECS)
def test_c5_sample_matches_oracle(c5, prec):
    part, o, g = c5
    pts = c5_points(part, 11, 1 << 14)
    dref, _, _ = o.sdf(pts)
    d, flags, _ = g.query(pts, prec)
    assert (flags == 0).all()
    assert ((d < 0) == (dref < 0)).all()
    assert np.abs(d - dref).max() <= (1e-9 if prec == FP64 else 1e-6 * part.diagonal)


@pytest.mark.parametrize("prec", PRECS)
def test_c5_far_points_match_oracle(c5, prec):
    """2^13 points uniform in the box grown by 5 mm, most of them millimetres from the
    surface: their blocks' candidates come in chunks (k_far), and those left undecided
    past their parts' certificates are settled in slices of their blocks (k_settle_far).
    Every answer against the oracle's."""
    part, o, g = c5
    pts = uniform_points(0x51ED27, 1 << 13, part.lo - 5, part.hi + 5)
    dref, _, _ = o.sdf(pts)
    d, flags, _ = g.query(pts, prec)
    assert (flags == 0).all()
    assert g.query_stats().settled > 0
    assert ((d < 0) == (dref < 0)).all()
    assert np.abs(d - dref).max() <= (1e-9 if prec == FP64 else 1e-6 * part.diagonal)


def test_c5_queries_by_properties(c5):
    """2^22 queries of the workload's mix in both precisions: every point decided, equal
    signs, distances within 1e-6 of the diagonal of each other, |d| no larger than the
    distance to the nearest of the near-surface points' own surface points."""
    import torch

    part, _, g = c5
    n = 1 << 22
    pts = c5_points(part, 21, n)
    dev = torch.device("cuda:0")
    d_pts = torch.from_numpy(pts).to(dev)
    out = {}
    for prec in PRECS:
        d = torch.empty(n, dtype=torch.float64, device=dev)
        f = torch.empty(n, dtype=torch.int32, device=dev)
        g.query_device(d_pts.data_ptr(), n, d.data_ptr(), f.data_ptr(), None, prec, None)
        torch.cuda.synchronize()
        out[prec] = (d.cpu().numpy(), f.cpu().numpy())
        assert (out[prec][1] == 0).all()
    dd, dm = out[FP64][0], out[MIXED][0]
    assert ((dd < 0) == (dm < 0)).all()
    assert np.abs(dd - dm).max() <= 1e-6 * part.diagonal
    # points drawn within 1 mm of the surface are within 1 mm of it
    assert (np.abs(dd[n * 3 // 10:]) <= 1.0 + 1e-9).all()


@pytest.mark.parametrize("prec", PRECS)
def test_c5_level_window_matches_oracle(c5, prec):
    """A 24 x 24 x 13-brick window of the 1024-node level (cell 0.393 mm) over pockets,
    holes and the plate's side: its map and atlas against the oracle's bake of the same
    window (a level of its own on the same node grid)."""
    part, o, g = c5
    full = c4_level(part, 1024)
    lv = c4_level(part, 1024)
    x0, y0 = 100, 70  # bricks: the window reaches the plate's +x side
    lv.lo[0] = full.lo[0] + 8 * full.h * x0
    lv.lo[1] = full.lo[1] + 8 * full.h * y0
    lv.nb[:] = [min(24, full.nb[0] - x0), min(24, full.nb[1] - y0), full.nb[2]]
    m_ref, a_ref, _ = o.bake(lv)
    m, a, st = g.bake(lv, prec)
    assert st.undecided == 0
    assert st.slots > 500
    assert_level(m, a, m_ref, a_ref, prec == FP64, f"c5 window precision {prec}")


@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_batch_with_a_level_past_the_workgroup_sign_pass(prec):
    """A batch holding the large mixed solid's 1024-node level (128 x 96 x 13 bricks, past
    the sign pass in workgroup memory, signed grid-wide in the graph) beside the box with a
    hole: each part's map and atlas bit-identical to its single-level bake."""
    c5 = Part("c5")
    lv5 = c4_level(c5, 1024)
    g5 = PartGpu(c5, 12 * lv5.h)
    c1 = Part("c1")
    lv1 = level_for(c1, 1.6, 16)
    g1 = PartGpu(c1, 12 * 1.6)
    b = Batch([g5, g1], [lv5, lv1], prec, tune=False)
    b.run()
    b.sync()
    for k, (g, lv) in enumerate(((g5, lv5), (g1, lv1))):
        m, a, st = b.download(k)
        m_ref, a_ref, _ = g.bake(lv, prec)
        assert (m == m_ref).all() and len(a) == len(a_ref) and (a == a_ref).all()
    b.close()
    g5.close()
    g1.close()

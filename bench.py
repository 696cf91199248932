"""Benchmark: the exact signed distance and its narrow-band brick bake on one MI355X.

Headline workload (a "step"): the large mixed solid -- a 400 x 300 x 40 mm plate with
400 pockets and 264 countersunk, filleted holes, 1,998 faces of which 800 are rational
bicubic B-splines with 32 x 32 nets (weights in [0.5, 2]), every face trimmed by loops
of 12-64 cubic spans -- baked on its 1024-node grid (1024 nodes along its longest side,
cell 0.393 mm: 128 x 96 x 13 bricks of 8^3 cells) within the narrow band: the decision
of every brick, the signs of the bricks kept whole, the 729 nodes of every band brick,
the settle of nodes past their certificates, flood fill, i16 quantization and packing
into the 365-word atlas, all on the device, in the FP32 + FP64-refine form through the
batched evaluator (the part's level as a batch of one: one launch of the captured graph
a step, its shape chosen by the autotuner; device-resident results, checked equal to the
single-level bake's). The plain FP64 form is timed beside it, and so are the
single-level entry point's bakes of the same level, the same part's scattered queries (30
% uniform, 70 % within 1 mm of the surface), the NURBS-top block's 256^3 level through
the batched evaluator, 2^20 queries on the analytic box-with-hole part, the trimmed
plate's 2^24 queries and the 64-part batched workload.

Prints one JSON line (rank 0, N = 1). `value` is the headline's band bricks baked per
second (mixed form).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32 = 157.3e12  # MI355X dense fp32, FLOP/s
PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12  # B/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def c5_window(part, side=24):
    """A side x side x full-height brick window of the large solid's 1024-node level
    (over pockets, holes and the plate's +x side), on the same node grid: the bounded
    sample the CPU baseline and the parity check bake."""
    from implicitcam_amd import c4_level

    full = c4_level(part, 1024)
    lv = c4_level(part, 1024)
    x0, y0 = 100, 70
    lv.lo[0] = full.lo[0] + 8 * full.h * x0
    lv.lo[1] = full.lo[1] + 8 * full.h * y0
    lv.nb[:] = [min(side, full.nb[0] - x0), min(side, full.nb[1] - y0), full.nb[2]]
    return lv


def large_solid(args, log, threads):
    """The headline: the large mixed solid's 1024-node level baked in both precisions
    (device-resident), its scattered queries in both precisions, and the CPU baseline,
    the parity check and the useful-work count on a window of the same level."""
    import torch

    from implicitcam_amd import FP64, MIXED, Batch, Part, PartGpu, c4_level, near_surface_points, uniform_points
    from implicitcam_amd.flops import useful_flops

    part = Part("c5")
    lv = c4_level(part, 1024)
    t0 = time.time()
    g = PartGpu(part, 12 * lv.h)
    pack_s = time.time() - t0
    log(f"c5: {part.n_faces} faces packed in {pack_s:.1f}s ({g.nodes} sub-patches and curve parts), level {list(lv.nb)} at {lv.h:.4f} mm")
    res = {"faces": part.n_faces, "pack_s": pack_s, "level": list(lv.nb), "h": lv.h}
    for prec, name in ((MIXED, "mixed"), (FP64, "fp64")):
        # the single-level autotuner (lanes a node of the nodes' search), outside the
        # timed steps; a profile run keeps the default shape
        tuning = None if args.profile else g.tune(lv, prec)
        if tuning:
            log(f"c5 tune {name}: {tuning}")
        _, _, st = g.bake(lv, prec)
        for _ in range(args.warmup):
            g.bake_device(lv, prec)
        torch.cuda.synchronize()
        stage_sum = {}
        t = time.perf_counter()
        for _ in range(args.steps):
            g.bake_device(lv, prec)
            for k, v in g.stage_times().items():
                stage_sum[k] = stage_sum.get(k, 0.0) + v
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / args.steps
        res[name] = {"ms": dt * 1e3, "bricks": st.bricks, "band": st.band, "slots": st.slots, "settled": st.settled, "refined": st.refined,
                     "undecided": st.undecided, "bricks_per_s": st.band / dt, "stages": {k: v / args.steps for k, v in stage_sum.items()},
                     "tuning": tuning}
        log(f"c5 bake {name}: {dt * 1e3:.2f} ms, band {st.band}, stored {st.slots}, settled {st.settled}: {st.band / dt / 1e6:.3f} M band bricks/s; "
            f"stages {res[name]['stages']}")
    res["mixed_over_fp64"] = res["fp64"]["ms"] / res["mixed"]["ms"]
    # the headline: the same level through the batched evaluator (one part: the captured
    # graph of the bake's launches, persistent queue stages, autotuned shape; a profile run
    # keeps the untuned shape), one graph launch a step; its map and atlas checked equal to
    # the single-level bake's
    res["batched"] = {}
    for prec, name in ((MIXED, "mixed"), (FP64, "fp64")):
        t0 = time.time()
        b = Batch([g], [lv], prec, tune=not args.profile)
        open_s = time.time() - t0
        for _ in range(max(args.warmup, 1)):
            b.run()
        b.sync()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            b.run()
        b.sync()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / args.steps
        st = b.stats(0)
        stages = b.stage_times()
        same = None
        if not args.profile:
            mb, ab, _ = b.download(0)
            ms_, as_, _ = g.bake(lv, prec)
            same = bool((mb == ms_).all() and len(ab) == len(as_) and (ab == as_).all())
        res["batched"][name] = {"ms": dt * 1e3, "bricks": st.bricks, "band": st.band, "slots": st.slots, "settled": st.settled,
                                "refined": st.refined, "undecided": st.undecided, "bricks_per_s": st.band / dt, "stages": stages,
                                "tuning": b.tuning, "open_and_tune_s": open_s, "equal_to_single_level": same}
        log(f"c5 batched {name}: {dt * 1e3:.2f} ms, {st.band / dt / 1e6:.3f} M band bricks/s, equal to the single-level bake: {same}; "
            f"{b.tuning}; stages {stages}")
        b.close()
    res["batched"]["mixed_over_fp64"] = res["batched"]["fp64"]["ms"] / res["batched"]["mixed"]["ms"]
    if not args.profile:
        # the brick cache's consumer: the viewer's ray cast (icam_hip_cast) of the baked level,
        # a 1024 x 1024 camera looking down at the plate from above one corner
        from implicitcam_amd import cast_device

        dm, da, _ = g.bake_device(lv, MIXED)
        c = 0.5 * (part.lo + part.hi)
        eye = c + np.array([-0.45, -0.55, 0.9]) * part.diagonal
        fwd = (c - eye) / np.linalg.norm(c - eye)
        right = np.cross(fwd, [0.0, 0.0, 1.0])
        right /= np.linalg.norm(right)
        up = np.cross(right, fwd)
        px = (np.arange(1024) + 0.5) / 512.0 - 1.0
        u, v = np.meshgrid(px * 0.45, px * 0.45)
        dirs = fwd + u.reshape(-1, 1) * right + v.reshape(-1, 1) * up
        dirs /= np.linalg.norm(dirs, axis=1)[:, None]
        rays = np.concatenate([np.broadcast_to(eye, dirs.shape), dirs], axis=1)
        nr = len(rays)
        d_rays = torch.from_numpy(np.ascontiguousarray(rays)).cuda()
        d_out = torch.empty((nr, 4), dtype=torch.float32, device="cuda")
        d_hit = torch.empty((nr, 2), dtype=torch.int32, device="cuda")
        for _ in range(max(args.warmup, 1)):
            cast_device([lv], [dm], [da], d_rays.data_ptr(), nr, d_out.data_ptr(), d_hit.data_ptr())
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            cast_device([lv], [dm], [da], d_rays.data_ptr(), nr, d_out.data_ptr(), d_hit.data_ptr())
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / args.steps
        hit_frac = float((d_out[:, 0] >= 0).float().mean().item())
        res["cast"] = {"rays": nr, "ms": dt * 1e3, "rays_per_s": nr / dt, "hit_fraction": hit_frac}
        log(f"c5 cast: {nr} rays in {dt * 1e3:.3f} ms, {nr / dt / 1e6:.1f} M rays/s, {hit_frac:.2f} of them meet the part")
    if args.profile:
        g.close()
        return res, None, None, None
    # scattered queries of the workload's mix, inputs resident in HBM
    n = 1 << args.c5_log2
    k = n * 3 // 10
    pts = np.concatenate([uniform_points(0x2545F4914F6CDD1D, k, part.lo - 5, part.hi + 5), near_surface_points(part, 0x9E3779B97F4A7C15, n - k, 1.0)])
    dev = torch.device("cuda:0")
    d_pts = torch.from_numpy(pts).to(dev)
    qout = {}
    res["queries"] = {"points": n, "near_surface": n - k}
    for prec, name in ((MIXED, "mixed"), (FP64, "fp64")):
        d_out = torch.empty(n, dtype=torch.float64, device=dev)
        d_flag = torch.empty(n, dtype=torch.int32, device=dev)
        g.query_device(d_pts.data_ptr(), n, d_out.data_ptr(), d_flag.data_ptr(), None, prec, None)
        torch.cuda.synchronize()
        t = time.perf_counter()
        g.query_device(d_pts.data_ptr(), n, d_out.data_ptr(), d_flag.data_ptr(), None, prec, None)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        st = g.query_stats()
        qout[name] = (d_out.cpu().numpy(), d_flag.cpu().numpy())
        # the stages' device times, from one more call through the host-buffer entry point
        g.query(pts, prec)
        stages = {key: round(val, 3) for key, val in g.stage_times().items()}
        res["queries"][name] = {"queries_per_s": n / dt, "ms": dt * 1e3, "settled": st.settled, "undecided": int((qout[name][1] != 0).sum()),
                                "stages": stages}
        log(f"c5 queries {name}: {n / dt / 1e6:.2f} M/s, {res['queries'][name]}")
    dm, dd = qout["mixed"][0], qout["fp64"][0]
    res["queries"]["mixed_vs_fp64"] = {"sign_mismatch": int(((dm < 0) != (dd < 0)).sum()), "max_rel_err": float(np.abs(dm - dd).max() / part.diagonal)}

    cpu = check = useful = None
    if not (args.no_cpu and args.no_check):
        from oracle import Oracle, counts

        o = Oracle(part)
        win = c5_window(part)
        # the decision's work and the band nodes' work, counted apart
        counts()
        o.decide(win, threads=threads)
        c_dec = counts()
        t = time.perf_counter()
        m_ref, a_ref, nread = o.bake(win, threads=threads)
        cpu_s = time.perf_counter() - t
        c_all = counts()
        c_nodes = {key: c_all[key] - c_dec[key] for key in c_all}
        nb_win = win.nb[0] * win.nb[1] * win.nb[2]
        cpu = {"value": nread / cpu_s, "unit": "bricks/s", "cores": threads, "kind": "port",
               "sample": f"a {win.nb[0]} x {win.nb[1]} x {win.nb[2]}-brick window of the same level ({nread} band bricks read), "
                         f"FP64 oracle on {threads} host threads, {cpu_s:.1f} s"}
        # one thread on the window's first band bricks
        small = c5_window(part, 6)
        t = time.perf_counter()
        _, _, nread1 = o.bake(small, threads=1)
        cpu1_s = time.perf_counter() - t
        cpu["single_thread"] = {"value": nread1 / cpu1_s, "unit": "bricks/s", "sample": f"a 6 x 6 x {win.nb[2]}-brick window, {nread1} band bricks, {cpu1_s:.1f} s"}
        log(f"cpu oracle: {nread / cpu_s:.1f} bricks/s on {threads} threads, {nread1 / cpu1_s:.1f} on one")
        # useful work a band brick (its nodes) and a brick (its decision), from the window
        useful = {"node_flop_per_band_brick": useful_flops(c_nodes) / max(nread, 1), "decide_flop_per_brick": useful_flops(c_dec) / nb_win,
                  "counts_window_nodes": c_nodes, "counts_window_decide": c_dec}
        if not args.no_check:
            check = {}
            for prec, name in ((MIXED, "mixed"), (FP64, "fp64")):
                m, a, _ = g.bake(win, prec)
                qa = a.view(np.int16).astype(np.int64)
                qr = a_ref.view(np.int16).astype(np.int64)
                same = len(qa) == len(qr)
                check["window_" + name] = {"map_equal": bool((m == m_ref).all()), "atlas_equal": bool(same and (qa == qr).all()),
                                           "max_dq": int(np.abs(qa - qr).max(initial=0)) if same else None,
                                           "nodes_off": int((qa != qr).sum()) if same else None,
                                           "sign_mismatch": int(((qa > 0) != (qr > 0)).sum()) if same else None}
            idx = np.concatenate([np.arange(0, k, max(k >> 11, 1)), np.arange(k, n, max((n - k) >> 12, 1))])
            dref, _, _ = o.sdf(pts[idx], threads=threads)
            for name in ("mixed", "fp64"):
                err = np.abs(qout[name][0][idx] - dref) / part.diagonal
                check["queries_" + name] = {"n": int(len(idx)), "max_rel_err": float(err.max()), "p99_rel_err": float(np.percentile(err, 99)),
                                            "p50_rel_err": float(np.percentile(err, 50)),
                                            "sign_mismatch": int(((qout[name][0][idx] < 0) != (dref < 0)).sum())}
            log(f"c5 check: {check}")
    g.close()
    return res, cpu, check, useful


def same_kernel(cell, kernel):
    """Whether a profile table's kernel cell names `kernel`, whose trailing template
    arguments (the nodes' search's lanes a node) may be left out: `k_search<1, float>`
    names `k_search<1, float, 4>`."""
    name = cell.strip("`")
    return name == kernel or (kernel.endswith(">") and name.startswith(kernel[:-1] + ", "))


def pmc_traffic(path, kernel):
    """HBM bytes per launch of `kernel` from a committed counter profile (its table row's
    scaled column), or None."""
    try:
        for line in open(path):
            cells = [c.strip() for c in line.split("|")]
            if len(cells) > 5 and same_kernel(cells[1], kernel):
                return float(cells[5].replace(",", ""))
    except OSError:
        return None
    return None


def pmc_valu(path, kernel):
    """FP32 and FP64 VALU rates (TFLOP/s) and lane use of `kernel` from a committed
    counter profile, or None."""
    try:
        for line in open(path):
            cells = [c.strip() for c in line.split("|")]
            if len(cells) > 8 and same_kernel(cells[1], kernel):
                return {"fp32_tflops": float(cells[4]), "fp64_tflops": float(cells[6]), "lane_use": float(cells[7]),
                        "frac_fp32_peak": float(cells[4]) * 1e12 / PEAK_FP32}
    except OSError:
        return None
    return None


def trimmed_plate_queries(args, log, threads):
    """The trimmed plate: 2^20 points (2^c3_log2), half uniform over its box (2 mm margin) and half
    within 0.05 mm of its planar faces' trim curves (where the mixed form takes its trim
    decisions in double precision), inputs resident in HBM, in both precisions; the two
    precisions' signs compared on every point and both against the FP64 oracle on a
    sample of 2^14. Scattered points are searched in cells of a 64^3 grid over their
    box (results in the caller's order), each block of them within the boundary's own
    bound over its box, and points past their sub-patches' convexity certificates are
    settled in double precision: this leg is far from the bake's rates."""
    import torch

    from implicitcam_amd import FP64, MIXED, Part, PartGpu, near_trim_points, uniform_points

    part = Part("c3")
    g = PartGpu(part, 20.0)
    n = 1 << args.c3_log2
    pts = np.concatenate([uniform_points(0x2545F4914F6CDD1D, n // 2, part.lo - 2, part.hi + 2), near_trim_points(part, 0x9E3779B97F4A7C15, n // 2, 0.05)])
    dev = torch.device("cuda:0")
    d_pts = torch.from_numpy(pts).to(dev)
    outs = {}
    res = {"points": n, "near_trim": n // 2, "faces": part.n_faces}
    for prec, name in ((MIXED, "mixed"), (FP64, "fp64")):
        d_out = torch.empty(n, dtype=torch.float64, device=dev)
        d_flag = torch.empty(n, dtype=torch.int32, device=dev)
        for _ in range(max(args.warmup, 1)):
            g.query_device(d_pts.data_ptr(), n, d_out.data_ptr(), d_flag.data_ptr(), None, prec, None)
        torch.cuda.synchronize()
        st = g.query_stats()
        t = time.perf_counter()
        for _ in range(args.steps):
            g.query_device(d_pts.data_ptr(), n, d_out.data_ptr(), d_flag.data_ptr(), None, prec, None)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / args.steps
        outs[name] = (d_out.cpu().numpy(), d_flag.cpu().numpy())
        # the stages' device times, from one more call through the host-buffer entry point
        # (the device one records none), outside the timed steps
        g.query(pts, prec)
        stages = g.stage_times()
        res[name] = {"queries_per_s": n / dt, "ms": dt * 1e3, "searched_again_fp64": st.refined, "near_trim_curve": st.unsure_trim,
                     "near_tie": st.unsure_tie, "settled": st.settled, "undecided": int((outs[name][1] != 0).sum()), "stages": stages,
                     "settle_share": (stages.get("points_settle64", 0.0) + stages.get("points_settle_far64", 0.0)) / max(sum(stages.values()), 1e-9)}
        log(f"c3 queries {name}: {n / dt / 1e6:.1f} M/s, {res[name]}")
    dm, dd = outs["mixed"][0], outs["fp64"][0]
    res["mixed_vs_fp64"] = {"sign_mismatch": int(((dm < 0) != (dd < 0)).sum()), "max_rel_err": float(np.abs(dm - dd).max() / part.diagonal),
                            "p99_rel_err": float(np.percentile(np.abs(dm - dd), 99) / part.diagonal)}
    res["mixed_over_fp64"] = res["mixed"]["queries_per_s"] / res["fp64"]["queries_per_s"]
    if not args.no_check:
        from oracle import Oracle

        idx = np.concatenate([np.arange(0, n // 2, max((n // 2) >> 13, 1)), np.arange(n // 2, n, max((n // 2) >> 13, 1))])
        dref, _, _ = Oracle(part).sdf(pts[idx], threads=threads)
        for name in ("mixed", "fp64"):
            err = np.abs(outs[name][0][idx] - dref) / part.diagonal
            res[name]["vs_oracle"] = {"n": int(len(idx)), "max_rel_err": float(err.max()), "p99_rel_err": float(np.percentile(err, 99)),
                                      "p50_rel_err": float(np.percentile(err, 50)), "sign_mismatch": int(((outs[name][0][idx] < 0) != (dref < 0)).sum())}
    log(f"c3 check: {res['mixed_vs_fp64']} {res.get('mixed', {}).get('vs_oracle')}")
    g.close()
    return res


def batched_workload(args, log):
    """64 parts (face counts log-uniform 6..400; turned parts, NURBS blocks with 4 x 4 to
    32 x 32 nets, plates with holes) at grids of 32^3 to 256^3 nodes, baked as one batch
    (one graph launch, autotuned queue grids) in both precisions, against the same 64
    levels baked one after another through the single-level entry point."""
    import torch

    from implicitcam_amd import FP64, MIXED, Batch, Part, PartGpu, c4_grid, c4_level

    t0 = time.time()
    parts = [Part(f"c4:{i}") for i in range(64)]
    levels = [c4_level(p, c4_grid(i)) for i, p in enumerate(parts)]
    gpus = [PartGpu(p, 12 * lv.h) for p, lv in zip(parts, levels)]
    log(f"c4: 64 parts packed in {time.time() - t0:.1f}s, faces {sum(p.n_faces for p in parts)}, "
        f"bricks {sum(lv.nb[0] * lv.nb[1] * lv.nb[2] for lv in levels)}")
    out = {"parts": 64, "faces": [p.n_faces for p in parts], "grids": [c4_grid(i) for i in range(64)]}
    for prec, name in ((MIXED, "mixed"), (FP64, "fp64")):
        t0 = time.time()
        b = Batch(gpus, levels, prec, tune=True)
        tune_s = time.time() - t0
        for _ in range(max(args.warmup, 1)):
            b.run()
        b.sync()
        t = time.perf_counter()
        for _ in range(args.steps):
            b.run()
        b.sync()
        ms = (time.perf_counter() - t) / args.steps * 1e3
        band = sum(b.stats(i).band for i in range(64))
        stages = b.stage_times()
        out[name] = {"ms": ms, "band_bricks": band, "bricks_per_s": band / (ms / 1e3), "tuning": b.tuning, "open_and_tune_s": tune_s,
                     "stages_ms": stages}
        log(f"c4 batch {name}: {ms:.3f} ms, {band} band bricks, {band / (ms / 1e3) / 1e6:.2f} M bricks/s; {b.tuning}; stages {stages}")
        b.close()
        # the same levels one after another (the single-level entry point: one bake each)
        for g, lv in zip(gpus, levels):
            g.bake_device(lv, prec)
        torch.cuda.synchronize()
        per = []
        t = time.perf_counter()
        for g, lv in zip(gpus, levels):
            t1 = time.perf_counter()
            g.bake_device(lv, prec)
            torch.cuda.synchronize()
            per.append((time.perf_counter() - t1) * 1e3)
        seq = (time.perf_counter() - t) * 1e3
        out[name].update({"sequential_ms": seq, "speedup_over_sequential": seq / ms, "per_part_ms_p50": float(np.percentile(per, 50)),
                          "per_part_ms_p99": float(np.percentile(per, 99))})
        log(f"c4 sequential {name}: {seq:.3f} ms (per part p50 {np.percentile(per, 50):.3f} ms, p99 {np.percentile(per, 99):.3f} ms): "
            f"batch {seq / ms:.2f}x faster")
    out["mixed_over_fp64"] = out["fp64"]["ms"] / out["mixed"]["ms"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true", help="skip the parity check against the oracle")
    ap.add_argument("--profile", action="store_true", help="a short run for a profiler: one leg's timed steps (--leg) and a copy of known size")
    ap.add_argument("--leg", choices=["c5", "c2", "c3"], default="c5",
                    help="with --profile: c5 the headline's bake, c2 the NURBS-top block's 256^3 bake (batched and single-level), "
                         "c3 the trimmed plate's queries")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-part batched workload")
    ap.add_argument("--no-c3", action="store_true", help="skip the trimmed plate's scattered queries")
    ap.add_argument("--c3-log2", type=int, default=24, help="log2 of the trimmed plate's query count (the configuration's 2^24)")
    ap.add_argument("--c5-log2", type=int, default=20, help="log2 of the large solid's query count (the configuration's 2^26)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only: the path shards by brick ranges across GPUs (see the design notes); "
                                   "run with --gpus 1"}))
        sys.exit(2)

    import torch

    from implicitcam_amd import FP64, MIXED, Batch, Part, PartGpu, level_for, uniform_points

    if not torch.cuda.is_available():
        raise SystemExit("no GPU: the benchmark runs the HIP kernels on an MI355X")
    torch.cuda.init()
    dev = torch.device("cuda:0")

    threads = len(os.sched_getaffinity(0))
    if args.profile:
        args.no_check = args.no_cpu = True

    def profile_end():
        # a plain copy of known size (256 MiB read, 256 MiB written): the scale for the
        # profiler's HBM byte counters
        x = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()

    if args.profile and args.leg == "c3":
        trimmed_plate_queries(args, log, threads)
        profile_end()
        return
    # the headline: the large mixed solid's 1024-node level, and its queries
    if not args.profile or args.leg == "c5":
        c5, c5_cpu, c5_check, c5_useful = large_solid(args, log, threads)
        if args.profile:
            profile_end()
            return

    h = 0.5
    part = Part("c2")
    lv = level_for(part, h, 32)
    lv.nb[:] = [32, 32, 32]  # the 256^3 node grid
    reach_need = 12 * h
    t0 = time.time()
    gpu = PartGpu(part, reach_need)
    log(f"packed c2: {gpu.nodes} sub-patches and curve parts in {time.time() - t0:.2f}s on {gpu.adapter}")

    def bake_step(prec):
        gpu.bake_device(lv, prec)

    # the level through the batched evaluator (one part: the captured graph of the
    # bake's launches, persistent queue stages, autotuned grids) and through the
    # single-level entry point (the same kernels' single-part forms, host-side sizing)
    res, single, batches = {}, {}, {}
    for prec, name in ((MIXED, "mixed"), (FP64, "fp64")):
        t0 = time.time()
        b = Batch([gpu], [lv], prec, tune=True)
        open_s = time.time() - t0
        for _ in range(max(args.warmup, 1)):
            b.run()
        b.sync()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            b.run()
        b.sync()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / args.steps
        st = b.stats(0)
        stages = b.stage_times()
        res[name] = dict(ms=dt * 1e3, band=st.band, slots=st.slots, bricks=st.bricks, undecided=st.undecided, refined=st.refined,
                         unsure_trim=st.unsure_trim, unsure_tie=st.unsure_tie, settled=st.settled, stages=stages, tuning=b.tuning,
                         open_and_tune_s=open_s)
        batches[name] = b
        log(f"bake {name} (batched evaluator, one part): {dt * 1e3:.3f} ms/step, band {st.band}, slots {st.slots}, undecided "
            f"{st.undecided}, searched again in fp64 {st.refined} (trim {st.unsure_trim}, ties {st.unsure_tie}), settled {st.settled}; "
            f"{b.tuning}; stages {stages}")
        for _ in range(args.warmup):
            bake_step(prec)
        torch.cuda.synchronize()
        t = time.perf_counter()
        stage_sum = {}
        for _ in range(args.steps):
            bake_step(prec)
            for k, v in gpu.stage_times().items():
                stage_sum[k] = stage_sum.get(k, 0.0) + v
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / args.steps
        single[name] = dict(ms=dt * 1e3, bricks_per_s=st.band / dt, stages={k: v / args.steps for k, v in stage_sum.items()})
        log(f"bake {name} (single-level entry point): {dt * 1e3:.3f} ms/step, stages {single[name]['stages']}")

    if args.profile:
        profile_end()
        return
    # 2^20 scattered queries on the analytic part (C1), inputs resident in HBM
    c1 = Part("c1")
    g1 = PartGpu(c1, 12 * 1.6)
    n = 1 << 20
    pts = uniform_points(0x2545F4914F6CDD1D, n, c1.lo - 5, c1.hi + 5)
    d_pts = torch.from_numpy(pts).to(dev)
    d_out = torch.empty(n, dtype=torch.float64, device=dev)
    d_flag = torch.empty(n, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    q = {}
    for prec, name in ((MIXED, "mixed"), (FP64, "fp64")):
        for _ in range(args.warmup):
            g1.query_device(d_pts.data_ptr(), n, d_out.data_ptr(), d_flag.data_ptr(), None, prec, stream)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            g1.query_device(d_pts.data_ptr(), n, d_out.data_ptr(), d_flag.data_ptr(), None, prec, stream)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / args.steps
        q[name] = n / dt
        log(f"c1 queries {name}: {n / dt / 1e6:.1f} M/s")

    check = {}
    c2_cpu = None
    if not args.no_check or not args.no_cpu:
        from oracle import Oracle

        o = Oracle(part)
        t = time.perf_counter()
        m_ref, a_ref, nread = o.bake(lv, threads=threads)
        cpu_s = time.perf_counter() - t
        c2_cpu = {"value": nread / cpu_s, "unit": "bricks/s", "cores": threads, "kind": "port",
                  "sample": f"the whole c2 level ({nread} band bricks of the 32^3), FP64 oracle on {threads} host threads, {cpu_s:.1f} s"}
        log(f"cpu oracle bake (c2): {cpu_s:.1f}s, {nread / cpu_s:.1f} bricks/s on {threads} threads")
        if not args.no_check:
            for prec, name in ((MIXED, "mixed"), (FP64, "fp64")):
                for how, (m, a, st) in (("", batches[name].download(0)), ("single_", gpu.bake(lv, prec))):
                    qa = a.view(np.int16).astype(np.int64)
                    qr = a_ref.view(np.int16).astype(np.int64)
                    same_len = len(qa) == len(qr)
                    check[how + name] = {"map_equal": bool((m == m_ref).all()), "atlas_equal": bool(same_len and (qa == qr).all()),
                                         "max_dq": int(np.abs(qa - qr).max()) if same_len and len(qa) else None,
                                         "sign_mismatch": int(((qa > 0) != (qr > 0)).sum()) if same_len else None}
            # query error against the oracle on a sample of the 2^20 points
            o1 = Oracle(c1)
            sample = pts[:: 64]
            dref, _, _ = o1.sdf(sample, threads=threads)
            for prec, name in ((MIXED, "mixed"), (FP64, "fp64")):
                dq, fl, _ = g1.query(sample, prec)
                err = np.abs(dq - dref) / c1.diagonal
                check["c1_" + name] = {"max_rel_err": float(err.max()), "p99_rel_err": float(np.percentile(err, 99)),
                                       "sign_mismatch": int(((dq < 0) != (dref < 0)).sum()), "n": int(len(sample))}
            log(f"check: {check}")

    # each leg's handles closed before the next opens its own
    for b in batches.values():
        b.close()
    gpu.close()
    g1.close()
    # the trimmed plate's queries, half of them at its trim curves
    c3 = None if args.no_c3 else trimmed_plate_queries(args, log, threads)
    # the batched workload: 64 parts of uneven face counts and grids in one graph
    c4 = None if args.no_batch else batched_workload(args, log)
    mixed, fp64 = res["mixed"], res["fp64"]
    cm, cf = c5["batched"]["mixed"], c5["batched"]["fp64"]
    # the roofline of the headline's dominant work, the band nodes' search (its launches:
    # single-precision search, dense pass, last correction, fallback, settle): the useful
    # floating-point work of reading every band brick's nodes -- the FP64 evaluator's
    # executed primitives a band brick, counted on a window of the same level, times the
    # device code's operations a primitive (implicitcam_amd/flops.py), times the level's
    # band bricks -- over those launches' summed HIP-event time, against the FP32 vector
    # peak (the bulk of the mixed form is single precision)
    node_stages = ["nodes_search32", "nodes_leaves32", "nodes_refine64", "nodes_fallback64", "nodes_mark", "nodes_settle64"]
    node_ms = sum(cm["stages"].get(k, 0.0) for k in node_stages)
    # HBM bytes a launch of those kernels, from the committed counter profile of the
    # same workload (profiles/bake_c5_1024_pmc.md)
    pmc = os.path.join(ROOT, "profiles", "bake_c5_1024_pmc.md")
    # (a kernel under the table's cut, such as the dense pass when the search runs several
    # lanes a node and defers nothing, counts as none; the search and the settle must be there)
    parts = [pmc_traffic(pmc, k) for k in ("kb_search<1, float>", "kb_leaves<float>", "kb_refine_bricks", "kb_redo<1, false>", "kb_mark_bricks",
                                           "kb_settle<1>")]
    node_traffic = sum(v for v in parts if v is not None) if parts[0] is not None and parts[-1] is not None else None
    roof = None
    if c5_useful:
        flop = c5_useful["node_flop_per_band_brick"] * cm["band"]
        achieved = flop / (node_ms / 1e3) / 1e12
        roof = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s", "frac": achieved * 1e12 / PEAK_FP32,
                "traffic": node_traffic,
                "kernel": "the band nodes' search: " + " + ".join(node_stages), "kernel_ms": node_ms, "useful_flop": flop,
                "useful_flop_per_band_brick": c5_useful["node_flop_per_band_brick"],
                "hbm": {"algorithmic_bytes": 1464.0 * cm["slots"], "achieved_GBps": 1464.0 * cm["slots"] / (cm["ms"] / 1e3) / 1e9,
                        "peak_GBps": PEAK_HBM / 1e9, "frac": 1464.0 * cm["slots"] / (cm["ms"] / 1e3) / PEAK_HBM},
                "valu_measured": {k: pmc_valu(os.path.join(ROOT, "profiles", "bake_c5_1024_valu.md"), k) for k in ("kb_search<1, float>", "kb_settle<1>")}}
    out = {
        "metric": "SDF queries/s + bricks/s, FP32+FP64-refine vs FP64 at <=1e-6 rel err, 1 GPU",
        "value": cm["bricks_per_s"],
        "unit": "bricks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": cm["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded generator)",
        "config": {"workload": "c5 large mixed solid: 400 x 300 x 40 mm plate, 1,998 faces (800 rational bicubic 32 x 32 nets, weights 0.5-2; "
                               "planes, cylinders, cones, tori), loops of 12-64 cubic spans; narrow-band bake of its 1024-node level "
                               f"(cell {c5['h']:.4f} mm, {c5['level'][0]} x {c5['level'][1]} x {c5['level'][2]} bricks of 8^3 cells), "
                               "FP32+FP64-refine, batched evaluator (one part: one captured-graph launch a step, autotuned), "
                               "device-resident",
                   "bricks": cm["bricks"], "band_bricks": cm["band"], "stored_bricks": cm["slots"]},
        "fp64_bricks_per_s": cf["bricks_per_s"],
        "mixed_over_fp64": c5["batched"]["mixed_over_fp64"],
        "cast_rays_per_s": c5.get("cast", {}).get("rays_per_s"),
        "single_level_entry_point": {"mixed_bricks_per_s": c5["mixed"]["bricks_per_s"], "fp64_bricks_per_s": c5["fp64"]["bricks_per_s"],
                                     "mixed_ms": c5["mixed"]["ms"], "fp64_ms": c5["fp64"]["ms"], "mixed_over_fp64": c5["mixed_over_fp64"]},
        "queries_per_s": {"mixed": c5["queries"]["mixed"]["queries_per_s"], "fp64": c5["queries"]["fp64"]["queries_per_s"],
                          "points": c5["queries"]["points"]},
        "roofline": roof,
        "cpu_baseline": c5_cpu,
        "check": c5_check,
        "c5": c5,
        "c2_bake": {"workload": "c2 freeform NURBS-top block, 256^3 nodes at 0.5 mm (32^3 bricks), batched evaluator (one part, captured graph)",
                    "bricks": mixed["bricks"], "band_bricks": mixed["band"], "stored_bricks": mixed["slots"],
                    "mixed_bricks_per_s": mixed["band"] / (mixed["ms"] / 1e3), "fp64_bricks_per_s": fp64["band"] / (fp64["ms"] / 1e3),
                    "mixed_ms": mixed["ms"], "fp64_ms": fp64["ms"], "mixed_over_fp64": fp64["ms"] / mixed["ms"],
                    "single_level_entry_point": {"mixed": single["mixed"], "fp64": single["fp64"]},
                    "tuning": {"mixed": mixed["tuning"], "fp64": fp64["tuning"]},
                    "stages_ms": {"mixed": mixed["stages"], "fp64": fp64["stages"]},
                    "undecided_nodes": mixed["undecided"], "refined_nodes": mixed["refined"],
                    "cpu_baseline": c2_cpu, "check": check},
        "c1_queries_per_s": {"mixed": q["mixed"], "fp64": q["fp64"]},
        "c3_queries": c3,
        "c4_batch": c4,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""Benchmark: the exact signed distance's narrow-band brick bake on one MI355X.

Workload (a "step"): the freeform single solid -- a 120 x 120 mm block whose top is
a rational bicubic NURBS with a 16 x 16 control net (seeded value-noise heights,
5 mm, weights in [0.8, 1.2]) -- baked on a 256^3 node grid at 0.5 mm into 8^3-cell
bricks (32^3 bricks) within the narrow band: the decision of every brick, the signs
of the bricks kept whole, the 729 nodes of every band brick, flood fill, i16
quantization and packing into the 365-word atlas, all on the device, in the
FP32 + FP64-refine form, run by the batched evaluator (the bake's launches captured
in one hipGraph, persistent queue stages with autotuned grids; here one part). The
plain FP64 form is timed beside it, so is the single-level entry point in both forms,
and so are 2^20 scattered SDF queries on the analytic box-with-hole part and the
64-part batched workload.

Prints one JSON line (rank 0, N = 1). `value` is band bricks baked per second.
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


def pmc_traffic(path, kernel):
    """HBM bytes per launch of `kernel` from a committed counter profile (its table row's
    scaled column), or None."""
    try:
        for line in open(path):
            cells = [c.strip() for c in line.split("|")]
            if len(cells) > 5 and cells[1] == f"`{kernel}`":
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
            if len(cells) > 8 and cells[1] == f"`{kernel}`":
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
        res[name] = {"queries_per_s": n / dt, "ms": dt * 1e3, "searched_again_fp64": st.refined, "near_trim_curve": st.unsure_trim,
                     "near_tie": st.unsure_tie, "settled": st.settled, "undecided": int((outs[name][1] != 0).sum())}
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
    ap.add_argument("--profile", action="store_true", help="a short run for a profiler: bake steps only")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-part batched workload")
    ap.add_argument("--c3", action="store_true", help="add scattered queries on the trimmed plate (half at its trim curves)")
    ap.add_argument("--c3-log2", type=int, default=20, help="log2 of the trimmed plate's query count (24: the configuration's 2^24)")
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
        # a plain copy of known size (256 MiB read, 256 MiB written): the scale for the
        # profiler's HBM byte counters
        x = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
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

    # the trimmed plate's queries, half of them at its trim curves (on request)
    c3 = trimmed_plate_queries(args, log, len(os.sched_getaffinity(0))) if args.c3 else None
    # the batched workload: 64 parts of uneven face counts and grids in one graph
    c4 = None if args.no_batch else batched_workload(args, log)

    check = {}
    cpu = None
    if not args.no_check or not args.no_cpu:
        from oracle import Oracle

        o = Oracle(part)
        threads = len(os.sched_getaffinity(0))
        t = time.perf_counter()
        m_ref, a_ref, nread = o.bake(lv, threads=threads)
        cpu_s = time.perf_counter() - t
        cpu = {"value": nread / cpu_s, "unit": "bricks/s", "cores": threads, "kind": "port",
               "sample": f"the same c2 level (all {nread} band bricks of the 32^3), FP64 oracle on {threads} host threads, {cpu_s:.1f} s"}
        log(f"cpu oracle bake: {cpu_s:.1f}s, {nread / cpu_s:.1f} bricks/s on {threads} threads")
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

    for b in batches.values():
        b.close()
    mixed, fp64 = res["mixed"], res["fp64"]
    bricks_ms = mixed["stages"].get("nodes_search32", float("nan"))
    # the dominant kernel (the mixed form's node search): its algorithmic bytes per launch
    # are the single-precision node records it must write (winning feature, two
    # parameters, distance, flags: 20 B for each of the band bricks' 729 nodes) and the
    # packed tables it reads (under 2.3 MB: left out); its measured HBM traffic comes from
    # the committed FETCH_SIZE / WRITE_SIZE profile of the same kernel and workload
    algo_bytes = mixed["band"] * 729 * 20.0
    traffic = pmc_traffic(os.path.join(ROOT, "profiles", "bake_c2_256_pmc.md"), "kb_search<1, float>")
    out = {
        "metric": "SDF queries/s + bricks/s, FP32+FP64-refine vs FP64 at <=1e-6 rel err, 1 GPU",
        "value": mixed["band"] / (mixed["ms"] / 1e3),
        "unit": "bricks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded generator)",
        "config": {"workload": "c2 freeform NURBS-top block, 256^3 nodes at 0.5 mm, 8^3-cell bricks, narrow band, FP32+FP64-refine, "
                               "batched evaluator (one part, captured graph)",
                   "bricks": mixed["bricks"], "band_bricks": mixed["band"], "stored_bricks": mixed["slots"]},
        "fp64_bricks_per_s": fp64["band"] / (fp64["ms"] / 1e3),
        "mixed_over_fp64": fp64["ms"] / mixed["ms"],
        "single_level_entry_point": {"mixed": single["mixed"], "fp64": single["fp64"]},
        "tuning": {"mixed": mixed["tuning"], "fp64": fp64["tuning"]},
        "c1_queries_per_s": {"mixed": q["mixed"], "fp64": q["fp64"]},
        "stages_ms": {"mixed": mixed["stages"], "fp64": fp64["stages"]},
        "undecided_nodes": mixed["undecided"],
        "refined_nodes": mixed["refined"],
        "roofline": {"bound": "hbm", "achieved": algo_bytes / (bricks_ms / 1e3) / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                     "frac": algo_bytes / (bricks_ms / 1e3) / PEAK_HBM, "traffic": traffic, "kernel": "kb_search<1, float> (nodes_search32)",
                     "kernel_ms": bricks_ms, "algorithmic_bytes": algo_bytes,
                     "valu": pmc_valu(os.path.join(ROOT, "profiles", "bake_c2_256_valu.md"), "kb_search<1, float>")},
        "cpu_baseline": cpu,
        "check": check,
        "c3_queries": c3,
        "c4_batch": c4,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

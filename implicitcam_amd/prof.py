"""Markdown tables from rocprofv3's CSV output, for the profiles kept with the sources.

    python -m implicitcam_amd.prof stats  <kernel_stats.csv>  [top]
    python -m implicitcam_amd.prof traffic <fetch counter_collection.csv> <write counter_collection.csv> <kernel_stats.csv> [top]
    python -m implicitcam_amd.prof valu <flops counter_collection.csv> <utilisation counter_collection.csv> <kernel_stats.csv> [top]

- stats: a kernel trace's summary (`rocprofv3 --kernel-trace --stats --output-format
  csv`): calls, average and total duration, share.
- traffic: HBM bytes per launch from two counter passes (`--pmc FETCH_SIZE`, then `--pmc
  WRITE_SIZE`), each scaled by what the same counter reads for the run's plain copy of
  256 MiB (`bench.py --profile` ends with one: `__amd_rocclr_copyBuffer` or torch's copy
  kernel), over the kernel trace's average duration.
- valu: floating-point rates per kernel from one pass of the VALU instruction counters
  (SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F{32,64}) and one of the utilisation counters
  (SQ_ACTIVE_INST_VALU, SQ_THREAD_CYCLES_VALU): FLOP = 64 x (ADD + MUL + TRANS + 2 FMA)
  instructions (whole waves: inactive lanes are not subtracted, so these bound the rate
  from above), lane use = SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU).
"""
import csv
import re
import sys
from collections import defaultdict

COPY_BYTES = 256 << 20
PEAK_FP32, PEAK_FP64 = 157.3, 78.6  # TFLOP/s, MI355X dense vector


def short(name):
    """A kernel's name without the namespace and argument list: `k_search<1, float>`."""
    n = name.replace("(anonymous namespace)::", "")
    n = re.sub(r"^void ", "", n)
    depth, out = 0, []
    for ch in n:
        if ch == "(" and depth == 0:
            break
        depth += ch == "<"
        depth -= ch == ">"
        out.append(ch)
    return "".join(out).strip()


def read_stats(path):
    """{kernel: (calls, average ns, total ns, percent)} from a kernel_stats.csv."""
    acc = defaultdict(lambda: [0, 0.0, 0.0])
    with open(path) as f:
        for row in csv.DictReader(f):
            a = acc[short(row["Name"])]
            a[0] += int(row["Calls"])
            a[1] += float(row["TotalDurationNs"])
            a[2] += float(row["Percentage"])
    return {k: (c, t / c, t, p) for k, (c, t, p) in acc.items()}


def read_counters(path):
    """{kernel: {counter: [value per dispatch]}} from a counter_collection.csv."""
    out = defaultdict(lambda: defaultdict(list))
    with open(path) as f:
        for row in csv.DictReader(f):
            out[short(row["Kernel_Name"])][row["Counter_Name"]].append(float(row["Counter_Value"]))
    return out


def mean(v):
    return sum(v) / len(v) if v else 0.0


def stats_table(stats, top=20):
    rows = sorted(stats.items(), key=lambda kv: -kv[1][2])[:top]
    lines = ["| kernel | calls | average (ms) | total (ms) | share |", "| --- | ---: | ---: | ---: | ---: |"]
    for k, (calls, avg, total, pct) in rows:
        lines.append(f"| `{k}` | {calls} | {avg / 1e6:.3f} | {total / 1e6:.2f} | {pct:.1f} % |")
    return "\n".join(lines)


def copy_kernel(counters):
    for k in counters:
        if "copyBuffer" in k or "copy" in k.lower():
            return k
    return None


def traffic_table(fetch_csv, write_csv, stats_csv, top=15):
    fetch, write, stats = read_counters(fetch_csv), read_counters(write_csv), read_stats(stats_csv)
    ck_f, ck_w = copy_kernel(fetch), copy_kernel(write)
    # the copy kernel's largest dispatch is the 256 MiB copy (the others copy results
    # back to the host, a few bytes each)
    cf = max(fetch[ck_f]["FETCH_SIZE"]) if ck_f else float("nan")
    cw = max(write[ck_w]["WRITE_SIZE"]) if ck_w else float("nan")
    sf = COPY_BYTES / 1024 / cf if ck_f else 1.0
    sw = COPY_BYTES / 1024 / cw if ck_w else 1.0
    lines = [f"Calibration: the 256 MiB copy (`{ck_f}`, its largest dispatch) reads FETCH_SIZE {cf:,.0f} KiB and WRITE_SIZE {cw:,.0f} KiB: "
             f"scale factors {sf:.3f} (fetch) and {sw:.3f} (write).", "",
             "| kernel | dispatches | FETCH_SIZE (KiB) | WRITE_SIZE (KiB) | HBM bytes per launch (scaled) | average (ms) | GB/s |",
             "| --- | ---: | ---: | ---: | ---: | ---: | ---: |"]
    rows = sorted(stats.items(), key=lambda kv: -kv[1][2])[:top]
    for k, (calls, avg, total, pct) in rows:
        if k not in fetch or k not in write:
            continue
        fk, wk = mean(fetch[k]["FETCH_SIZE"]), mean(write[k]["WRITE_SIZE"])
        b = (fk * sf + wk * sw) * 1024
        lines.append(f"| `{k}` | {len(fetch[k]['FETCH_SIZE'])} | {fk:,.1f} | {wk:,.1f} | {b:,.0f} | {avg / 1e6:.3f} | {b / avg:,.0f} |")
    return "\n".join(lines)


def valu_table(flops_csv, util_csv, stats_csv, top=15):
    fl, ut, stats = read_counters(flops_csv), read_counters(util_csv), read_stats(stats_csv)
    lines = ["| kernel | average (ms) | FP32 GFLOP | FP32 TFLOP/s | FP64 GFLOP | FP64 TFLOP/s | lane use | FP32 share of peak | FP64 share of peak |",
             "| --- | ---: | ---: | ---: | ---: | ---: | ---: | ---: | ---: |"]
    rows = sorted(stats.items(), key=lambda kv: -kv[1][2])[:top]
    for k, (calls, avg, total, pct) in rows:
        if k not in fl:
            continue
        c = {n: mean(v) for n, v in fl[k].items()}

        def flop(p):
            return 64.0 * (c.get(f"SQ_INSTS_VALU_ADD_{p}", 0) + c.get(f"SQ_INSTS_VALU_MUL_{p}", 0) + c.get(f"SQ_INSTS_VALU_TRANS_{p}", 0)
                           + 2 * c.get(f"SQ_INSTS_VALU_FMA_{p}", 0))

        f32, f64 = flop("F32"), flop("F64")
        u = {n: mean(v) for n, v in ut.get(k, {}).items()}
        lane = u.get("SQ_THREAD_CYCLES_VALU", 0) / (64 * u["SQ_ACTIVE_INST_VALU"]) if u.get("SQ_ACTIVE_INST_VALU") else float("nan")
        r32, r64 = f32 / avg / 1e3, f64 / avg / 1e3  # TFLOP/s (FLOP / ns / 1e3)
        lines.append(f"| `{k}` | {avg / 1e6:.3f} | {f32 / 1e9:.3f} | {r32:.2f} | {f64 / 1e9:.3f} | {r64:.2f} | {lane:.2f} | "
                     f"{r32 / PEAK_FP32 * 100:.1f} % | {r64 / PEAK_FP64 * 100:.1f} % |")
    return "\n".join(lines)


def main(argv):
    if len(argv) < 2:
        print(__doc__)
        return 2
    what = argv[1]
    if what == "stats":
        print(stats_table(read_stats(argv[2]), int(argv[3]) if len(argv) > 3 else 20))
    elif what == "traffic":
        print(traffic_table(argv[2], argv[3], argv[4], int(argv[5]) if len(argv) > 5 else 15))
    elif what == "valu":
        print(valu_table(argv[2], argv[3], argv[4], int(argv[5]) if len(argv) > 5 else 15))
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))

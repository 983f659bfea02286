"""Per-kernel roofline of one mixed-precision AMG-FCG solve from a rocprofv3
kernel trace.

    rocprofv3 --kernel-trace --stats -d OUT -o run --output-format csv -- \\
        python3 bench.py --steps 1 --warmup 1 --mixed-only > bench.json
    python -m unap_amd.kernel_report OUT/run_kernel_trace.csv bench.json

With the trace alone (python -m unap_amd.kernel_report TRACE), the last
solve's kernels with their launch counts and times only (any solver).

The last solve in the trace starts at its kStateInit dispatch.  For every
kernel name in it: launches, total and mean time, algorithmic bytes (each
distinct array a launch reads or writes, counted once at its unpadded length,
summed over the hierarchy levels the kernel runs on) and the resulting GB/s
against the 8 TB/s HBM peak.  Level sizes come from bench.py's "hierarchy"
field (level 0 = the finest matrix).
"""
import collections
import csv
import json
import re
import sys

PEAK = 8000.0  # GB/s


def _levels(h):
    return [(int(n), int(f)) for n, f in h]


def alg_bytes(kernel, lv, iters, launches=None):
    """Algorithmic bytes over the whole solve for one kernel name, or None
    when the kernel moves no per-cell data (finalizers, state updates).
    Kernels that do not run every iteration are counted per launch."""
    n0, f0 = lv[0]
    L = len(lv) - 1              # coarse levels; the last is the coarsest
    smoothed = lv[:L]            # finest + coarse levels 0 .. L-2
    s = 4                        # FP32 hierarchy
    k = kernel
    if "BCheby<float>" in k:     # 4 sweeps a level a V-cycle
        return iters * 4 * sum((6 * s + 4) * n + (2 * s + 8) * f for n, f in smoothed)
    if "BSfAmul<float>" in k:    # coarse levels 0 .. L-3 and the finest
        return iters * sum((4 * s + 4) * n + (2 * s + 8) * f for n, f in [lv[0]] + lv[1:L - 1])
    if "BJacobiScale<float>" in k:
        return iters * sum(5 * s * n for n, _ in lv[1:L - 1])
    if "BPsiScaled<float>" in k:  # fixed cycle: psi is written, not read (reads fc, fres, Apsi, d; writes fc, psi)
        return iters * 6 * s * n0
    if "BRestrict<float, float>" in k:
        return iters * sum((s + 4) * (lv[i][0] + lv[i + 1][0]) for i in range(L))
    if "BProlong<float>" in k:
        return iters * sum((s + 4) * lv[i][0] + s * lv[i + 1][0] for i in range(L))
    if "BCopy<double, float>" in k:  # the FP32 source cast, when the solver does not supply it
        return (iters if launches is None else launches) * 12 * n0
    if "BFill<float>" in k:
        return iters * 4 * n0
    if "kDenseSolve" in k:
        nc = lv[-1][0]
        return iters * (8 * nc * nc + 8 * nc)
    if "BFcgRz<float>" in k:      # r (FP64), z (FP32), q = A p (FP64)
        return iters * 20 * n0
    if "BFcgP<float>" in k:
        return iters * 20 * n0
    if "BAmulDot" in k:
        return iters * (28 * n0 + 24 * f0)
    if "BPcgXR" in k:
        return iters * 48 * n0
    if "BResidual<double>" in k:
        return 36 * n0 + 24 * f0
    if "BGather<double, double>" in k or "BScatter<double, double>" in k:
        return 20 * n0
    return None


def normalize(name):
    """One name for the load-shape instantiations of a kernel body
    (BCheby<float, 0> and BCheby<float, 1> -> BCheby<float>)."""
    return re.sub(r"<(float|double), \d+>", r"<\1>", name)


def short(name):
    name = name.replace("unap_hip::", "")
    m = re.match(r"void (\w+)<(.*)>\(", name)
    return f"{m.group(1)}<{m.group(2)}>" if m else name.split("(")[0]


def report(trace_csv, bench_json):
    rows = sorted(csv.DictReader(open(trace_csv)), key=lambda r: int(r["Start_Timestamp"]))
    for r in rows:
        r["Kernel_Name"] = normalize(r["Kernel_Name"])
    starts = [i for i, r in enumerate(rows) if "kStateInit" in r["Kernel_Name"]]
    if not starts:
        raise SystemExit("no kStateInit dispatch in the trace")
    seg = rows[starts[-1]:]
    # bench.py prints one JSON line; take the last one in the file
    info = json.loads([ln for ln in open(bench_json) if ln.startswith("{")][-1])
    lv = _levels(info["hierarchy"])
    iters = int(info["iterations"])
    t0, t1 = int(seg[0]["Start_Timestamp"]), int(seg[-1]["End_Timestamp"])
    agg = collections.defaultdict(lambda: [0, 0])
    for r in seg:
        d = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        a = agg[r["Kernel_Name"]]
        a[0] += 1
        a[1] += d
    busy = sum(v[1] for v in agg.values())
    out = [f"solve span {(t1 - t0) / 1e6:.1f} ms, kernel busy {busy / 1e6:.1f} ms, "
           f"{len(seg)} dispatches, {iters} iterations, {len(lv) - 1} coarse levels",
           "",
           "| kernel | launches | total ms | mean us | alg. GB | GB/s | of 8 TB/s |",
           "|---|---|---|---|---|---|---|"]
    tot_b = 0.0
    for name, (c, d) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        b = alg_bytes(name, lv, iters, c)
        if b is None:
            out.append(f"| {short(name)} | {c} | {d / 1e6:.2f} | {d / c / 1e3:.1f} | - | - | - |")
            continue
        tot_b += b
        gbs = b / (d * 1e-9) / 1e9
        out.append(f"| {short(name)} | {c} | {d / 1e6:.2f} | {d / c / 1e3:.1f} | {b / 1e9:.2f} | "
                   f"{gbs:.0f} | {gbs / PEAK:.0%} |")
    out.append("")
    out.append(f"all kernels: {tot_b / 1e9:.1f} GB algorithmic in {busy / 1e6:.1f} ms busy = "
               f"{tot_b / (busy * 1e-9) / 1e9:.0f} GB/s")
    out.append("")
    out += per_level(seg, lv)
    return "\n".join(out)


def _by_position(seg, pattern, order):
    """Mean duration (ns) per level of the launches whose name contains
    `pattern`, given the level of each launch within one V-cycle."""
    ds = [int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in seg if pattern in r["Kernel_Name"]]
    acc = collections.defaultdict(list)
    for i, d in enumerate(ds):
        acc[order[i % len(order)]].append(d)
    return {lvl: sum(v) / len(v) for lvl, v in acc.items()}


def per_level(seg, lv):
    """Per-level means from the fixed launch order of a V-cycle: Chebyshev
    sweeps go coarse to fine (4 a level, the coarsest level unsmoothed),
    restrictions fine to coarse, prolongations coarse to fine."""
    L = len(lv) - 1
    cheb = _by_position(seg, "BCheby<float>", [lvl for lvl in range(L - 1, 0, -1) for _ in range(4)] + [0] * 4)
    rst = _by_position(seg, "BRestrict<float, float>", list(range(L)))
    pro = _by_position(seg, "BProlong<float>", list(range(L - 1, -1, -1)))
    out = ["| level | cells | entries/row | Chebyshev us | Chebyshev GB/s | restrict to next us | "
           "prolong from next us |",
           "|---|---|---|---|---|---|---|"]
    for lvl in range(L + 1):
        n, f = lv[lvl]
        c = cheb.get(lvl)
        cs = f"{c / 1e3:.1f} | {(28 * n + 16 * f) / c:.0f}" if c else "- | -"
        r = f"{rst[lvl] / 1e3:.1f}" if lvl in rst else "-"
        p = f"{pro[lvl] / 1e3:.1f}" if lvl in pro else "-"
        out.append(f"| {lvl} | {n} | {2 * f / n:.2f} | {cs} | {r} | {p} |")
    return out


def time_table(trace_csv):
    """Launches, total and mean time per kernel of the last solve in a trace."""
    rows = sorted(csv.DictReader(open(trace_csv)), key=lambda r: int(r["Start_Timestamp"]))
    starts = [i for i, r in enumerate(rows) if "kStateInit" in r["Kernel_Name"]]
    if not starts:
        raise SystemExit("no kStateInit dispatch in the trace")
    seg = rows[starts[-1]:]
    agg = collections.defaultdict(lambda: [0, 0])
    for r in seg:
        a = agg[normalize(r["Kernel_Name"])]
        a[0] += 1
        a[1] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    busy = sum(v[1] for v in agg.values())
    t0, t1 = int(seg[0]["Start_Timestamp"]), int(seg[-1]["End_Timestamp"])
    out = [f"solve span {(t1 - t0) / 1e6:.1f} ms, kernel busy {busy / 1e6:.1f} ms, {len(seg)} dispatches", "",
           "| kernel | launches | total ms | mean us | % of busy |", "|---|---|---|---|---|"]
    for name, (c, d) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        out.append(f"| {short(name)} | {c} | {d / 1e6:.2f} | {d / c / 1e3:.1f} | {100 * d / busy:.1f} |")
    return "\n".join(out)


if __name__ == "__main__":
    if len(sys.argv) == 2:
        print(time_table(sys.argv[1]))
    elif len(sys.argv) == 3:
        print(report(sys.argv[1], sys.argv[2]))
    else:
        raise SystemExit(__doc__)

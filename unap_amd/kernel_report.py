"""Per-kernel roofline of one mixed-precision AMG-FCG solve from a rocprofv3
kernel trace.

    rocprofv3 --kernel-trace --stats -d OUT -o run --output-format csv -- \\
        python3 bench.py --steps 1 --warmup 1 --mixed-only > bench.json
    python -m unap_amd.kernel_report OUT/run_kernel_trace.csv bench.json

With the trace alone (python -m unap_amd.kernel_report TRACE), the last
solve's kernels with their launch counts and times only (any solver).

The last solve in the trace starts at the kStateInit before its
initial-residual finalizer.  For every
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


def solve_start(rows):
    """Index of the last solve's first dispatch: the kStateInit right before
    its initial-residual finalizer (eigen estimates inside a solve launch
    kStateInit too)."""
    fin = [i for i, r in enumerate(rows) if "FinInitRes" in r["Kernel_Name"]]
    i = fin[-1] if fin else len(rows) - 1  # without it: the last kStateInit
    while i >= 0 and "kStateInit" not in rows[i]["Kernel_Name"]:
        i -= 1
    if i < 0:
        raise SystemExit("no kStateInit before the last solve")
    return i


def normalize(name):
    """One name for the load-shape instantiations of a kernel body
    (BCheby<float, 0> and BCheby<float, 1> -> BCheby<float>) and for its two
    row walks (kRowsXcd, the XCD-slab walk of the large levels -> kRows)."""
    return re.sub(r"<(float|double), \d+>", r"<\1>", name).replace("kRowsXcd<", "kRows<")


def short(name):
    name = name.replace("unap_hip::", "")
    m = re.match(r"void (\w+)<(.*)>\(", name)
    return f"{m.group(1)}<{m.group(2)}>" if m else name.split("(")[0]


def report(trace_csv, bench_json):
    rows = sorted(csv.DictReader(open(trace_csv)), key=lambda r: int(r["Start_Timestamp"]))
    for r in rows:
        r["Kernel_Name"] = normalize(r["Kernel_Name"])
    seg = rows[solve_start(rows):]
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
    out += per_level(seg, lv, iters, int(info.get("pair_steps", 2)))
    return "\n".join(out)


def _by_position(seg, pattern, order):
    """Mean duration (ns) per level of the launches whose name contains
    `pattern`, given the level of each launch within one V-cycle."""
    ds = [int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in seg if pattern in r["Kernel_Name"]]
    acc = collections.defaultdict(list)
    for i, d in enumerate(ds):
        acc[order[i % len(order)]].append(d)
    return {lvl: sum(v) / len(v) for lvl, v in acc.items()}


def restrict_order(seg, L, iters, pair_steps):
    """The level each restriction launch of one V-cycle fills: with pair
    steps above 2 a level is reached in sub-steps of two pairwise levels
    (ceil(pair_steps / 2) launches; the last level takes the rest), else one
    launch a level."""
    r = sum(1 for x in seg if "BRestrict<float, float>" in x["Kernel_Name"]) // max(1, iters)
    per = [1] * L if pair_steps <= 2 else [(pair_steps + 1) // 2] * L
    per[-1] = max(1, r - sum(per[:-1]))
    return [lvl for lvl in range(L) for _ in range(per[lvl])]


def _by_group(seg, pattern, order):
    """Mean duration (ns) per V-cycle and level of the launches whose name
    contains `pattern`, summed over the launches a level takes (order: the
    level of each launch within one V-cycle)."""
    ds = [int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in seg if pattern in r["Kernel_Name"]]
    acc = collections.defaultdict(float)
    cycles = len(ds) // len(order)
    for i, d in enumerate(ds[:cycles * len(order)]):
        acc[order[i % len(order)]] += d
    return {lvl: v / cycles for lvl, v in acc.items()} if cycles else {}


def per_level(seg, lv, iters=1, pair_steps=2):
    """Per-level means from the fixed launch order of a V-cycle: Chebyshev
    sweeps go coarse to fine (4 a level, the coarsest level unsmoothed),
    restrictions fine to coarse (all the launches into a level summed),
    prolongations coarse to fine."""
    L = len(lv) - 1
    cheb = _by_position(seg, "BCheby<float>", [lvl for lvl in range(L - 1, 0, -1) for _ in range(4)] + [0] * 4)
    rst = _by_group(seg, "BRestrict<float, float>", restrict_order(seg, L, iters, pair_steps))
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


def report64(trace_csv, bench_json):
    """The FP64 drop-in solve (PCG around the reference V-cycle, bench.py
    --fp64-only): per-kernel algorithmic bytes and GB/s, with the V-cycle
    kernels attributed to levels by their fixed launch order."""
    rows = sorted(csv.DictReader(open(trace_csv)), key=lambda r: int(r["Start_Timestamp"]))
    for r in rows:
        r["Kernel_Name"] = normalize(r["Kernel_Name"])
    seg = rows[solve_start(rows):]
    info = json.loads([ln for ln in open(bench_json) if ln.startswith("{")][-1])
    lv = _levels(info["hierarchy"])
    npost, growth, nfin = info["n_post_sweeps"], info["level_sweep_growth"], info["n_finest_sweeps"]
    L = len(lv) - 1                       # coarse levels; coarse level l is lv[l + 1]
    n0, f0 = lv[0]
    s = 8

    def nf(l):                           # (cells, faces) of coarse level l, "f" the finest
        return lv[0] if l == "f" else lv[l + 1]

    def fine_of(l):
        return lv[0] if l == 0 else lv[l]

    # launch order of one V-cycle (Vcycle.cpp; the first cycle estimates
    # eigenvalues instead of Chebyshev sweeps)
    order = {
        "BCheby<double>": [l for l in range(L - 2, -1, -1) for _ in range(npost + growth * l)] + ["f"] * nfin,
        "BSfAmul<double>": list(range(L - 3, -1, -1)) + ["f"],
        "BJacobiScale<double>": list(range(L - 3, -1, -1)),
        "BRestrict<double, double>": list(range(L)),
        "BProlong<double>": list(range(L - 2, -1, -1)) + ["f"],
    }
    per_launch = {
        "BCheby<double>": lambda l: (6 * s + 4) * nf(l)[0] + (2 * s + 8) * nf(l)[1],
        "BSfAmul<double>": lambda l: (4 * s + 4) * nf(l)[0] + (2 * s + 8) * nf(l)[1],
        "BJacobiScale<double>": lambda l: 5 * s * nf(l)[0],
        "BRestrict<double, double>": lambda l: (s + 4) * (fine_of(l)[0] + lv[l + 1][0]),
        "BProlong<double>": lambda l: (s + 4) * nf(l)[0] + s * (lv[1][0] if l == "f" else lv[l + 2][0]),
    }
    flat = {"BPsiScaled<double>": 7 * s * n0, "BFill<double>": s * n0, "BDot": 2 * s * n0, "BPcgP": 3 * s * n0,
            "BAmulDot": 28 * n0 + 24 * f0, "BPcgXR": 48 * n0, "BResidual<double>": 36 * n0 + 24 * f0}
    agg = collections.defaultdict(lambda: [0, 0, 0.0])
    lvl = collections.defaultdict(lambda: [0, 0, 0.0])  # (kernel, level) -> launches, ns, bytes
    seen = collections.Counter()
    for r in seg:
        name = r["Kernel_Name"]
        d = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        a = agg[name]
        a[0] += 1
        a[1] += d
        key = next((k for k in order if k in name), None)
        if key:
            pat = order[key]
            l = pat[seen[key] % len(pat)]
            seen[key] += 1
            b = per_launch[key](l)
            a[2] += b
            e = lvl[(key, l)]
            e[0] += 1
            e[1] += d
            e[2] += b
            continue
        n2 = name.replace("unap_hip::", "")
        fk = next((k for k in flat if any(("<" + k + sep) in n2 for sep in (">", ",", " >"))), None)
        if fk:
            a[2] += flat[fk]
    busy = sum(v[1] for v in agg.values())
    t0, t1 = int(seg[0]["Start_Timestamp"]), int(seg[-1]["End_Timestamp"])
    out = [f"solve span {(t1 - t0) / 1e6:.1f} ms, kernel busy {busy / 1e6:.1f} ms, {len(seg)} dispatches, "
           f"{info['iterations']} iterations, {L} coarse levels", "",
           "| kernel | launches | total ms | mean us | alg. GB | GB/s | of 8 TB/s |", "|---|---|---|---|---|---|---|"]
    tot_b = 0.0
    for name, (c, d, b) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        if b <= 0:
            out.append(f"| {short(name)} | {c} | {d / 1e6:.2f} | {d / c / 1e3:.1f} | - | - | - |")
            continue
        tot_b += b
        gbs = b / (d * 1e-9) / 1e9
        out.append(f"| {short(name)} | {c} | {d / 1e6:.2f} | {d / c / 1e3:.1f} | {b / 1e9:.2f} | {gbs:.0f} | "
                   f"{gbs / PEAK:.0%} |")
    out += ["", f"kernels with a byte count: {tot_b / 1e9:.1f} GB algorithmic in {busy / 1e6:.1f} ms busy "
            f"(all kernels) = {tot_b / (busy * 1e-9) / 1e9:.0f} GB/s", "",
            "| level | cells | Chebyshev us (GB/s) | restrict into us (GB/s) | prolong into us (GB/s) |",
            "|---|---|---|---|---|"]
    for l in ["f"] + list(range(L)):
        cells = nf(l)[0]
        cols = []
        for key in ("BCheby<double>", "BRestrict<double, double>", "BProlong<double>"):
            e = lvl.get((key, l))
            cols.append(f"{e[1] / e[0] / 1e3:.1f} ({e[2] / (e[1] * 1e-9) / 1e9:.0f})" if e else "-")
        out.append(f"| {'finest' if l == 'f' else l} | {cells} | " + " | ".join(cols) + " |")
    return "\n".join(out)


def _counter_rows(path):
    rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Dispatch_Id"]))
    for r in rows:
        r["Kernel_Name"] = normalize(r["Kernel_Name"])
        r["v"] = float(r["Counter_Value"]) * 1024.0  # KiB per dispatch
    return rows


def _copy_scale(fetch_csv, write_csv, n=16777216, s=4):
    """Bytes per counted byte, from the plain copy y = x over n FP32 values
    (reads and writes exactly s * n bytes): mean of the later half of its
    launches (warm caches)."""
    out = []
    for path in (fetch_csv, write_csv):
        v = [r["v"] for r in _counter_rows(path) if "BCopy<float, float>" in r["Kernel_Name"]]
        v = v[len(v) // 2:]
        out.append(s * n / (sum(v) / len(v)))
    return out


def traffic(fetch_csv, write_csv, copy_fetch_csv, copy_write_csv, bench_json):
    """HBM traffic (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, one pass each)
    of the kernels of one mixed solve (bench.py --mixed-only), per kernel and
    level, against their algorithmic bytes."""
    kf, kw = _copy_scale(copy_fetch_csv, copy_write_csv)
    info = json.loads([ln for ln in open(bench_json) if ln.startswith("{")][-1])
    lv = _levels(info["hierarchy"])
    L = len(lv) - 1
    rows0 = _counter_rows(fetch_csv)
    rorder = restrict_order(rows0[solve_start(rows0):], L, int(info["iterations"]), int(info.get("pair_steps", 2)))
    orders = {"BCheby<float>": [lvl for lvl in range(L - 1, 0, -1) for _ in range(4)] + [0] * 4,
              "BRestrict<float, float>": rorder,
              "BProlong<float>": list(range(L - 1, -1, -1)),
              "BAmulDot": [0]}
    alg = {"BCheby<float>": lambda l: 28 * lv[l][0] + 16 * lv[l][1],
           "BRestrict<float, float>": lambda l: 8 * (lv[l][0] + lv[l + 1][0]),
           "BProlong<float>": lambda l: 8 * lv[l][0] + 4 * lv[l + 1][0],
           "BAmulDot": lambda l: 28 * lv[0][0] + 24 * lv[0][1]}
    labels = {"BCheby<float>": "Chebyshev sweep (FP32)", "BRestrict<float, float>": "restriction to level l+1",
              "BProlong<float>": "prolongation from level l+1", "BAmulDot": "FCG Amul + dot (FP64)"}
    acc = collections.defaultdict(lambda: [0, 0.0, 0.0])
    for path, col in ((fetch_csv, 1), (write_csv, 2)):
        rows = _counter_rows(path)
        seg = rows[solve_start(rows):]
        seen = collections.Counter()
        for r in seg:
            key = next((k for k in orders if k in r["Kernel_Name"]), None)
            if not key:
                continue
            pat = orders[key]
            lvl = pat[seen[key] % len(pat)]
            seen[key] += 1
            a = acc[(key, lvl)]
            if col == 1:
                a[0] += 1
            a[col] += r["v"] * (kf if col == 1 else kw)
    out = [f"copy-kernel scale: FETCH_SIZE x {kf:.4f}, WRITE_SIZE x {kw:.4f}", "",
           "| kernel | level | cells | launches | read MB | written MB | traffic MB | algorithmic MB | traffic / alg. |",
           "|---|---|---|---|---|---|---|---|---|"]
    for (key, lvl), (c, rd, wr) in sorted(acc.items(), key=lambda kv: (list(orders).index(kv[0][0]), kv[0][1])):
        b = alg[key](lvl)
        # a restriction into a level may take several launches (sub-steps):
        # its row is per restriction; the other rows are per launch
        k = c / orders[key].count(lvl) if key.startswith("BRestrict") else c
        t = (rd + wr) / k
        out.append(f"| {labels[key]} | {lvl} | {lv[lvl][0]} | {c} | {rd / k / 1e6:.1f} | {wr / k / 1e6:.1f} | "
                   f"{t / 1e6:.1f} | {b / 1e6:.1f} | {t / b:.2f} |")
    return "\n".join(out)


def sweep_traffic(fetch_csv, write_csv, copy_fetch_csv, copy_write_csv, kind, n=16777216, f=50135040):
    """HBM traffic of one finest-level FP64 sweep application (bench.py
    --kernel-only dic|gs): the sum over its level launches, mean of the later
    half of the applications, against its algorithmic bytes."""
    kf, kw = _copy_scale(copy_fetch_csv, copy_write_csv)
    per = {"dic": (34, 60 * n + 24 * f), "gs": (18, 36 * n + 24 * f)}
    nd, alg_b = per[kind]
    tot = []
    for path, k in ((fetch_csv, kf), (write_csv, kw)):
        v = [r["v"] * k for r in _counter_rows(path) if "kSweepLevel" in r["Kernel_Name"] or
             "kSweepTail" in r["Kernel_Name"]]
        apps = [sum(v[i:i + nd]) for i in range(0, len(v) - nd + 1, nd)]
        apps = apps[len(apps) // 2:]
        tot.append(sum(apps) / len(apps))
    t = tot[0] + tot[1]
    return (f"| {kind.upper()} apply (FP64, finest, all levels) | {tot[0] / 1e6:.1f} | {tot[1] / 1e6:.1f} | "
            f"{t / 1e6:.1f} | {alg_b / 1e6:.1f} | {t / alg_b:.2f} |")


def time_table(trace_csv):
    """Launches, total and mean time per kernel of the last solve in a trace."""
    rows = sorted(csv.DictReader(open(trace_csv)), key=lambda r: int(r["Start_Timestamp"]))
    seg = rows[solve_start(rows):]
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
    if len(sys.argv) >= 2 and sys.argv[1] == "--traffic":
        # --traffic FETCH.csv WRITE.csv COPY_FETCH.csv COPY_WRITE.csv bench.json
        print(traffic(*sys.argv[2:7]))
    elif len(sys.argv) >= 2 and sys.argv[1] == "--sweep-traffic":
        # --sweep-traffic dic|gs FETCH.csv WRITE.csv COPY_FETCH.csv COPY_WRITE.csv
        print(sweep_traffic(*sys.argv[3:7], sys.argv[2]))
    elif len(sys.argv) == 2:
        print(time_table(sys.argv[1]))
    elif len(sys.argv) == 3:
        info = json.loads([ln for ln in open(sys.argv[2]) if ln.startswith("{")][-1])
        print(report64(sys.argv[1], sys.argv[2]) if "fp64_s" in info else report(sys.argv[1], sys.argv[2]))
    else:
        raise SystemExit(__doc__)

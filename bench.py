"""Benchmark of the solve path on one MI355X.

Workload (C4): a 256^3 (16.8M cells, 50.1M faces) randomly relabelled
variable-coefficient Poisson system in LDU form, face coefficients
log-uniform in [1e-4, 1], solved by AMG-preconditioned CG to relative
residual 1e-8 from x = 0.

  * mixed (the headline `value`): flexible CG in FP64 (true residual,
    dot products and norms in FP64) around one V-cycle of the agglomeration
    AMG in FP32 (every second level of the reference's pairwise hierarchy,
    Chebyshev(4) smoothers on every level with estimates taken at setup over
    the interval [lambda / 5, lambda], restriction and prolongation in FP32;
    coarsest level solved in FP64 by a dense inverse);
  * fp64 drop-in: the reference's algorithm in FP64 on the GPU (PCG around
    MGSolver V-cycles over the full pairwise hierarchy with Chebyshev
    smoothers, nPost = nFinest = 4 with the reference's nPost + level
    schedule, coarsest level by PCG+DIC to 1e-2);
  * fp64 same algorithm: the mixed path's algorithm with the hierarchy in
    FP64, so that the mixed/FP64 ratio of precision alone is reported too.

Inputs are resident in HBM before the timed region; each step is one full
solve.  Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "AMG-PCG time to rel.res 1e-8 (256^3), mixed vs FP64; HBM GB/s"
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--n-coarsest", type=int, default=500, help="coarsest-level threshold of the FP64 drop-in")
    ap.add_argument("--mixed-n-coarsest", type=int, default=4000,
                    help="coarsest-level threshold of the mixed algorithm (dense FP64 inverse)")
    ap.add_argument("--pair-steps", type=int, default=7, help="pairwise steps per level of the mixed hierarchy")
    ap.add_argument("--mixed-cheby-ratio", type=float, default=3.0,
                    help="Chebyshev interval of the mixed cycle's coarse levels: lower end lambda / ratio "
                         "(the reference: 1.5 on every level)")
    ap.add_argument("--mixed-cheby-ratio-finest", type=float, default=12.0,
                    help="the same on the finest level (0: --mixed-cheby-ratio)")
    ap.add_argument("--mixed-cheby-boost", type=float, default=1.0,
                    help="Chebyshev interval of the mixed cycle: upper end lambda * boost (the reference: 1.1)")
    ap.add_argument("--mixed-scaling", type=int, default=0,
                    help="correction scaling of the mixed cycle: 0 none (FCG's step length already scales each "
                         "correction), 1 coarse levels, 2 every level but the two coarsest (the reference)")
    ap.add_argument("--fp64-steps", type=int, default=3,
                    help="timed solves of each FP64 leg (median; one untimed warm-up solve before them)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--cpu-size", type=int, default=None, help="edge length of the CPU sample (default: --size)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "hbm_traffic_256.json"),
                    help="JSON with the HBM bytes per launch of the dominant kernel, measured by rocprofv3 --pmc")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-systems section")
    ap.add_argument("--no-configs", action="store_true", help="skip the C2 / C3 FP64 drop-in solves")
    ap.add_argument("--config-size", type=int, default=128, help="edge length of the C2 / C3 systems")
    ap.add_argument("--configs", default="c1,c2,c3", help="which drop-in configurations to time (c1, c2, c3)")
    ap.add_argument("--configs-only", action="store_true",
                    help="time only the drop-in configurations (for kernel traces of exactly those solves)")
    ap.add_argument("--configs-fp64-only", action="store_true",
                    help="with --configs-only: skip the FP32-preconditioner and FP64-reference solves, so the "
                         "last solve in a kernel trace is a timed FP64 one")
    ap.add_argument("--batch-launches", type=int, default=20)
    ap.add_argument("--batch-systems", type=int, default=512)
    ap.add_argument("--kernel-only", default=None,
                    help="only time one finest-level kernel (amul, cheby, gs, dic, residual, axpy_norm, dot, copy) "
                         "on the 256^3 matrix, for hardware-counter runs")
    ap.add_argument("--precision", type=int, default=32, help="precision of --kernel-only")
    ap.add_argument("--reps", type=int, default=5, help="launches of --kernel-only")
    ap.add_argument("--mixed-only", action="store_true",
                    help="stop after the timed mixed solves (for kernel traces of exactly those solves)")
    ap.add_argument("--fp64-only", action="store_true",
                    help="time only the FP64 drop-in solve (for kernel traces of exactly that solve)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU design is "
              "documented with the library (sharding over point-to-point xGMI) and not run here.", file=sys.stderr)
        sys.exit(2)

    from unap_amd import hip, meshgen

    dev_name, cus = hip.device_info()
    log("device:", dev_name, cus, "CUs")
    if args.configs_only:
        print(json.dumps({"configs": config_legs(args.config_size, not args.no_cpu, args.configs.split(","),
                                                 with_mixed=not args.configs_fp64_only)}), flush=True)
        return
    N = args.size
    if args.kernel_only:
        s = meshgen.config_system(4, N)
        A = hip.LduMatrix.from_system(s)
        ms, by = hip.profile_kernel(A, args.kernel_only, args.precision, reps=args.reps)
        print(json.dumps({"kernel": args.kernel_only, "precision": args.precision, "n": s.n, "n_faces": s.nfaces,
                          "reps": args.reps, "ms": ms, "alg_bytes": by, "GBps": by / (ms * 1e-3) / 1e9}), flush=True)
        return
    t0 = time.time()
    s = meshgen.config_system(4, N)
    t_gen = time.time() - t0
    log(f"generated {N}^3: n={s.n} faces={s.nfaces} in {t_gen:.1f}s")

    t0 = time.time()
    A = hip.LduMatrix.from_system(s)
    t_create = time.time() - t0
    info = A.info(levels=False)
    log(f"ldu_create {t_create:.1f}s", info)

    common = dict(smoother=2, n_vcycles=1, n_post_sweeps=4, n_finest_sweeps=4)
    mixed = dict(level_sweep_growth=0, pair_steps=args.pair_steps, n_cells_coarsest=args.mixed_n_coarsest,
                 cycle=hip.CYCLE_FIXED, scaling=args.mixed_scaling, cheby_eig_ratio=args.mixed_cheby_ratio,
                 cheby_eig_ratio_finest=args.mixed_cheby_ratio_finest, cheby_boost=args.mixed_cheby_boost, **common)
    t0 = time.time()
    P32 = hip.Precond(A, "AMG", amg=hip.amg_params(precision=32, **mixed))
    t_setup32 = time.time() - t0
    log(f"mixed hierarchy setup {t_setup32:.1f}s (device layout, agglomeration, coarse layouts, estimates)")
    t0 = time.time()
    P64 = hip.Precond(A, "AMG", amg=hip.amg_params(precision=64, level_sweep_growth=1, n_cells_coarsest=args.n_coarsest,
                                                   **common))
    t_setup64 = time.time() - t0
    log(f"fp64 hierarchy setup {t_setup64:.1f}s")
    P64f = None
    if not (args.mixed_only or args.fp64_only):
        P64f = hip.Precond(A, "AMG", amg=hip.amg_params(precision=64, **mixed))

    b = hip.DeviceArray.from_numpy(s.b)
    x = hip.DeviceArray(s.n)
    S32 = hip.Solver("FCG", P32, rel_tol=1e-8, max_iter=1000)
    S64 = hip.Solver("PCG", P64, rel_tol=1e-8, max_iter=1000)
    S64f = hip.Solver("FCG", P64f, rel_tol=1e-8, max_iter=1000) if P64f is not None else None

    def solve(S):
        x.zero()
        hip.synchronize()
        t = time.perf_counter()
        _, perf = S.solve(A, b, x=x)
        hip.synchronize()
        return time.perf_counter() - t, perf

    levels32 = P32.amg_levels()
    if args.fp64_only:
        solve(S64)  # warm-up
        times64 = []
        for k in range(max(1, args.fp64_steps)):
            dt, p64 = solve(S64)
            times64.append(dt)
            log(f"fp64 step {k}: {dt * 1e3:.1f} ms, {p64['nIterations']} iterations, converged={p64['converged']}")
        print(json.dumps({"fp64_s": times64, "iterations": p64["nIterations"], "hierarchy": P64.amg_levels(),
                          "n_post_sweeps": 4, "level_sweep_growth": 1, "n_finest_sweeps": 4}), flush=True)
        return
    for _ in range(args.warmup):
        solve(S32)
    log("warmup done")
    times32 = []
    for k in range(args.steps):
        dt, p32 = solve(S32)
        times32.append(dt)
        log(f"mixed step {k}: {dt * 1e3:.1f} ms, {p32['nIterations']} iterations, converged={p32['converged']}")
    x32 = x.numpy()
    ms_dev32 = S32.last_ms()
    if args.mixed_only:
        print(json.dumps({"mixed_s": times32, "iterations": p32["nIterations"], "hierarchy": levels32,
                          "pair_steps": args.pair_steps}), flush=True)
        return

    solve(S64)  # warm-up
    times64 = []
    for k in range(max(1, args.fp64_steps)):
        dt, p64 = solve(S64)
        times64.append(dt)
        log(f"fp64 step {k}: {dt * 1e3:.1f} ms, {p64['nIterations']} iterations, converged={p64['converged']}")
    x64 = x.numpy()
    t64 = float(np.median(times64))
    solve(S64f)  # warm-up
    times64f = []
    for k in range(max(1, args.fp64_steps)):
        dt, p64f = solve(S64f)
        times64f.append(dt)
        log(f"fp64 same-algorithm step {k}: {dt * 1e3:.1f} ms, {p64f['nIterations']} iterations, "
            f"converged={p64f['converged']}")
    x64f = x.numpy()
    t64f = float(np.median(times64f))

    # accuracy: against an FP64 solve to 1e-12 (flexible CG around the FP64
    # hierarchy: the reference's V-cycle is not a fixed linear operator, so
    # plain PCG stalls before 1e-12), and the true residual
    Sref = hip.Solver("FCG", P64, rel_tol=1e-12, max_iter=2000)
    _, pref = solve(Sref)
    xs = x.numpy()
    err32 = float(np.linalg.norm(x32 - xs) / np.linalg.norm(xs))
    err64 = float(np.linalg.norm(x64 - xs) / np.linalg.norm(xs))
    err64f = float(np.linalg.norm(x64f - xs) / np.linalg.norm(xs))
    r32 = s.b - A.spmv(x32)
    true_res32 = float(np.linalg.norm(r32) / np.linalg.norm(s.b))
    log(f"errors vs FP64@1e-12 ({pref['nIterations']} it): mixed {err32:.2e}, fp64@1e-8 {err64:.2e}; "
        f"mixed true rel. residual {true_res32:.2e}")

    # per-kernel rooflines on the finest level (HIP events on the library stream)
    kernels = {}
    for name, prec in [("cheby", 32), ("cheby", 64), ("amul", 32), ("amul", 64), ("residual", 64),
                       ("axpy_norm", 64), ("dot", 64), ("gs", 64), ("dic", 64), ("copy", 64), ("copy", 32)]:
        ms, by = hip.profile_kernel(A, name, prec, reps=20)
        gbs = by / (ms * 1e-3) / 1e9
        kernels[f"{name}_fp{prec}"] = {"ms": round(ms, 4), "alg_bytes": by, "GBps": round(gbs, 1),
                                       "frac": round(gbs / PEAK_HBM_GBS, 3)}
        log(f"kernel {name} fp{prec}: {ms:.3f} ms, {gbs:.0f} GB/s ({gbs / PEAK_HBM_GBS:.0%} of 8 TB/s)")
    dom = kernels["cheby_fp32"]
    traffic = None
    if args.traffic and os.path.exists(args.traffic) and N == 256:
        traffic = json.load(open(args.traffic))["cheby_fp32"]["traffic_bytes"]
    roofline = {"kernel": "Chebyshev sweep, FP32, finest level (fused Amul + update)", "bound": "hbm",
                "achieved": dom["GBps"], "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": dom["frac"],
                "alg_bytes_per_launch": dom["alg_bytes"], "traffic": traffic,
                "traffic_source": ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes of bench.py --kernel-only cheby, "
                                   "scaled by the copy kernel (profiles/hbm_traffic_256.md)") if traffic else None}

    # CPU baseline: the reference's own x86 solve path (compiled from its
    # sources into oracle/_ref) or, where that is absent, the oracle restatement.
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_size or N, p64["nIterations"], args.n_coarsest)

    del P32, P64, P64f, S32, S64, S64f, Sref, A, x, b
    configs = None
    if not args.no_configs:
        configs = config_legs(args.config_size, not args.no_cpu, args.configs.split(","))

    batch = None
    if not args.no_batch:
        batch = batch_bench(args.batch_systems, args.batch_launches, not args.no_cpu)

    t32 = float(np.median(times32))
    out = {
        "metric": METRIC, "value": round(t32, 5), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(t32 * 1e3, 3), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"C4: {N}^3 permuted log-uniform(1e-4..1) Poisson LDU, AMG-PCG to rel.res 1e-8",
                   "n_cells": s.n, "n_faces": s.nfaces,
                   "mixed": f"FCG(FP64) + FP32 AMG V(0,4) Chebyshev(4) over [lambda/"
                            f"{args.mixed_cheby_ratio_finest or args.mixed_cheby_ratio:g}, "
                            f"{args.mixed_cheby_boost:g} lambda] on the finest level and [lambda/"
                            f"{args.mixed_cheby_ratio:g}, {args.mixed_cheby_boost:g} lambda] on the coarse ones, "
                            f"{args.pair_steps} pairwise steps a level, "
                            f"coarsest threshold {args.mixed_n_coarsest}, dense FP64 inverse (GPU Gauss-Jordan), "
                            f"correction scaling {['none', 'coarse levels', 'reference'][args.mixed_scaling]}",
                   "fp64": "PCG + FP64 AMG V(0,4+level) Chebyshev(4) over [lambda/1.5, 1.1 lambda] (the reference's), "
                           "reference pairwise hierarchy, coarsest PCG+DIC 1e-2",
                   "fp64_same_algorithm": "the mixed algorithm with the hierarchy in FP64",
                   "n_cells_coarsest_fp64": args.n_coarsest, "n_cells_coarsest_mixed": args.mixed_n_coarsest},
        "hierarchy_levels": levels32,
        "mixed": {"s": round(t32, 5), "device_ms": round(ms_dev32, 3), "iterations": p32["nIterations"],
                  "converged": p32["converged"], "err_vs_fp64_1e-12": err32, "true_rel_residual": true_res32},
        "fp64_drop_in": {"s": round(t64, 5), "steps_s": [round(t, 5) for t in times64],
                         "iterations": p64["nIterations"], "converged": p64["converged"],
                         "err_vs_fp64_1e-12": err64},
        "fp64_same_algorithm": {"s": round(t64f, 5), "steps_s": [round(t, 5) for t in times64f],
                                "iterations": p64f["nIterations"],
                                "converged": p64f["converged"], "err_vs_fp64_1e-12": err64f},
        "speedup_mixed_vs_fp64": round(t64 / t32, 3),
        "speedup_mixed_vs_fp64_same_algorithm": round(t64f / t32, 3),
        "setup_s": {"generate": round(t_gen, 2), "ldu_create": round(t_create, 2),
                    "mixed_hierarchy": round(t_setup32, 2), "fp64_hierarchy": round(t_setup64, 2)},
        "device": dev_name, "roofline": roofline, "kernels": kernels, "cpu_baseline": cpu,
        "configs": configs, "batch": batch,
    }
    print(json.dumps(out), flush=True)


def config_legs(N, with_cpu, which=("c1", "c2", "c3"), reps=3, with_mixed=True):
    """The FP64 drop-in configurations: C1 (PCG + DIC, 32^3 constant Poisson,
    rel. tol 1e-6), and at edge N C2 (PCG + AMG with Gauss-Seidel smoothing,
    coarsest level <= 500 cells) and C3 (PBiCGStab + DILU), rel. tol 1e-8.
    Each is timed on the GPU (median of `reps` solves from x = 0, setup
    excluded) beside the reference's x86 solve of the same system (one and
    three iterations timed, extrapolated to the GPU's count)."""
    from unap_amd import hip, meshgen

    legs = {
        "c1": (32, 1e-6, "PCG", "DIC", None, None, "PCG + DIC"),
        "c2": (N, 1e-8, "PCG", "AMG", {"smoother": 1, "n_cells_coarsest": 500, "n_vcycles": 1}, [1, 500, 1, 0, 2, 2],
               "PCG + FP64 AMG, Gauss-Seidel V(0,2), coarsest <= 500 cells"),
        "c3": (N, 1e-8, "PBiCGStab", "DILU", None, None, "PBiCGStab + DILU"),
    }
    # the precision pattern on the same systems: the preconditioner in FP32
    # inside an FP64 flexible CG / BiCGStab
    mixed_legs = {
        "c1": ("FCG", "DIC", {"precision": 32}, "FCG + FP32 DIC"),
        "c2": ("FCG", "AMG", {"precision": 32, "smoother": 1, "n_cells_coarsest": 4000, "n_vcycles": 1,
                              "cycle": hip.CYCLE_FIXED, "pair_steps": 4, "level_sweep_growth": 0},
               "FCG + FP32 AMG (Gauss-Seidel V(0,2), fixed cycle, 4 pairwise steps a level, coarsest threshold "
               "4000, dense FP64 coarsest)"),
        "c3": ("FBiCGStab", "DILU", {"precision": 32}, "flexible BiCGStab + FP32 DILU"),
    }
    # the FP64 solution the errors are measured against
    ref_legs = {
        "c1": ("PCG", "DIC", None),
        "c2": ("FCG", "AMG", {"smoother": 1, "n_cells_coarsest": 500, "n_vcycles": 1}),
        "c3": ("PBiCGStab", "DILU", None),
    }

    def timed(S, A, b, x):
        _, perf = S.solve(A, b, x=x)  # warm-up (first-call estimates)
        times = []
        for _ in range(reps):
            x.zero()
            hip.synchronize()
            t1 = time.perf_counter()
            _, perf = S.solve(A, b, x=x)
            hip.synchronize()
            times.append(time.perf_counter() - t1)
        return float(np.median(times)), perf
    out = {}
    for key, (size, tol, solver, pre, amg, mg, what) in legs.items():
        if key not in which:
            continue
        s = meshgen.config_system(int(key[1]), size)
        A = hip.LduMatrix.from_system(s)
        t0 = time.time()
        P = hip.Precond(A, pre, amg=hip.amg_params(**amg) if amg else None)
        S = hip.Solver(solver, P, rel_tol=tol, max_iter=1000)
        b = hip.DeviceArray.from_numpy(s.b)
        x = hip.DeviceArray(s.n)
        t, perf = timed(S, A, b, x)
        t_setup = time.time() - t0 - reps * t
        x64 = x.numpy()
        res = _ldu_residual(s, x64)
        leg = {"workload": f"{key.upper()}: {size}^3 {what}, rel. tol {tol:g}", "s": round(t, 5),
               "iterations": perf["nIterations"],
               "converged": perf["converged"], "true_rel_residual": res, "setup_and_first_solve_s": round(t_setup, 2)}
        log(f"{key}: {t * 1e3:.1f} ms, {perf['nIterations']} iterations, converged={perf['converged']}, "
            f"true rel. residual {res:.2e}")
        # FP32 preconditioner, and both solutions against an FP64 solve to 1e-12
        msolver, mpre, mamg, mwhat = mixed_legs[key]
        if not with_mixed:
            msolver = None
        if msolver is not None:
            Pm = hip.Precond(A, mpre, amg=hip.amg_params(**mamg))
            tm, pm = timed(hip.Solver(msolver, Pm, rel_tol=tol, max_iter=3000), A, b, x)
            x32 = x.numpy()
            rsolver, rpre, ramg = ref_legs[key]
            Pr = hip.Precond(A, rpre, amg=hip.amg_params(**ramg) if ramg else None)
            x.zero()
            hip.Solver(rsolver, Pr, rel_tol=1e-12, max_iter=5000).solve(A, b, x=x)
            xr = x.numpy()
            e32 = float(np.linalg.norm(x32 - xr) / np.linalg.norm(xr))
            e64 = float(np.linalg.norm(x64 - xr) / np.linalg.norm(xr))
            leg["err_vs_fp64_1e-12"] = e64
            leg["mixed"] = {"solver": mwhat, "s": round(tm, 5), "iterations": pm["nIterations"],
                            "converged": pm["converged"], "true_rel_residual": _ldu_residual(s, x32),
                            "err_vs_fp64_1e-12": e32, "speedup_vs_fp64": round(t / tm, 3)}
            log(f"{key} mixed ({mwhat}): {tm * 1e3:.1f} ms, {pm['nIterations']} iterations; error vs FP64 1e-12: "
                f"mixed {e32:.2e}, fp64 {e64:.2e}")
            del Pm, Pr
        if with_cpu:
            from oracle import oracle as O
            path = O.reference()
            kind = "reference"
            if path is None:
                path, kind = O.oracle(), "port"
            M = path.matrix(s)
            if s.n <= 100000:  # small enough to solve whole on the CPU
                cpu_perf, _, _, est = O.timed_solve(path, M, solver, pre, s.b, 1, 1000, tol, mg=mg)
                sample = f"the whole solve ({cpu_perf['nIterations']} iterations) after a cold one-iteration solve"
                log(f"{key} cpu ({kind}): {est * 1e3:.1f} ms, {cpu_perf['nIterations']} iterations")
            else:
                _, _, _, t_one = O.timed_solve(path, M, solver, pre, s.b, 1, 1, tol, mg=mg)
                _, _, _, t_three = O.timed_solve(path, M, solver, pre, s.b, 1, 3, tol, mg=mg)
                per_it = (t_three - t_one) / 2
                est = t_one + (perf["nIterations"] - 1) * per_it
                sample = (f"solves of 1 and 3 iterations ({t_one:.2f}s, {t_three:.2f}s) after a cold one, "
                          f"extrapolated to {perf['nIterations']} iterations")
                log(f"{key} cpu ({kind}): {per_it:.2f}s an iteration -> {est:.1f}s")
            leg["cpu_baseline"] = {"value": round(est, 4), "unit": "s", "cores": 1, "kind": kind, "sample": sample}
            leg["speedup_vs_cpu"] = round(est / t, 1)
        out[key] = leg
        del S, P, A, b, x
    return out


def _ldu_residual(s, x):
    """||b - A x|| / ||b|| with A applied from the LDU arrays (host numpy)."""
    y = s.diag * x
    np.add.at(y, s.l, s.upper * x[s.u])
    np.add.at(y, s.u, s.lower_or_upper() * x[s.l])
    return float(np.linalg.norm(s.b - y) / np.linalg.norm(s.b))


def batch_bench(nsys, launches, with_cpu):
    """Workload C5: nsys independent permuted LDU box systems (edges 8..48,
    60 % SPD Poisson by PCG + DIC, 40 % convection-diffusion by PBiCGStab +
    DILU, FP32 preconditioners in FP64 Krylov loops) in one persistent launch
    per solve; systems/s and launch-latency percentiles over `launches`."""
    from unap_amd import hip, meshgen
    t0 = time.time()
    specs = meshgen.batch_specs(nsys=nsys)
    systems = [meshgen.batch_system(sp) for sp in specs]
    solvers = ["PCG" if sp[3] == meshgen.POISSON_LOGUNIF else "PBiCGStab" for sp in specs]
    t_gen = time.time() - t0
    t0 = time.time()
    B = hip.Batch(systems, solvers)
    t_create = time.time() - t0
    ncells = int(B.cell_off[-1])
    log(f"batch: {nsys} systems, {ncells} cells, generated {t_gen:.1f}s, created {t_create:.1f}s")
    bh = np.concatenate([s.b for s in systems])
    b = hip.DeviceArray.from_numpy(bh)
    x = hip.DeviceArray(ncells)
    t0 = time.time()
    _, perfs = B.solve(b, x=x, rel_tol=1e-8, max_iter=1000)  # first solve autotunes
    t_tune = time.time() - t0
    bs, wg, rw = B.config()
    log(f"batch autotune {t_tune:.1f}s -> block {bs}, {wg} workgroups/CU, {rw} rows a wavefront pass")
    lat = []
    for _ in range(launches):
        x.zero()
        _, perfs = B.solve(b, x=x, rel_tol=1e-8, max_iter=1000)
        lat.append(B.last_ms())
    lat = np.array(lat)
    xh = x.numpy()
    off = B.cell_off
    res = [_ldu_residual(s, xh[off[i]:off[i + 1]]) for i, s in enumerate(systems)]
    its = np.array([p["nIterations"] for p in perfs])
    conv = all(p["converged"] for p in perfs)
    p50 = float(np.percentile(lat, 50))
    # roofline of the launch: algorithmic bytes per system and iteration
    # (each array a step reads or writes, once, unpadded), summed over the
    # systems' iterations, over the median launch time.  PCG + FP32 DIC:
    # 152 n + 40 f (ELL Amul 28 n + 24 f, FP32 sweep tables 36 n + 16 f, dots
    # and updates 88 n); PBiCGStab + FP32 DILU: 296 n + 80 f (two of each).
    alg = 0.0
    for sp, sysm, it in zip(specs, systems, its):
        n, f = sysm.n, sysm.nfaces
        per = (152 * n + 40 * f) if sp[3] == meshgen.POISSON_LOGUNIF else (296 * n + 80 * f)
        alg += per * int(it)
    gbs = alg / (p50 * 1e-3) / 1e9
    log(f"batch launch ms p50 {p50:.2f} p90 {np.percentile(lat, 90):.2f} p99 {np.percentile(lat, 99):.2f}; "
        f"{nsys / (p50 * 1e-3):.0f} systems/s; all converged {conv}; max true rel. residual {max(res):.2e}; "
        f"{gbs:.0f} GB/s algorithmic ({gbs / PEAK_HBM_GBS:.1%} of 8 TB/s)")
    out = {"workload": f"C5: {nsys} permuted LDU boxes, edges 8..48 ({ncells} cells), 60% PCG+DIC / "
                       "40% PBiCGStab+DILU, FP32 preconditioners, FP64 Krylov, rel.tol 1e-8, one launch",
           "systems_per_s": round(nsys / (p50 * 1e-3), 1), "launches": launches,
           "latency_ms": {"p50": round(p50, 3), "p90": round(float(np.percentile(lat, 90)), 3),
                          "p99": round(float(np.percentile(lat, 99)), 3), "min": round(float(lat.min()), 3)},
           "roofline": {"bound": "hbm", "alg_bytes_per_launch": alg, "achieved": round(gbs, 1),
                        "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(gbs / PEAK_HBM_GBS, 4)},
           "all_converged": conv, "max_true_rel_residual": max(res),
           "iterations": {"min": int(its.min()), "median": float(np.median(its)), "max": int(its.max())},
           "autotuned": {"block": bs, "workgroups_per_cu": wg, "rows_per_wavefront": rw, "tune_s": round(t_tune, 2)},
           "setup_s": {"generate": round(t_gen, 2), "create": round(t_create, 2)}}
    if with_cpu:
        out["cpu_baseline"] = batch_cpu_baseline(systems, solvers, p50)
        out["cpu_baseline_all_cores"] = batch_cpu_all_cores(nsys)
    return out


def _cpu_solve_systems(args):
    """Worker: solve the given systems of the batch on the CPU path, one after
    another; returns the seconds spent (preconditioner setup + solve)."""
    nsys, idx = args
    from oracle import oracle as O
    from unap_amd import meshgen
    path = O.reference() or O.oracle()
    specs = meshgen.batch_specs(nsys=nsys)
    t = 0.0
    for i in idx:
        s = meshgen.batch_system(specs[i])
        kind = "PCG" if specs[i][3] == meshgen.POISSON_LOGUNIF else "PBiCGStab"
        _, t_setup, t_solve, _ = O.timed_solve(path, path.matrix(s), kind, "DIC" if kind == "PCG" else "DILU",
                                               s.b, 1000, 1, 1e-8)
        t += t_setup + t_solve
    return t


def batch_cpu_all_cores(nsys):
    """All nsys systems solved by the CPU path, one independent single-threaded
    solve per worker process (the reference runs one system per MPI rank),
    over min(16, cores) worker processes; systems/s from the wall clock."""
    import concurrent.futures as cf
    import multiprocessing as mp
    from oracle import oracle as O
    kind = "reference" if O.reference() is not None else "port"
    workers = max(1, min(16, os.cpu_count() or 1))
    chunks = [(nsys, list(range(w, nsys, workers))) for w in range(workers)]
    t0 = time.perf_counter()
    with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        busy = list(ex.map(_cpu_solve_systems, chunks))
    wall = time.perf_counter() - t0
    log(f"batch cpu all cores ({kind}): {nsys} systems on {workers} processes in {wall:.1f}s wall")
    return {"value": round(nsys / wall, 2), "unit": "systems/s", "cores": workers, "kind": kind,
            "sample": f"all {nsys} systems, one FP64 solve per system (preconditioner setup + solve), round-robin "
                      f"over {workers} worker processes; wall clock {wall:.1f}s including process start, "
                      f"per-process busy {min(busy):.1f}-{max(busy):.1f}s"}


def batch_cpu_baseline(systems, solvers, gpu_ms, stride=16):
    """The same systems solved one after another by the CPU path (FP64 DIC /
    DILU as the reference runs them) on one core: every `stride`-th system
    timed (preconditioner setup + solve), scaled by `stride`."""
    from oracle import oracle as O
    path = O.reference()
    kind = "reference"
    if path is None:
        path = O.oracle()
        kind = "port"
    idx = list(range(0, len(systems), stride))
    t = 0.0
    for k, i in enumerate(idx):
        s = systems[i]
        M = path.matrix(s)
        pre = "DIC" if solvers[i] == "PCG" else "DILU"
        _, t_setup, t_solve, _ = O.timed_solve(path, M, solvers[i], pre, s.b, 1000, 1, 1e-8)
        t += t_setup + t_solve
        if k % 8 == 7:
            log(f"batch cpu: {k + 1}/{len(idx)} systems, {t:.2f}s")
    est = t * len(systems) / len(idx)
    log(f"batch cpu ({kind}): {len(idx)} systems in {t:.2f}s -> {est:.1f}s for {len(systems)}")
    return {"value": round(len(systems) / est, 2), "unit": "systems/s", "cores": 1, "kind": kind,
            "sample": f"every {stride}th system ({len(idx)} of {len(systems)}) solved in FP64 on one core "
                      f"(preconditioner setup + solve, {t:.2f}s), scaled by {len(systems) / len(idx):.0f}"}


def cpu_baseline(N, iters_fp64, n_coarsest):
    """UNAP's x86 path timed on this host: a bounded sample of the FP64
    drop-in's algorithm, extrapolated to the GPU FP64 iteration count."""
    from oracle import oracle as O
    from unap_amd import meshgen

    path = O.reference()
    kind = "reference"
    if path is None:
        path = O.oracle()
        kind = "port"
    s = meshgen.config_system(4, N)
    M = path.matrix(s)
    log(f"cpu baseline ({kind}) at {N}^3 ...")
    perf, t_setup, t_cold, t_warm = O.timed_solve(path, M, "PCG", "AMG", s.b, 1, 1, 1e-8,
                                                  mg=[2, n_coarsest, 1, 0, 4, 4])
    est = t_cold + (iters_fp64 - 1) * t_warm
    log(f"cpu: setup {t_setup:.1f}s, first iteration {t_cold:.1f}s, next iteration {t_warm:.1f}s -> "
        f"{est:.1f}s for {iters_fp64} iterations")
    return {"value": round(est, 2), "unit": "s", "cores": 1, "kind": kind,
            "sample": (f"{N}^3 C4 system, FP64 PCG + AMG(Chebyshev 4, nPost+level) as the reference runs it: "
                       f"one cold iteration ({t_cold:.2f}s, includes the eigen estimates) and one warm iteration "
                       f"({t_warm:.2f}s) timed, extrapolated to the {iters_fp64} iterations of the GPU FP64 "
                       f"solve; setup {t_setup:.1f}s excluded"),
            "host_cpu": _cpu_model()}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip() + f" (nproc {os.cpu_count()})"
    except OSError:
        pass
    return None


if __name__ == "__main__":
    main()

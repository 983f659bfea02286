"""Cell numbering and locality: the random relabelling of the synthetic
systems, reverse Cuthill-McKee of it (src/tools/rcm/cooRCM.cpp, here
unap_hip_rcm_order + unap_hip_ldu_renumber) and the box's own lexicographic
numbering, on one matrix.

For each numbering:
  * bandwidth max |l - u| and mean |l - u| over the faces;
  * depth of the sweep level schedules (forward / backward levels: the DIC,
    DILU and Gauss-Seidel sweeps run level by level);
  * on the GPU, with the matrix stored in the caller's numbering
    (UNAP_HIP_CELL_ORDER=natural) and in the library's own orders (default:
    breadth-first for Amul, level-major for the sweeps): the FP64 Amul and
    one DIC application (HIP events, mean of 20 launches), and a PCG + DIC
    solve to 1e-6 (iterations, device time).

    python -m unap_amd.order_study [--config 2] [--size 128] > out.md

Each cell-order setting runs in a child process of its own (the library
reads the variable once).
"""
import argparse
import json
import os
import subprocess
import sys

import numpy as np

from unap_amd import hip, meshgen

NUMBERINGS = ("random", "rcm", "lexicographic")


def _system(config, size, numbering):
    s = meshgen.box_system(size, size, size, *_law(config), with_perm=True)
    if numbering == "random":
        return s
    if numbering == "rcm":
        new_of_old = hip.rcm_order(s.n, s.l, s.u)
    else:  # the box's lexicographic order: perm[natural cell] is its label
        new_of_old = np.empty(s.n, np.int64)
        new_of_old[s.perm] = np.arange(s.n)
    lo = None if s.symmetric else s.lower
    l, u, _, lo2, d2, up2 = hip.renumber(s.n, s.l, s.u, new_of_old, s.diag, s.upper, lo)
    b2 = np.empty(s.n)
    b2[new_of_old] = s.b
    return meshgen.LduSystem(s.n, l, u, lo2, d2, up2, b2, s.symmetric, s.dims)


def _law(config):
    """(kind, p0, p1, shift, seed) of the box generator for workload C<config>."""
    if config == 1:
        return meshgen.POISSON_CONST, 0.0, 0.0, 1e-6, 1001
    if config == 2:
        return meshgen.POISSON_LOGUNIF, -3.0, 0.0, 1e-6, 1002
    if config == 4:
        return meshgen.POISSON_LOGUNIF, -4.0, 0.0, 1e-6, 1004
    raise ValueError("symmetric configurations only (1, 2, 4): the study solves with PCG + DIC")


def _inner(config, size):
    out = {}
    for name in NUMBERINGS:
        s = _system(config, size, name)
        A = hip.LduMatrix.from_system(s)
        info = A.info()
        amul_ms, amul_b = hip.profile_kernel(A, "amul", 64, reps=20)
        dic_ms, dic_b = hip.profile_kernel(A, "dic", 64, reps=20)
        P = hip.Precond(A, "DIC")
        S = hip.Solver("PCG", P, rel_tol=1e-6, max_iter=5000)
        b = hip.DeviceArray.from_numpy(s.b)
        x = hip.DeviceArray(s.n)
        S.solve(A, b, x=x)  # warm-up: layouts, schedules, first launches
        x.zero()
        _, perf = S.solve(A, b, x=x)
        out[name] = {"bandwidth": int(np.max(np.abs(s.u - s.l))), "mean_offset": float(np.mean(np.abs(s.u - s.l))),
                     "fwd_levels": info["fwd_levels"], "bwd_levels": info["bwd_levels"],
                     "amul_ms": amul_ms, "amul_GBps": amul_b / amul_ms / 1e6, "dic_ms": dic_ms,
                     "pcg_dic_iterations": perf["nIterations"], "pcg_dic_ms": S.last_ms()}
        print(f"{name}: done", file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--inner", action="store_true", help=argparse.SUPPRESS)
    a = ap.parse_args()
    if a.inner:
        print(json.dumps(_inner(a.config, a.size)))
        return
    res = {}
    for mode, env in (("caller", {"UNAP_HIP_CELL_ORDER": "natural"}), ("library", {})):
        e = dict(os.environ)
        e.pop("UNAP_HIP_CELL_ORDER", None)
        e.update(env)
        print(f"cell order: {mode}", file=sys.stderr, flush=True)
        r = subprocess.run([sys.executable, "-m", "unap_amd.order_study", "--inner", "--config", str(a.config),
                            "--size", str(a.size)], env=e, stdout=subprocess.PIPE, text=True, timeout=900)
        if r.returncode != 0:
            raise SystemExit(f"inner run ({mode}) failed with {r.returncode}")
        res[mode] = json.loads(r.stdout.strip().splitlines()[-1])
    n = a.size ** 3
    print(f"C{a.config} matrix at {a.size}^3 ({n} cells), FP64; mean of 20 launches (HIP events); "
          "PCG + DIC to rel. tol 1e-6, second solve\n")
    print("| numbering | bandwidth | mean \\|l-u\\| | fwd / bwd levels | Amul, stored as numbered | "
          "Amul, library order | DIC apply, as numbered | DIC apply, library order | "
          "PCG+DIC it. | PCG+DIC ms, as numbered | PCG+DIC ms, library order |")
    print("|---|---|---|---|---|---|---|---|---|---|---|")
    for name in NUMBERINGS:
        c, l = res["caller"][name], res["library"][name]
        print(f"| {name} | {c['bandwidth']} | {c['mean_offset']:.0f} | {c['fwd_levels']} / {c['bwd_levels']} | "
              f"{c['amul_ms'] * 1e3:.0f} us ({c['amul_GBps']:.0f} GB/s) | "
              f"{l['amul_ms'] * 1e3:.0f} us ({l['amul_GBps']:.0f} GB/s) | {c['dic_ms'] * 1e3:.0f} us | "
              f"{l['dic_ms'] * 1e3:.0f} us | {c['pcg_dic_iterations']} | {c['pcg_dic_ms']:.1f} | "
              f"{l['pcg_dic_ms']:.1f} |")
    print("\n```json\n" + json.dumps(res) + "\n```")


if __name__ == "__main__":
    main()

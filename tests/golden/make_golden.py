"""Write the golden vectors of the solve path from the compiled reference.

Run in the container that holds /root/reference, after
`make -C oracle -f Makefile.ref` has built oracle/_ref/libunap_ref.so:

    python tests/golden/make_golden.py

Inputs are not stored: every case names the seeded generator call that
rebuilds it (unap_amd.meshgen.config_system(c, N)), plus a checksum of the
regenerated inputs.  Outputs are the reference's own results on those inputs.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from unap_amd import meshgen  # noqa: E402

CASES = [(1, 8), (2, 8), (3, 8), (2, 10)]


def input_digest(s):
    h = hashlib.sha256()
    for a in (s.l, s.u, s.diag, s.upper, s.lower_or_upper(), s.b):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def probe(n, seed):
    return np.random.default_rng(seed).uniform(-1, 1, n)


def main():
    ref = O.reference()
    if ref is None:
        raise SystemExit("oracle/_ref/libunap_ref.so is not built")
    for c, N in CASES:
        s = meshgen.config_system(c, N)
        R = ref.matrix(s)
        out = {"case": np.array([c, N]), "digest": np.array(input_digest(s))}
        x = probe(s.n, 1)
        out["spmv"] = R.spmv(x)
        for k in ("diag", "DIC", "DILU"):
            w, rd = R.precond(k, x)
            out[f"precond_{k}_w"] = w
            out[f"precond_{k}_rd"] = rd
        out["gs2"] = R.smooth("GS", x, s.b, 2)
        out["cheby_2calls_3sweeps"] = R.smooth("Cheby", x, s.b, 3, calls=2)
        lam, xe = R.eigen(x, s.b)
        out["eigen_lambda"] = np.array(lam)
        out["eigen_x"] = xe
        w = np.abs(s.upper)
        levels = R.agglomerate(w, n_coarsest=20)
        out["aggl_nlevels"] = np.array(len(levels))
        for i, lv in enumerate(levels):
            for k in ("restrict", "face_restrict", "l", "u", "diag", "upper", "lower"):
                out[f"aggl{i}_{k}"] = lv[k]
        runs = {"pcg_dic": ("PCG", "DIC", None), "bicg_dilu": ("PBiCGStab", "DILU", None),
                "pcg_diag": ("PCG", "diag", None), "amgpcg_gs": ("PCG", "AMG", [1, 20, 1, 0, 2, 2]),
                "amgpcg_cheby": ("PCG", "AMG", [2, 20, 1, 0, 2, 2]),
                "mg_gs": ("MG", None, [1, 20, 1, 0, 2, 2]), "mg_cheby": ("MG", None, [2, 20, 1, 0, 2, 2])}
        for name, (solver, pre, mg) in runs.items():
            if solver == "PCG" and c == 3:
                continue  # PCG is for symmetric systems
            xs, p = R.solve(solver, pre, s.b, rel_tol=1e-8, max_iter=200, mg=mg)
            out[f"{name}_x"] = xs
            out[f"{name}_hist"] = p["history"]
            out[f"{name}_niter"] = np.array(p["nIterations"])
        np.savez_compressed(os.path.join(HERE, f"ref_c{c}_n{N}.npz"), **out)
        print("wrote", c, N, len(out), "arrays")


if __name__ == "__main__":
    main()

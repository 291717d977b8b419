"""Writes tests/golden/ref_runs.json from the reference GCGE compiled in
oracle/_ref (oracle/Makefile), on the seeded stencil problems of
gcge_amd.problems.

Each record: problem, sizes, shift, the reference's eigenvalues (first
nevConv + 5), numIter and nevConv, for the CCS app (app/app_ccs.c) and,
for small problems, the dense LAPACK app (app/app_lapack.c).

Run in the build container (the reference sources and MKL are needed to
build oracle/_ref):  python tests/golden/make_golden.py

With --cfg2 it instead writes tests/golden/ref_cfg2.json: one run of the
OpenMP build (oracle/_ref/libgcge_ref_omp.so, OMP_NUM_THREADS threads,
MKL one thread each) on the 128^3 7-point problem with nev 100, block 120,
nevInit = nevMax = 220, autoshift on, srand(0) before the solve as
gcge.c:508 does.  It takes about half an hour on 8 cores.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

from gcge_amd import problems as P  # noqa: E402
from ref_binding import RefLib  # noqa: E402

CASES = [
    # name, n, dim, nev, shift, apps
    ("cfg1_2d_128", 128, 2, 20, 1, ("ccs",)),
    ("cfg1_2d_128", 128, 2, 20, 0, ("ccs",)),
    ("2d_32", 32, 2, 10, 1, ("ccs", "lapack")),
    ("2d_32", 32, 2, 10, 0, ("ccs", "lapack")),
    ("2d_24", 24, 2, 20, 1, ("ccs", "lapack")),
    ("3d_12", 12, 3, 30, 1, ("ccs",)),
    ("3d_12", 12, 3, 30, 0, ("ccs",)),
    ("3d_16", 16, 3, 40, 0, ("ccs",)),
    # variable coefficients: dim 2 is the 5-point operator of the batched
    # configuration (sigma 0.5), dim 27 the 27-point one of config 3 (sigma 1)
    ("vc2d_40", 40, "vc2", 16, 1, ("ccs",)),
    ("vc2d_40", 40, "vc2", 16, 0, ("ccs",)),
    ("vc27_16", 16, "vc27", 20, 1, ("ccs",)),
    ("vc27_16", 16, "vc27", 20, 0, ("ccs",)),
]


def problem(name, n, dim, nev):
    """The seeded problem of a case (also used by the tests)."""
    if dim == "vc2":
        A, B = P.varcoef_pair_2d(n, 0.5, np.random.default_rng([P.SEED_BASE, 2000 + n]))
    elif dim == "vc27":
        A, B = P.varcoef27_pair(n, 1.0, np.random.default_rng([P.SEED_BASE, 2700 + n]))
    else:
        return P.stencil_problem(n, dim, nev)
    bs, ni, nm = P.slepc_block_policy(nev)
    return P.CCSProblem(name, A, B, nev, bs, ni, nm)


def main_cfg2():
    """The headline problem through the reference's CCS app (app/app_ccs.c),
    its GCG driver (src/ops_eig_sol_gcg.c:1259-1562) and the plug-in's call
    sequence (gcge.c:442-535, via oracle/ref_harness.c)."""
    from ref_binding import DEFAULT_OMP
    ref = RefLib(DEFAULT_OMP)
    prob = P.config2()
    r = ref.solve_ccs(prob, shift=1)
    rec = {"name": prob.name, "n": 128, "dim": 3, "nev": prob.nev, "shift": 1, "app": "ccs",
           "block_size": prob.block_size, "nev_init": prob.nev_init, "nev_max": prob.nev_max,
           "tol": [prob.tol_abs, prob.tol_rel], "gap_min": prob.gap_min,
           "num_iter": r["num_iter"], "nev_conv": r["nev_conv"], "seconds": r["seconds"],
           "threads": os.environ.get("OMP_NUM_THREADS"),
           "eval": [float(x) for x in r["eval"][:prob.nev_max]]}
    print(prob.name, r["num_iter"], r["nev_conv"], r["seconds"], flush=True)
    with open(os.path.join(HERE, "ref_cfg2.json"), "w") as f:
        json.dump({"source": "reference GCGE compiled from its sources by oracle/Makefile "
                             "(OpenMP variant), MKL 2021.4 BLAS/LAPACK, one process, no MPI",
                   "run": rec}, f, indent=1)


def main():
    if "--cfg2" in sys.argv:
        return main_cfg2()
    ref = RefLib()
    out = []
    for name, n, dim, nev, shift, apps in CASES:
        prob = problem(name, n, dim, nev)
        for app in apps:
            if app == "ccs":
                r = ref.solve_ccs(prob, shift=shift)
            else:
                ops = ref.lapack_table()
                hA, ka = ref.wrap_dense(prob.A.toarray())
                hB, kb = ref.wrap_dense(prob.B.toarray())
                r = ref.solve(ops, hA, hB, prob.n, prob.nev, prob.block_size, prob.nev_init,
                              prob.nev_max, (prob.tol_abs, prob.tol_rel), prob.gap_min, shift)
            k = min(r["nev_conv"] + 5, prob.nev_max)
            rec = {"name": name, "n": n, "dim": dim, "nev": nev, "shift": shift, "app": app,
                   "block_size": prob.block_size, "nev_init": prob.nev_init,
                   "nev_max": prob.nev_max, "num_iter": r["num_iter"],
                   "nev_conv": r["nev_conv"], "eval": [float(x) for x in r["eval"][:k]]}
            print(name, app, shift, r["num_iter"], r["nev_conv"], flush=True)
            out.append(rec)
    with open(os.path.join(HERE, "ref_runs.json"), "w") as f:
        json.dump({"source": "reference GCGE compiled from its sources by oracle/Makefile, "
                             "MKL 2021.4 BLAS/LAPACK, one process, no MPI",
                   "runs": out}, f, indent=1)


if __name__ == "__main__":
    main()

"""Builds the launch-shape table shipped beside the library.

Runs solves of the shapes the configurations produce with the tuner on
(gcge_tune_set_mode): the first call of each (operation, shape class)
times every candidate launch shape on its own operands and keeps the
fastest.  The table is then saved as text; the library loads
tuning_gfx950.txt from its own directory when the first device context is
created.

    python -m gcge_amd.tuning --out gcge_amd/tuning_gfx950.txt
"""
from __future__ import annotations

import argparse
import sys
import time

from . import abi, problems as P
from . import solver as S


def workloads(quick: bool = False):
    """(label, problem, mode) triples covering the configurations' shapes:
    2D variable-coefficient problems of 2k-64k rows, 3D 7-point problems of
    48^3-96^3 and 128^3, the 27-point operator of 160^3."""
    out = []
    for i in range(0, 64 if not quick else 8, 1 if not quick else 1):
        g = P.config4_grid(i)
        out.append((g.name, g, "mixed"))
    for n in ((48, 64, 80, 96) if not quick else (48,)):
        bs, ni, nm = P.slepc_block_policy(64)
        out.append((f"3d_{n}", P.stencil_grid_problem(n, 3, 64, bs, ni, nm, f"3d_{n}"), "mixed"))
    if not quick:
        g = P.config2_grid()
        out.append((g.name, g, "mixed"))
        out.append((g.name, g, "fp64"))
        g3 = P.config3_grid()
        out.append((g3.name, g3, "mixed"))
    return out


def tune(path: str, quick: bool = False, log=print) -> int:
    lib = abi.load()
    lib.gcge_hip_synchronize()      # the context exists: the shipped table is loaded
    lib.gcge_tune_clear()
    lib.gcge_tune_set_mode(1)
    try:
        for label, prob, mode in workloads(quick):
            t = time.perf_counter()
            r = S.solve_problem(prob, shift=1 if mode == "fp64" else 0, mode=mode,
                                rand="device", seed=1)
            log(f"[tune] {label} {mode}: N={prob.n} nconv {r.nev_conv} "
                f"{time.perf_counter() - t:.2f}s")
    finally:
        lib.gcge_tune_set_mode(0)
    return lib.gcge_tune_save(path.encode())


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--quick", action="store_true")
    a = ap.parse_args(argv)
    n = tune(a.out, a.quick, log=lambda s: print(s, file=sys.stderr, flush=True))
    print(f"{n} entries written to {a.out}")


if __name__ == "__main__":
    main()

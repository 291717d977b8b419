"""Bounded CPU sample of the reference GCGE on the 128^3 problem.

TEST INFRASTRUCTURE: the cpu_baseline leg of bench.py runs this in a child
process.  It solves with the reference driver and the reference CCS app
(oracle/_ref, compiled from the reference sources) for a fixed number of
outer iterations and prints one JSON line with the wall time and the
reference's own phase timers (src/ops_eig_sol_gcg.c:27-41).

usage: python oracle/ref_sample.py <libgcge_ref_omp.so> <outer iterations>
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

from gcge_amd import problems as P  # noqa: E402
from ref_binding import RefLib  # noqa: E402


def main():
    lib, iters = sys.argv[1], int(sys.argv[2])
    ref = RefLib(lib)
    prob = P.config2()
    r = ref.solve_ccs(prob, shift=1, max_iter=iters)
    print(json.dumps({"seconds": r["seconds"], "num_iter": r["num_iter"],
                      "nev_conv": r["nev_conv"], "phases": ref.phase_times(),
                      "threads": os.environ.get("OMP_NUM_THREADS")}))


if __name__ == "__main__":
    main()

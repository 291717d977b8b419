"""Bounded CPU samples of the reference GCGE on the benchmark workloads.

TEST INFRASTRUCTURE: the cpu_baseline leg of bench.py runs this in a child
process.  It solves with the reference driver and the reference CCS app
(oracle/_ref, compiled from the reference sources) and prints one JSON line
with the wall time, the outer iterations, the converged pairs and the
reference's own phase timers (src/ops_eig_sol_gcg.c:27-41).

usage: python oracle/ref_sample.py <libgcge_ref_omp.so> <workload> <size>
  cfg2, cfg3: the 128^3 / 160^3 problem for <size> outer iterations (the
              driver's InitializeX plus that many iterations);
  cfg4:       the first <size> problems of the 512-problem batch, each
              solved to convergence, one after another;
  cfg5:       the first <size> requests of the served stream, likewise.
The reference runs with its SLEPc plug-in's defaults (autoshift on,
gcge.c:733), FP64 throughout.
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

from gcge_amd import problems as P  # noqa: E402
from ref_binding import RefLib  # noqa: E402


def main():
    lib, what, size = sys.argv[1], sys.argv[2], int(sys.argv[3])
    ref = RefLib(lib)
    if what in ("cfg2", "cfg3"):
        prob = P.config2() if what == "cfg2" else P.config3()
        print(f"{what}: N = {prob.n}, InitializeX + {size} outer iteration(s)", flush=True)
        r = ref.solve_ccs(prob, shift=1, max_iter=size, print_level=1)
        out = {"seconds": r["seconds"], "num_iter": r["num_iter"], "nev_conv": r["nev_conv"],
               "phases": ref.phase_times()}
    else:
        if what == "cfg4":
            probs = P.config4_batch(size)
        else:
            probs = [p for _, p in P.config5_stream(size, 10.0)]
        secs, conv, iters = 0.0, 0, 0
        for prob in probs:
            print(f"{prob.name}: N = {prob.n}, nev {prob.nev}", flush=True)
            t0 = time.perf_counter()
            r = ref.solve_ccs(prob, shift=1)
            secs += time.perf_counter() - t0
            conv += r["nev_conv"]
            iters += r["num_iter"]
        out = {"seconds": secs, "nev_conv": conv, "num_iter": iters, "problems": len(probs)}
    out["threads"] = os.environ.get("OMP_NUM_THREADS")
    print(json.dumps(out))


if __name__ == "__main__":
    main()

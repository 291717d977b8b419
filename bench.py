"""Benchmark: converged eigenpairs/s of the GCG eigensolver on one MI355X.

Workload (default): the 3D 7-point problem on a 128^3 interior grid,
N = 2,097,152, A = 6I - Adj, B = (12I + Adj)/18, CCS FP64, the nev = 100
smallest pairs to 1e-8 relative residual with block size 120
(nevInit = nevMax = 220).  One step = one complete GCG solve with the
matrices resident in HBM (upload timed separately).

Two paths are run on the same problem:
  fp64  -- the reference-equivalent path: every operation FP64 on the GPU,
           the reference's default autoshift (shift = 1);
  mixed -- FP32 device BlockPCG in the MultiLinearSolver slot, FP64
           everywhere else, shift = 0 (the reference forbids autoshift
           with a user solver).
`value` is the mixed path's whole-job eigenpairs/s; the fp64 path's is
reported beside it.

The roofline object is for the dominant kernel class of the mixed path,
timed with HIP events on the library's stream over the timed region; its
`traffic` (measured HBM bytes per launch) is read from the hardware-counter
summary profiles/pmc_traffic_cfg2_mixed.txt when that file is present.
`--copy-probe` runs only a column copy of known size (the counter
calibration kernel) and exits.
The CPU baseline runs the reference GCGE (compiled from its sources in
oracle/_ref, OpenMP build) on the same workload for a bounded sample and
extrapolates with the GPU run's iteration count.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
FP32_PEAK_TFS = 157.3


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def tune_entries() -> int:
    """Entries of the launch-shape table in use (the library loads the
    table beside it when its first device context is created)."""
    import ctypes as C
    from gcge_amd import abi
    lib = abi.load()
    n = 0
    z = [C.c_int() for _ in range(5)] + [C.c_double(), C.c_double()]
    while lib.gcge_tune_entry(n, *[C.byref(x) for x in z]):
        n += 1
    return n


def make_problem(name: str):
    from gcge_amd import problems as P
    if name == "cfg2":
        return P.config2(), 128, 3
    if name == "cfg1":
        return P.config1(), 128, 2
    if name == "cfg3":
        return P.config3(), 160, None      # no closed form
    if name.startswith("3d_"):
        n = int(name[3:])
        return P.stencil_problem(n, 3, 100, 120, 220, 220, name=f"3d_{n}"), n, 3
    raise ValueError(name)


def make_grid_problem(name: str):
    """The grid form of make_problem's problem (assembled on the device)."""
    from gcge_amd import problems as P
    if name == "cfg2":
        return P.config2_grid()
    if name == "cfg1":
        return P.config1_grid()
    if name == "cfg3":
        return P.config3_grid()
    if name.startswith("3d_"):
        n = int(name[3:])
        return P.stencil_grid_problem(n, 3, 100, 120, 220, 220, f"3d_{n}")
    return None


def run_path(prob, n, dim, mode, steps, warmup, mats, want_check=True):
    from gcge_amd import abi, problems as P, solver as S
    lib = abi.load()
    # fp64: the reference's algorithm (autoshift, column-by-column MGS);
    # fp64_cholqr: the same with the block Cholesky QR orthogonalization;
    # mixed: FP32 BlockPCG without shift, Cholesky QR
    shift = 0 if mode == "mixed" else 1
    kw = dict(shift=shift, mode="mixed" if mode == "mixed" else "fp64", rand="device", seed=1,
              mats=mats, print_level=1, orth="cholqr" if mode != "fp64" else "mgs")
    for w in range(warmup):
        log(f"{mode}: warmup {w + 1}/{warmup}")
        S.solve_problem(prob, **kw)
    lib.gcge_hip_profile_reset()
    lib.gcge_hip_profile_enable(1)
    res, secs = None, []
    for k in range(steps):
        last = (k == steps - 1)
        res = S.solve_problem(prob, want_vectors=(want_check and last), **kw)
        secs.append(res.seconds)
        log(f"{mode}: step {k + 1}/{steps}: {res.seconds:.3f} s, iters {res.num_iter}, "
            f"nconv {res.nev_conv}, bpcg {res.bpcg_iters}")
    lib.gcge_hip_profile_enable(0)
    prof = prof_totals(lib.gcge_hip_profile_read)
    k = res.nev_conv
    out = {"seconds": secs, "num_iter": res.num_iter, "nev_conv": k,
           "bpcg_iters": res.bpcg_iters, "phases": res.phases, "kernels": prof}
    out["eval"] = [float(x) for x in res.eval[:k]]
    if want_check and res.evec is not None:
        if dim is not None:
            exact = P.analytic_eigenvalues(n, dim, k)
            out["max_rel_eig_err_vs_analytic"] = float(np.max(np.abs(res.eval[:k] - exact) /
                                                              exact))
        X = res.evec[:, :k]
        out["_evec"] = X  # kept for the principal angles between paths
        lam = res.eval[:k]
        R = prob.A @ X - (prob.B @ X) * lam[None, :]
        rr = np.sqrt(np.einsum("ij,ij->j", R, R)) / np.abs(lam)
        out["max_rel_residual"] = float(rr.max())
        out["eval_first_last"] = [float(lam[0]), float(lam[min(prob.nev, k) - 1])]
        del R
    return out


def copy_probe(reps: int = 5):
    """A plain device copy of known size through the table (MatDotMultiVec
    with a NULL matrix): 120 columns of 2,097,152 doubles, read once and
    written once, 2,013,265,920 bytes each way per launch."""
    import ctypes as C
    from gcge_amd import abi, problems as P
    from gcge_amd.solver import DeviceMatrix, HipTable
    t = HipTable()
    lib = t.lib
    prob = P.stencil_problem(128, 3, 100, 120, 220, 220)
    A = DeviceMatrix(prob.A)
    lib.gcge_hip_set_random_mode(abi.GCGE_HIP_RAND_DEVICE, 5)
    X, Y = t.create_multivec(120, A), t.create_multivec(120, A)
    t.slot("MultiVecSetRandomValue")(X, 0, 120, t.ops)
    s, e = (C.c_int * 2)(0, 0), (C.c_int * 2)(120, 120)
    lib.gcge_hip_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        t.slot("MatDotMultiVec")(None, X, Y, s, e, t.ops)
    lib.gcge_hip_synchronize()
    dt = (time.perf_counter() - t0) / reps
    nbytes = 120 * prob.n * 8
    print(json.dumps({"copy_probe": {"bytes_read": nbytes, "bytes_written": nbytes,
                                     "ms": 1e3 * dt, "GBs": 2 * nbytes / dt / 1e9}}))
    t.destroy_multivec(X, 120)
    t.destroy_multivec(Y, 120)


def rocprof_summary(d: str):
    """Condenses rocprofv3 CSV output under directory `d` (any depth) into
    one line per kernel: calls and average duration from *_kernel_stats.csv,
    mean counter values per dispatch from *_counter_collection.csv."""
    import csv
    import glob
    for p in sorted(glob.glob(os.path.join(d, "**", "*_kernel_stats.csv"), recursive=True)):
        print(f"# kernel stats: {os.path.relpath(p, d)}")
        print(f"{'calls':>8} {'total_ms':>10} {'avg_us':>10} {'pct':>6}  kernel")
        for r in csv.DictReader(open(p)):
            print(f"{int(r['Calls']):8d} {float(r['TotalDurationNs']) / 1e6:10.2f} "
                  f"{float(r['AverageNs']) / 1e3:10.2f} {float(r['Percentage']):6.2f}  "
                  f"{r['Name'][:110]}")
    for p in sorted(glob.glob(os.path.join(d, "**", "*_counter_collection.csv"),
                              recursive=True)):
        agg = {}
        for r in csv.DictReader(open(p)):
            key = (r["Kernel_Name"][:110], r["Counter_Name"])
            # one row per dispatch and counter
            a = agg.setdefault(key, [0, 0.0])
            a[0] += 1
            a[1] += float(r["Counter_Value"])
        print(f"# counters per dispatch: {os.path.relpath(p, d)}")
        print(f"{'dispatches':>10} {'counter':>12} {'mean':>16}  kernel")
        for (k, c), (cnt, tot) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
            print(f"{cnt:10d} {c:>12} {tot / cnt:16.1f}  {k}")


# kernel class of the library profiler -> device kernel name prefix
# (one profiler launch of these classes is one dispatch of one of the kernels)
# the library GEMM of the large LinearComb calls (rocBLAS's Tensile FP64 NN
# kernel that it picks for 2 M rows on MI355X; a trailing '*' matches a
# prefix of the full kernel name)
GEMM_KERNEL = "Cijk_Ailk_Bljk_DB_MT128x128x8_MI16x16x4x1*"
KERNEL_OF_CLASS = {"lincomb": ["k_lincomb_mfma"], "gemm": [GEMM_KERNEL],
                   "spmm64": ["k_spmm_sell", "k_stencil7"],
                   "spmm32": ["k_spmm_pw", "k_stencil_pw"],
                   "pcgvec": ["k_update_p", "k_update_xr"]}


def _counter_means(path: str, counters=("FETCH_SIZE", "WRITE_SIZE")):
    """{kernel name: (dispatches, mean)} from a --rocprof-summary output,
    over the lines of the named counters."""
    out = {}
    for ln in open(path):
        f = ln.split(None, 3)
        if len(f) == 4 and f[0].isdigit() and f[1] in counters:
            out[f[3].strip()] = (int(f[0]), float(f[2]))
    return out


def _short(kernel: str) -> str:
    """'void gcge::(anonymous namespace)::k_x<8>(long, ...)' -> 'k_x<8>'."""
    k = kernel.replace("(anonymous namespace)::", "")
    if k.startswith("void "):
        k = k[5:]
    return k.split("(")[0].split("::")[-1]


def _match(kernel: str, pf: str) -> bool:
    if pf.endswith("*"):
        return _short(kernel).startswith(pf[:-1])
    return _short(kernel).split("<")[0] == pf.split("<")[0] and (
        "<" not in pf or _short(kernel) == pf)


def _find(means: dict, prefix):
    prefixes = [prefix] if isinstance(prefix, str) else prefix
    hits = [(k, v) for k, v in means.items() for pf in prefixes if _match(k, pf)]
    if not hits:
        return None
    # several template instances: dispatch-weighted mean
    n = sum(v[0] for _, v in hits)
    return n, sum(v[0] * v[1] for _, v in hits) / n


def pmc_traffic(fetch: str, write: str, cfetch: str, cwrite: str, bench_json: str):
    """HBM bytes per launch from FETCH_SIZE/WRITE_SIZE summaries (KiB per
    dispatch), each scaled by what the counter reads for the copy probe of
    known size, next to the algorithmic bytes of the bench line."""
    copy_bytes = 120 * 2097152 * 8
    mf, mw = _counter_means(fetch), _counter_means(write)
    cf = _find(_counter_means(cfetch), "k_copy_cols")[1]
    cw = _find(_counter_means(cwrite), "k_copy_cols")[1]
    ff, fw = copy_bytes / (cf * 1024), copy_bytes / (cw * 1024)
    line = json.loads(open(bench_json).read().strip().splitlines()[-1])
    kern = line["paths"]["mixed"]["kernels"]
    print("# calibration: the copy probe (k_copy_cols) reads and writes "
          f"{copy_bytes:,} B per launch")
    print(f"#   FETCH_SIZE reads {cf:.1f} KiB -> factor {ff:.4f}")
    print(f"#   WRITE_SIZE reads {cw:.1f} KiB -> factor {fw:.4f}")
    print("# per kernel class of the bench line (mean over dispatches): FETCH_SIZE KiB, "
          "WRITE_SIZE KiB, calibrated bytes, algorithmic bytes")
    print(f"{'class':8s} {'kernels':24s} {'dispatches':>10} {'FETCH KiB':>12} {'WRITE KiB':>12} "
          f"{'measured B':>12} {'algorithmic B':>13} {'ratio':>6}")
    result = {}
    for cls, kname in KERNEL_OF_CLASS.items():
        if cls not in kern:
            continue
        a, b = _find(mf, kname), _find(mw, kname)
        if a is None or b is None:
            continue
        meas = a[1] * 1024 * ff + b[1] * 1024 * fw
        alg = kern[cls]["bytes"] / kern[cls]["launches"]
        result[cls] = meas
        print(f"{cls:8s} {'+'.join(kname):24s} {a[0]:10d} {a[1]:12.1f} {b[1]:12.1f} {meas:12.4e} "
              f"{alg:13.4e} {meas / alg:6.3f}")
    print("# top kernels by calibrated bytes per dispatch")
    for k in sorted(set(mf) & set(mw), key=lambda k: -(mf[k][1] * ff + mw[k][1] * fw))[:16]:
        print(f"{mf[k][0]:10d} {mf[k][1] * 1024 * ff + mw[k][1] * 1024 * fw:12.4e}  {_short(k)}")
    for cls, meas in result.items():
        print(f"traffic_bytes_per_launch {cls} {meas:.6e}")


# kernel class of the library profiler -> its MFMA kernel
MFMA_CLASSES = {"gram": "k_gram_mfma", "lincomb": "k_lincomb_mfma", "gemm": GEMM_KERNEL}


def pmc_mfma(summary: str, bench_json: str):
    """MFMA work per dispatch from SQ_INSTS_VALU_MFMA_MOPS_F64 (one count is
    512 FP64 flops) and the busy fraction of the matrix cores (the derived
    counter MfmaUtil), next to the algorithmic flops of the bench line."""
    mops = _counter_means(summary, ("SQ_INSTS_VALU_MFMA_MOPS_F64",))
    util = _counter_means(summary, ("MfmaUtil",))
    busy = _counter_means(summary, ("SQ_VALU_MFMA_BUSY_CYCLES",))
    line = json.loads(open(bench_json).read().strip().splitlines()[-1])
    kern = line["paths"]["mixed"]["kernels"]
    print("# per kernel class (mean over dispatches): MFMA flops counted, algorithmic flops "
          "(2*rows*m*n), their ratio, MfmaUtil (%)")
    print(f"{'class':8s} {'kernel':16s} {'dispatches':>10} {'MFMA flops':>12} "
          f"{'algorithmic':>12} {'ratio':>6} {'MfmaUtil %':>10} {'TF/s (events)':>13}")
    for cls, kname in MFMA_CLASSES.items():
        if cls not in kern:
            continue
        m, u = _find(mops, kname), _find(util, kname)
        if m is None:
            continue
        alg = kern[cls]["flops"] / kern[cls]["launches"]
        uu = u[1] if u is not None else float("nan")
        print(f"{cls:8s} {kname:16s} {m[0]:10d} {m[1] * 512:12.4e} {alg:12.4e} "
              f"{m[1] * 512 / alg:6.3f} {uu:10.1f} {kern[cls]['TFs']:13.1f}")
    print("# every kernel with MFMA work: dispatches, MFMA flops per dispatch, "
          "MFMA busy cycles per dispatch, MfmaUtil (%)")
    for k in sorted(mops, key=lambda k: -mops[k][1]):
        if mops[k][1] <= 0:
            continue
        b = busy.get(k, (0, float("nan")))[1]
        uu = util.get(k, (0, float("nan")))[1]
        print(f"{mops[k][0]:10d} {mops[k][1] * 512:12.4e} {b:12.4e} {uu:6.1f}  {_short(k)}")


def measured_traffic(kernel: str, config: str):
    """HBM bytes per launch of `kernel` from the committed counter summary
    of the 128^3 mixed solve (a line 'traffic_bytes_per_launch <kernel>
    <bytes>'); None for other workloads, whose traffic was not measured."""
    p = os.path.join(ROOT, "profiles", "pmc_traffic_cfg2_mixed.txt")
    if config != "cfg2" or not os.path.exists(p):
        return None
    for ln in open(p):
        f = ln.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
            return float(f[2])
    return None


def _ref_sample(what: str, size: int):
    """Runs oracle/ref_sample.py (the compiled reference on the CPU) in a
    child process; returns (json, cores) or (None, reason)."""
    refso = os.path.join(ROOT, "oracle", "_ref", "libgcge_ref_omp.so")
    if not os.path.exists(refso):
        return None, "oracle/_ref/libgcge_ref_omp.so not built"
    cores = min(os.cpu_count() or 1, 16)
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), MKL_NUM_THREADS="1")
    cmd = [sys.executable, os.path.join(ROOT, "oracle", "ref_sample.py"), refso, what,
           str(size)]
    log(f"cpu baseline: reference GCGE, {cores} threads, {what} sample of size {size}")
    # the reference's own monitor lines come through as they are printed
    p = subprocess.Popen(cmd, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                         text=True)
    last = None
    t0 = time.perf_counter()
    for ln in p.stdout:
        ln = ln.rstrip()
        if ln.startswith("{"):
            last = ln
        elif ln:
            log(f"reference: {ln[:160]}")
        if time.perf_counter() - t0 > 1200:
            p.kill()
            return None, "reference sample timed out"
    if p.wait() != 0 or last is None:
        return None, "reference sample failed"
    r = json.loads(last)
    r["cores"] = cores
    return r, None


def cpu_baseline(config, gpu_iters, nev_conv, iters_sample):
    """Reference GCGE (OpenMP build) on the same single-problem workload for
    a bounded number of outer iterations; extrapolated to the GPU run's
    iteration count."""
    r, why = _ref_sample(config, iters_sample)
    if r is None:
        return None, why
    t_init = r["phases"]["initX"]
    t_sample = r["seconds"]
    per_iter = (t_sample - t_init) / max(r["num_iter"], 1)
    t_full = t_init + per_iter * gpu_iters
    val = nev_conv / t_full
    size = "128^3" if config == "cfg2" else "160^3 27-point"
    return {"value": val, "unit": "eigenpairs/s", "cores": r["cores"], "kind": "reference",
            "sample": (f"reference GCGE (OpenMP build, MKL BLAS/LAPACK 1 thread) on the same "
                       f"{size} problem: InitializeX + {r['num_iter']} outer iteration(s) in "
                       f"{t_sample:.1f} s (initX {t_init:.1f} s, {per_iter:.1f} s/iteration), "
                       f"extrapolated to the GPU run's {gpu_iters} iterations"),
            "seconds_extrapolated": t_full}, None


def cpu_baseline_served(config, count):
    """Reference GCGE on the first `count` problems of the batched/served
    workload, solved one after another on all the host cores."""
    r, why = _ref_sample(config, count)
    if r is None:
        return None, why
    return {"value": r["nev_conv"] / r["seconds"], "unit": "eigenpairs/s", "cores": r["cores"],
            "kind": "reference",
            "sample": (f"reference GCGE (OpenMP build, FP64, autoshift on) on the first "
                       f"{r['problems']} problems of the same workload, one after another: "
                       f"{r['nev_conv']} pairs in {r['seconds']:.1f} s "
                       f"({r['num_iter']} outer iterations)")}, None


PROF_KINDS = ("spmm64", "spmm32", "gram", "lincomb", "blas1", "pcgvec", "gemm")


def roofline_of(kern, config):
    """The roofline object of the dominant kernel class (by total time)."""
    dom = max(kern, key=lambda k: kern[k]["total_ms"]) if kern else None
    if not dom:
        return None
    kd = kern[dom]
    # FP64 ridge point: 78.6 TFLOP/s / 8 TB/s = 9.8 flop/byte
    intensity = kd["flops"] / max(kd["bytes"], 1.0)
    common = {"kernel": dom, "avg_launch_us": kd["avg_us"], "launches": kd["launches"],
              "algorithmic_bytes_per_launch": kd["bytes"] / kd["launches"],
              "algorithmic_flops_per_launch": kd["flops"] / kd["launches"],
              "flop_per_byte": intensity,
              "traffic": measured_traffic(dom, config)}
    if intensity > FP64_PEAK_TFS / (HBM_PEAK_GBS / 1e3):
        return {"bound": "mfma", "achieved": kd["TFs"], "peak": FP64_PEAK_TFS,
                "unit": "TFLOP/s", "frac": kd["TFs"] / FP64_PEAK_TFS, **common}
    return {"bound": "hbm", "achieved": kd["GBs"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": kd["GBs"] / HBM_PEAK_GBS, **common}


def prof_totals(read):
    import ctypes as C
    prof = {}
    for kind, name in enumerate(PROF_KINDS):
        la, ms, by, fl = C.c_int64(), C.c_double(), C.c_double(), C.c_double()
        read(kind, C.byref(la), C.byref(ms), C.byref(by), C.byref(fl))
        if la.value:
            prof[name] = {"launches": la.value, "total_ms": ms.value,
                          "avg_us": 1e3 * ms.value / la.value,
                          "bytes": by.value, "flops": fl.value,
                          "GBs": by.value / (ms.value * 1e-3) / 1e9 if ms.value else 0.0,
                          "TFs": fl.value / (ms.value * 1e-3) / 1e12 if ms.value else 0.0}
    return prof


def served(args):
    """The batched (cfg4) and served (cfg5) configurations through the
    solve service: one JSON line with throughput and latency."""
    from gcge_amd import problems as P
    from gcge_amd.service import Service
    import scipy.sparse.linalg  # noqa: F401  (residual checks below)
    t0 = time.perf_counter()
    if args.config == "cfg4":
        items = [(0.0, p) for p in P.config4_batch(args.count)]
        what = (f"cfg4: {args.count} 2D variable-coefficient 5-point problems, N in "
                f"[2048, 65536], nev in 8..64, submitted together")
    else:
        items = P.config5_stream(args.count, args.qps)
        what = (f"cfg5: {args.count} requests, Poisson arrivals at {args.qps} per s, 80 % "
                f"cfg4-like, 20 % 3D 7-point 48^3..96^3 with nev 32..128")
    log(f"{args.config}: generated {len(items)} problems in {time.perf_counter() - t0:.1f}s")
    check_every = max(1, len(items) // 8)
    res = [None] * len(items)
    from gcge_amd import abi
    lib = abi.load()
    lib.gcge_hip_profile_all(1)  # every worker's and the batcher's launches, timed
    with Service(workers=args.workers, hbm_budget=args.hbm_budget,
                 batching=args.batching == "on", orth=args.service_orth) as svc:
        # warm the workers' kernels and handles on a small problem each
        warm = [svc.submit(P.stencil_problem(24, 2, 8), seed=1) for _ in range(args.workers)]
        for t in warm:
            svc.wait(t)
        tickets = []
        start = time.perf_counter()
        for k, (t_arr, prob) in enumerate(items):
            delay = start + t_arr - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets.append(svc.submit(prob, mode="mixed", seed=1000 + k,
                                      want_vectors=(k % check_every == 0)))
        for k, t in enumerate(tickets):
            res[k] = svc.wait(t)
        wall = time.perf_counter() - start
        st = svc.stats()
        bst = svc.batch_stats()
    kern = prof_totals(lib.gcge_hip_profile_read_all)  # the workers have been released
    lib.gcge_hip_profile_all(0)
    roof = roofline_of(kern, args.config)
    cpu, why = (None, "disabled")
    if args.cpu_baseline == "on":
        cpu, why = cpu_baseline_served(args.config, args.cpu_count)
    lat = np.array([r.total_seconds for r in res])
    conv = sum(r.nev_conv for r in res)
    short = sum(1 for (_, p), r in zip(items, res) if r.nev_conv < p.nev)
    worst = 0.0
    for k, ((_, p), r) in enumerate(zip(items, res)):
        if r.evec is not None:
            R = p.A @ r.evec - (p.B @ r.evec) * r.eval[None, :]
            rr = np.linalg.norm(R, axis=0) / np.abs(r.eval)
            worst = max(worst, float(rr.max()))
    line = {
        "metric": "converged eigenpairs/s at 1e-8 residual, " +
                  ("batched" if args.config == "cfg4" else "served") +
                  "; p50/p99 request latency",
        "value": conv / wall, "unit": "eigenpairs/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": 1e3 * wall, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": what, "path": "mixed", "workers": args.workers,
                   "batched_inner_solves": args.batching == "on",
                   "orthonormalization": args.service_orth,
                   "hbm_budget_bytes": args.hbm_budget, "parallelism": "single GPU"},
        "batched_rounds": bst,
        "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)),
                       "p99": 1e3 * float(np.percentile(lat, 99))},
        "requests": len(items), "eigenpairs_converged": conv,
        "requests_short_of_nev": short, "max_concurrent": st["max_concurrent"],
        "max_rel_residual_sampled": worst,
        "queue_s_p50": float(np.percentile([r.queue_seconds for r in res], 50)),
        "solve_s_p50": float(np.percentile([r.solve_seconds for r in res], 50)),
        # a stream is kept up with when the last request is done within
        # the arrivals' span plus a median request's latency
        "arrival_span_s": float(items[-1][0]),
        "served_requests_per_s": len(items) / wall,
        "keeps_up": (None if args.config == "cfg4" else
                     bool(wall <= 1.05 * float(items[-1][0]) + float(np.percentile(lat, 50)))),
        "kernels": kern,
        "roofline": roof, "cpu_baseline": cpu,
        "tuning_entries": tune_entries(),
    }
    if cpu is None:
        line["cpu_baseline_note"] = why
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="cfg2")
    ap.add_argument("--modes", default="fp64,fp64_cholqr,mixed")
    ap.add_argument("--cpu-count", type=int, default=6,
                    help="cfg4/cfg5: problems in the reference's CPU sample")
    ap.add_argument("--cpu-baseline", default="on", choices=("on", "off"))
    ap.add_argument("--cpu-iters", type=int, default=1)
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--count", type=int, default=512, help="cfg4/cfg5 requests")
    ap.add_argument("--qps", type=float, default=20.0, help="cfg5 arrival rate")
    ap.add_argument("--workers", type=int, default=16, help="cfg4/cfg5 service workers")
    ap.add_argument("--service-orth", default="cholqr_host", choices=("cholqr", "cholqr_host"),
                    help="cfg4/cfg5 mixed requests: Cholesky QR factorized on the device or host")
    ap.add_argument("--batching", default="on", choices=("on", "off"),
                    help="cfg4/cfg5: batched inner solves across requests")
    ap.add_argument("--hbm-budget", type=float, default=200e9)
    ap.add_argument("--tune-table", default=None,
                    help="launch-shape table to load (default: the one beside the library; "
                         "'none': built-in heuristics only)")
    ap.add_argument("--rocprof-summary", metavar="DIR")
    ap.add_argument("--pmc-traffic", nargs=5,
                    metavar=("FETCH", "WRITE", "COPY_FETCH", "COPY_WRITE", "BENCH_JSON"))
    ap.add_argument("--pmc-mfma", nargs=2, metavar=("SUMMARY", "BENCH_JSON"))
    args = ap.parse_args()
    if args.tune_table is not None:
        # read by the library when its first device context is created
        os.environ["GCGE_TUNE_TABLE"] = "" if args.tune_table == "none" else args.tune_table
    if args.rocprof_summary:
        rocprof_summary(args.rocprof_summary)
        return
    if args.pmc_traffic:
        pmc_traffic(*args.pmc_traffic)
        return
    if args.pmc_mfma:
        pmc_mfma(*args.pmc_mfma)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.copy_probe:
        copy_probe()
        return
    if args.config in ("cfg4", "cfg5"):
        served(args)
        return
    from gcge_amd import abi, problems as P, solver as S
    abi.load().gcge_host_set_print_target(1)  # monitor lines to stderr
    t0 = time.perf_counter()
    prob, n, dim = make_problem(args.config)
    log(f"problem {prob.name}: N={prob.n}, nnz={prob.A.nnz}, gen {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    mats = (S.DeviceMatrix(prob.A, fp32=True), S.DeviceMatrix(prob.B, fp32=True))
    t_upload = time.perf_counter() - t0
    log(f"upload {t_upload:.2f}s")
    # the same matrices assembled on the device from the grid and the node
    # coefficients; they must equal the uploaded ones bit for bit
    assembly = None
    gp = make_grid_problem(args.config)
    if gp is not None:
        abi.load().gcge_hip_synchronize()
        t0 = time.perf_counter()
        gA = S.device_operator(gp.A, fp32=True)
        gB = S.device_operator(gp.B, fp32=True)
        abi.load().gcge_hip_synchronize()
        t_asm = time.perf_counter() - t0
        same = gA.equal(mats[0]) and gB.equal(mats[1])
        gA.free()
        gB.free()
        assembly = {"seconds": t_asm, "equal_to_upload": same}
        log(f"device assembly {t_asm:.3f}s, equal to the upload: {same}")
        if not same:
            raise SystemExit("device-assembled matrices differ from the uploaded ones")
    if args.warmup == 0:
        # load the kernels and the dense eigensolver once (not a step)
        pp = P.stencil_problem(12, 3, 30)
        S.solve_problem(pp, shift=1, rand="device", eig="device")
        S.solve_problem(pp, shift=0, mode="mixed", rand="device")

    results = {}
    for mode in args.modes.split(","):
        results[mode] = run_path(prob, n, dim, mode, args.steps, args.warmup, mats)

    head = "mixed" if "mixed" in results else list(results)[0]
    hr = results[head]
    secs = np.array(hr["seconds"])
    value = hr["nev_conv"] / secs.mean()
    # dominant kernel class of the headline path
    roof = roofline_of(hr["kernels"], args.config)
    cpu, why = (None, "disabled")
    if args.cpu_baseline == "on" and args.config in ("cfg2", "cfg3"):
        cpu, why = cpu_baseline(args.config, hr["num_iter"], hr["nev_conv"], args.cpu_iters)
    line = {
        "metric": "converged eigenpairs/s at 1e-8 residual, mixed vs FP64; p50/p99 solve latency",
        "value": value, "unit": "eigenpairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * float(secs.mean()),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32" if head == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": (f"{prob.name}: generalized eigenproblem, N={prob.n}, "
                                f"nnz(A)={prob.A.nnz}, nnz(B)={prob.B.nnz}, nev={prob.nev}, "
                                f"block {prob.block_size}, nevInit=nevMax={prob.nev_max}, "
                                f"tol 1e-8 rel"),
                   "path": head, "parallelism": "single GPU"},
        "latency_ms": {"p50": 1e3 * float(np.percentile(secs, 50)),
                       "p99": 1e3 * float(np.percentile(secs, 99))},
        "paths": {m: {"eigenpairs_per_s": r["nev_conv"] / float(np.mean(r["seconds"])),
                      "seconds": r["seconds"], "num_iter": r["num_iter"],
                      "nev_conv": r["nev_conv"], "bpcg_iters": r["bpcg_iters"],
                      "max_rel_eig_err_vs_analytic": r.get("max_rel_eig_err_vs_analytic"),
                      "max_rel_residual": r.get("max_rel_residual"),
                      "phases_s": r["phases"], "kernels": r["kernels"],
                      "eval_first_last": r.get("eval_first_last")}
                  for m, r in results.items()},
        "upload_s": t_upload,
        "device_assembly": assembly,
        "tuning_entries": tune_entries(),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if "fp64_cholqr" in results and "mixed" in results:
        # the same orthogonalization on both sides: precision (and the
        # autoshift that FP32 inner solves cannot use) alone
        line["speedup_mixed_vs_fp64_cholqr"] = (np.mean(results["fp64_cholqr"]["seconds"]) /
                                                np.mean(results["mixed"]["seconds"]))
    if "fp64" in results and "mixed" in results:
        line["speedup_mixed_vs_fp64"] = (np.mean(results["fp64"]["seconds"]) /
                                         np.mean(results["mixed"]["seconds"]))
        a, b = (np.array(results[m]["eval"]) for m in ("fp64", "mixed"))
        k = min(len(a), len(b))
        line["max_rel_eig_diff_fp64_vs_mixed"] = float(np.max(np.abs(a[:k] - b[:k]) / a[:k]))
        X1, X2 = results["fp64"].get("_evec"), results["mixed"].get("_evec")
        if X1 is not None and X2 is not None:
            from gcge_amd import checks
            kc = checks.cluster_cut(a, min(X1.shape[1], X2.shape[1]))
            line["subspace"] = {
                "pairs": kc,
                "max_sine_fp64_vs_mixed": checks.max_subspace_sine(X1[:, :kc], X2[:, :kc],
                                                                   prob.B),
                "note": "largest principal angle in the B inner product over the first "
                        "pairs up to a cluster boundary"}
    for r in results.values():
        r.pop("_evec", None)
    if cpu is None:
        line["cpu_baseline_note"] = why
    print(json.dumps(line))


if __name__ == "__main__":
    main()

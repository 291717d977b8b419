"""The headline workloads at full size on the GPU.

* 128^3 7-point problem (N = 2,097,152, nev 100, block 120, nevInit =
  nevMax = 220, tol 1e-8): the FP64 path with the libc initial block
  against the compiled reference's run on the same problem
  (tests/golden/ref_cfg2.json, written by tests/golden/make_golden.py
  --cfg2 from the reference's CCS app and GCG driver,
  src/ops_eig_sol_gcg.c:1259-1562, app/app_ccs.c:50-139, srand(0) as at
  gcge.c:508) to 1e-9; both paths against the closed-form spectrum to
  1e-10, every returned pair's relative residual <= 1e-8 (the check of
  src/ops_eig_sol_gcg.c:239-240), B-orthonormal eigenvectors, and the
  principal angles between the two paths' eigenvector sets over whole
  clusters.
* the 512-problem batch (2D variable coefficients, N in [2048, 65536],
  nev 8..64) through the solve service in mixed precision: every request
  converges with every pair's residual <= 1e-8, dense LAPACK (dsygv)
  eigenvalues on the smallest problems, and the FP64 path's eigenvalues on
  the largest ones.
* a served stream of 100 requests (Poisson arrivals at 10 per second, 20 %
  3D 7-point problems with closed-form spectra).
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sla

import gcge_oracle as O
from gcge_amd import checks, problems as P
from gcge_amd import solver as S
from gcge_amd.service import Service

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_cfg2.json")


def b_orth_error(B, X):
    BX = B @ X
    return float(np.max(np.abs(X.T @ BX - np.eye(X.shape[1]))))


@pytest.fixture(scope="module")
def cfg2_runs(hip):
    prob = P.config2_grid()
    out = {}
    for mode in ("fp64", "mixed"):
        print(f"# 128^3 solve, {mode} path", flush=True)
        kw = dict(mode=mode, want_vectors=True)
        if mode == "fp64":
            kw.update(shift=1, rand="libc", seed=0)
        else:
            kw.update(shift=0, rand="device", seed=7)
        out[mode] = S.solve_problem(prob, **kw)
    A, B = P.stencil_pair(128, 3)
    return prob, A, B, out


@pytest.mark.timeout(900)
def test_cfg2_fp64_matches_reference_run(cfg2_runs):
    prob, A, B, out = cfg2_runs
    gold = json.load(open(GOLD))["run"]
    assert (gold["nev"], gold["block_size"], gold["nev_init"], gold["nev_max"]) == \
        (prob.nev, prob.block_size, prob.nev_init, prob.nev_max)
    r = out["fp64"]
    assert r.nev_conv >= prob.nev and gold["nev_conv"] >= prob.nev
    k = min(r.nev_conv, gold["nev_conv"])
    ref = np.array(gold["eval"][:k])
    err = np.max(np.abs(r.eval[:k] - ref) / ref)
    print(f"# fp64 vs reference run: {k} pairs, max rel diff {err:.2e}; iterations "
          f"{r.num_iter} vs {gold['num_iter']}, converged {r.nev_conv} vs {gold['nev_conv']}")
    assert err <= 1e-9
    assert abs(r.num_iter - gold["num_iter"]) <= 6


@pytest.mark.timeout(900)
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_cfg2_analytic_residual_orthonormality(cfg2_runs, mode):
    prob, A, B, out = cfg2_runs
    r = out[mode]
    k = r.nev_conv
    assert k >= prob.nev
    exact = P.analytic_eigenvalues(128, 3, k)
    err = np.max(np.abs(r.eval[:k] - exact) / exact)
    X = r.evec[:, :k]
    rr = O.relative_residuals(A, B, r.eval[:k], X)
    orth = b_orth_error(B, X)
    print(f"# {mode}: {k} pairs, {r.num_iter} iterations, {r.seconds:.2f} s, max rel error "
          f"{err:.2e}, max rel residual {rr.max():.2e}, max|X^T B X - I| {orth:.2e}")
    assert err <= 1e-10
    assert rr.max() <= 1.01e-8
    assert orth <= 1e-10


@pytest.mark.timeout(900)
def test_cfg2_subspace_angles_between_paths(cfg2_runs):
    prob, A, B, out = cfg2_runs
    r1, r2 = out["fp64"], out["mixed"]
    k = checks.cluster_cut(r1.eval, min(r1.nev_conv, r2.nev_conv))
    assert k >= prob.nev
    s = checks.max_subspace_sine(r1.evec[:, :k], r2.evec[:, :k], B)
    print(f"# largest principal angle (B inner product) over the first {k} pairs: sine {s:.2e}")
    assert s <= 1e-6


def residual_ok(p, r):
    rr = O.relative_residuals(p.A, p.B, r.eval, r.evec)
    return float(rr.max())


@pytest.mark.timeout(900)
def test_cfg4_full_batch(hip):
    probs = P.config4_batch(512)
    small = [i for i, p in enumerate(probs) if p.n <= 3000]
    large = sorted(range(len(probs)), key=lambda i: -probs[i].n)[:3]
    assert len(small) >= 2
    worst, conv = 0.0, 0
    evals = {}
    with Service(workers=8) as svc:
        tickets = [svc.submit(p, mode="mixed", seed=1000 + i, want_vectors=True)
                   for i, p in enumerate(probs)]
        for i, (p, t) in enumerate(zip(probs, tickets)):
            r = svc.wait(t)
            assert r.nev_conv >= p.nev, (p.name, r.nev_conv, p.nev)
            worst = max(worst, residual_ok(p, r))
            conv += r.nev_conv
            if i in small or i in large:
                evals[i] = r.eval
            r.evec = None
            if i % 64 == 63:
                print(f"# {i + 1} of 512 done", flush=True)
        fp64 = {i: svc.wait(svc.submit(probs[i], mode="fp64", seed=1000 + i)) for i in large}
    print(f"# 512 problems: {conv} pairs converged, max rel residual {worst:.2e}")
    assert worst <= 1.01e-8
    for i in small:
        p = probs[i]
        k = len(evals[i])
        ref = sla.eigh(p.A.toarray(), p.B.toarray(), eigvals_only=True,
                       subset_by_index=[0, k - 1])
        assert np.max(np.abs(evals[i] - ref) / ref) <= 1e-9, p.name
    for i in large:
        k = min(len(evals[i]), fp64[i].nev_conv)
        assert k >= probs[i].nev
        d = np.max(np.abs(evals[i][:k] - fp64[i].eval[:k]) / fp64[i].eval[:k])
        assert d <= 1e-9, (probs[i].name, d)


@pytest.mark.timeout(900)
def test_cfg5_served_stream(hip):
    import time
    items = P.config5_stream(100, 10.0)
    res = []
    with Service(workers=8) as svc:
        tickets = []
        start = time.perf_counter()
        for k, (t_arr, prob) in enumerate(items):
            delay = start + t_arr - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets.append(svc.submit(prob, mode="mixed", seed=5000 + k, want_vectors=True))
        for (t_arr, p), t in zip(items, tickets):
            r = svc.wait(t)
            assert r.nev_conv >= p.nev, (p.name, r.nev_conv)
            rr = residual_ok(p, r)
            assert rr <= 1.01e-8, (p.name, rr)
            if p.name.startswith("cfg5_3d"):
                n = round(p.n ** (1 / 3))
                exact = P.analytic_eigenvalues(n, 3, r.nev_conv)
                assert np.max(np.abs(r.eval - exact) / exact) <= 1e-10, p.name
            res.append(r)
            r.evec = None
        st = svc.stats()
    lat = np.array([r.total_seconds for r in res])
    print(f"# 100 requests: max concurrent {st['max_concurrent']}, latency p50 "
          f"{np.percentile(lat, 50):.2f} s, p99 {np.percentile(lat, 99):.2f} s")
    assert st["max_concurrent"] >= 2


@pytest.mark.timeout(900)
def test_cfg2_reference_driver_over_hip_table(hip, reflib):
    """The literal drop-in at the headline size: the reference's own GCG
    driver (src/ops_eig_sol_gcg.c:1259-1562, compiled from its sources into
    oracle/_ref) runs unchanged over the table OPS_HIP_Set fills, on the
    128^3 problem with the plug-in's call sequence and srand(0), and
    reproduces the reference's CPU run (tests/golden/ref_cfg2.json): the same
    eigenvalues to 1e-9 and nearly the same iteration count."""
    import ctypes as C
    from gcge_amd import abi
    gold = json.load(open(GOLD))["run"]
    prob = P.config2_grid()
    dA, dB = S.device_operator(prob.A), S.device_operator(prob.B)
    ops = reflib.empty_table()
    hip.OPS_HIP_Set(C.cast(C.c_void_p(ops), abi.OPS_p))
    reflib.setup(ops)
    hip.gcge_hip_set_random_mode(abi.GCGE_HIP_RAND_LIBC, 0)
    r = reflib.solve(ops, dA.handle.value, dB.handle.value, 128 ** 3, prob.nev, prob.block_size,
                     prob.nev_init, prob.nev_max, (prob.tol_abs, prob.tol_rel), prob.gap_min,
                     1, 1000, host_vectors=False)
    k = min(r["nev_conv"], gold["nev_conv"])
    ref = np.array(gold["eval"][:k])
    err = np.max(np.abs(r["eval"][:k] - ref) / ref)
    exact = P.analytic_eigenvalues(128, 3, k)
    err_a = np.max(np.abs(r["eval"][:k] - exact) / exact)
    print(f"# reference driver over the HIP table: {r['seconds']:.2f} s, {r['num_iter']} "
          f"iterations (CPU run {gold['num_iter']}), {r['nev_conv']} converged "
          f"({gold['nev_conv']}), max rel diff {err:.2e}, vs analytic {err_a:.2e}")
    assert r["nev_conv"] >= prob.nev
    assert err <= 1e-9 and err_a <= 1e-10
    assert abs(r["num_iter"] - gold["num_iter"]) <= 3

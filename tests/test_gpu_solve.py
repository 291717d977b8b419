"""End-to-end parity of the GPU eigensolver.

* This library's driver over the HIP table (the FP64 path) and with the
  FP32 device BlockPCG (the mixed path), against the closed-form spectra,
  the reference's recorded runs (tests/golden/ref_runs.json) and the
  residual bound the driver promises (src/ops_eig_sol_gcg.c:239-240).
* The reference driver itself (oracle/_ref) running unchanged over the HIP
  table: the drop-in.
"""
import ctypes as C
import json
import os
import sys

import numpy as np
import pytest

import gcge_oracle as O
from gcge_amd import abi, problems as P
from gcge_amd import solver as S

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_runs.json")))


def golden(name, shift, app="ccs"):
    for r in GOLD["runs"]:
        if r["name"] == name and r["shift"] == shift and r["app"] == app:
            return r
    raise KeyError(name)


def check_pairs(prob, res, dim, n, iter_ref=None, iter_slack=6, nconv_ref=None):
    k = res.nev_conv
    assert k >= prob.nev
    lam = res.eval[:k]
    exact = P.analytic_eigenvalues(n, dim, k)
    rel = np.abs(lam - exact) / exact
    assert rel.max() <= 1e-10, rel.max()
    if res.evec is not None:
        rr = O.relative_residuals(prob.A, prob.B, lam, res.evec[:, :k])
        assert rr.max() <= 1e-8 * 1.01, rr.max()
        # B-orthonormal eigenvectors (the B inner product of src/ops_orth.c);
        # a block that lost B-orthogonality shows here before it shows as a
        # spurious eigenvalue
        X = res.evec[:, :k]
        G = X.T @ (prob.B @ X)
        assert np.abs(G - np.eye(k)).max() <= 1e-10
    if iter_ref is not None:
        assert abs(res.num_iter - iter_ref) <= iter_slack, (res.num_iter, iter_ref)
    if nconv_ref is not None:
        # how many pairs beyond nev pass the test in the last iteration
        # depends on the rounding path (the projected eigensolver here is
        # not the reference's dsyevx)
        assert abs(res.nev_conv - nconv_ref) <= 3, (res.nev_conv, nconv_ref)


@pytest.mark.parametrize("shift", [1, 0])
def test_fp64_cfg1(hip, shift):
    prob = P.config1()
    g = golden("cfg1_2d_128", shift)
    res = S.solve_problem(prob, shift=shift, want_vectors=True)
    check_pairs(prob, res, 2, 128, g["num_iter"], 6, g["nev_conv"])
    k = min(res.nev_conv, g["nev_conv"])
    ref = np.array(g["eval"][:k])
    assert np.max(np.abs(res.eval[:k] - ref) / ref) <= 1e-9


def test_fp64_3d_small(hip):
    prob = P.stencil_problem(12, 3, 30)
    g = golden("3d_12", 1)
    res = S.solve_problem(prob, shift=1, want_vectors=True)
    check_pairs(prob, res, 3, 12, g["num_iter"], 8, g["nev_conv"])


def test_mixed_cfg1(hip):
    prob = P.config1()
    res = S.solve_problem(prob, shift=0, mode="mixed", want_vectors=True)
    check_pairs(prob, res, 2, 128)


@pytest.mark.parametrize("orth", ["cholqr", "mgs"])
def test_mixed_3d(hip, orth):
    prob = P.stencil_problem(16, 3, 40)
    res = S.solve_problem(prob, shift=0, mode="mixed", want_vectors=True, orth=orth)
    check_pairs(prob, res, 3, 16)
    if orth == "cholqr":
        assert res.cholqr_blocks >= res.num_iter and res.cholqr_fallbacks == 0


def test_fp64_cholqr(hip):
    """The FP64 path with the block Cholesky QR orthogonalization."""
    prob = P.stencil_problem(12, 3, 30)
    g = golden("3d_12", 1)
    res = S.solve_problem(prob, shift=1, want_vectors=True, orth="cholqr")
    check_pairs(prob, res, 3, 12, g["num_iter"], 10)


def test_device_random_and_eig(hip):
    """Device RNG for the initial block and the device dense eigensolver:
    a different trajectory, the same pairs."""
    prob = P.stencil_problem(32, 2, 10)
    res = S.solve_problem(prob, shift=1, rand="device", seed=3, eig="device",
                          want_vectors=True)
    check_pairs(prob, res, 2, 32)


def test_dropin_reference_driver(hip, reflib):
    """The reference GCG driver (compiled from its sources, oracle/_ref)
    runs unchanged over the table filled by OPS_HIP_Set."""
    prob = P.stencil_problem(32, 2, 10)
    g = golden("2d_32", 1)
    ops = reflib.empty_table()
    hip.OPS_HIP_Set(C.cast(C.c_void_p(ops), abi.OPS_p))
    reflib.setup(ops)
    dA, dB = S.DeviceMatrix(prob.A), S.DeviceMatrix(prob.B)
    hip.gcge_hip_set_random_mode(abi.GCGE_HIP_RAND_LIBC, 0)
    r = reflib.solve(ops, dA.handle.value, dB.handle.value, prob.n, prob.nev, prob.block_size,
                     prob.nev_init, prob.nev_max, (prob.tol_abs, prob.tol_rel), prob.gap_min,
                     1, 1000, host_vectors=False)
    k = r["nev_conv"]
    assert k >= prob.nev
    exact = P.analytic_eigenvalues(32, 2, k)
    assert np.max(np.abs(r["eval"][:k] - exact) / exact) <= 1e-10
    assert abs(r["num_iter"] - g["num_iter"]) <= 4


def vc_problem(name, n, kind, nev):
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    from make_golden import problem
    return problem(name, n, kind, nev)


@pytest.mark.parametrize("name,n,kind,nev", [("vc2d_40", 40, "vc2", 16),
                                             ("vc27_16", 16, "vc27", 20)])
def test_fp64_variable_coefficients(hip, name, n, kind, nev):
    """No closed form: the compiled reference's runs and dense dsygv."""
    prob = vc_problem(name, n, kind, nev)
    g = golden(name, 1)
    res = S.solve_problem(prob, shift=1, want_vectors=True)
    assert res.nev_conv >= nev
    assert abs(res.num_iter - g["num_iter"]) <= 8
    k = min(res.nev_conv, g["nev_conv"])
    ref = np.array(g["eval"][:k])
    assert np.max(np.abs(res.eval[:k] - ref) / ref) <= 1e-9
    rr = O.relative_residuals(prob.A, prob.B, res.eval[:res.nev_conv],
                              res.evec[:, :res.nev_conv])
    assert rr.max() <= 1.01e-8


def test_mixed_variable_coefficients_27pt(hip):
    prob = vc_problem("vc27_16", 16, "vc27", 20)
    res = S.solve_problem(prob, shift=0, mode="mixed", want_vectors=True)
    k = res.nev_conv
    assert k >= 20
    ex = O.dense_generalized_eigenvalues(prob.A, prob.B, k)
    assert np.max(np.abs(res.eval[:k] - ex) / ex) <= 1e-9
    rr = O.relative_residuals(prob.A, prob.B, res.eval[:k], res.evec[:, :k])
    assert rr.max() <= 1.01e-8


def test_27pt_32cubed_fp64_and_mixed_agree(hip):
    """The config-3 family at 32^3 (N = 32768, rows of 27 entries: the
    32-wide SELL path): the two precisions agree and meet the residual."""
    A, B = P.varcoef27_pair(32, 1.0, np.random.default_rng([P.SEED_BASE, 3]))
    bs, ni, nm = P.slepc_block_policy(40)
    prob = P.CCSProblem("cfg3_27pt_32", A, B, 40, bs, ni, nm)
    r64 = S.solve_problem(prob, shift=1, want_vectors=True)
    rmx = S.solve_problem(prob, shift=0, mode="mixed", want_vectors=True)
    k = min(r64.nev_conv, rmx.nev_conv)
    assert k >= 40
    assert np.max(np.abs(r64.eval[:k] - rmx.eval[:k]) / r64.eval[:k]) <= 1e-9
    for r in (r64, rmx):
        rr = O.relative_residuals(A, B, r.eval[:r.nev_conv], r.evec[:, :r.nev_conv])
        assert rr.max() <= 1.01e-8


def test_eps_entry_point_matches_reference_run(hip):
    """gcge_eps_solve reproduces the plug-in's sequence: block policy,
    tolerances, autoshift, srand(0) -- the reference's recorded run on the
    same problem."""
    prob = P.stencil_problem(32, 2, 10)
    g = golden("2d_32", 1)
    r = S.eps_solve(prob.A, prob.B, 10, want_vectors=True)
    assert r.reason == abi.GCGE_EPS_CONVERGED_TOL and r.nconv == 10
    assert abs(r.its - (g["num_iter"] + 1)) <= 3
    ref = np.array(g["eval"][:10])
    assert np.max(np.abs(r.eigr - ref) / ref) <= 1e-9
    rr = O.relative_residuals(prob.A, prob.B, r.eigr, r.evec)
    assert rr.max() <= 1.01e-8


def test_eps_entry_point_mixed_and_diverged(hip):
    prob = P.stencil_problem(32, 2, 10)
    r = S.eps_solve(prob.A, prob.B, 10, mode="mixed")
    exact = P.analytic_eigenvalues(32, 2, 10)
    assert r.nconv == 10 and np.max(np.abs(r.eigr - exact) / exact) <= 1e-10
    r1 = S.eps_solve(prob.A, prob.B, 10, max_it=1)
    assert r1.reason == abi.GCGE_EPS_DIVERGED_ITS and r1.its == 1 and r1.nconv < 10


def test_mixed_cholqr_degenerate_cluster(hip):
    """3D 85^3, nev 32, block 6 (a request of the served stream): the 6-fold
    eigenvalue cluster at position 27 makes new directions nearly dependent
    on converged ones.  Block Cholesky QR normalizes what remains of them,
    so the earlier columns are removed once more between its two passes;
    without that the Rayleigh-Ritz step returned spurious near-zero
    eigenvalues and the solve stopped at 26 pairs after 1000 iterations."""
    prob = P.stencil_problem(85, 3, 32, name="3d85_nev32")
    assert (prob.block_size, prob.nev_init, prob.nev_max) == (6, 18, 50)
    r = S.solve_problem(prob, shift=0, mode="mixed", rand="device", seed=1072, max_iter=400,
                        orth="cholqr", want_vectors=True)
    k = r.nev_conv
    assert k >= prob.nev
    exact = P.analytic_eigenvalues(85, 3, k)
    assert np.max(np.abs(r.eval[:k] - exact) / exact) <= 1e-10
    X = r.evec[:, :k]
    G = X.T @ (prob.B @ X)
    assert np.abs(G - np.eye(k)).max() <= 1e-10


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_eigenvector_subspace_angles_vs_dense(hip, mode):
    """Principal angles (B inner product) between the computed eigenvectors
    and dense LAPACK's (scipy.linalg.eigh of the pair), over whole
    eigenvalue clusters: residual 1e-8 over relative gaps of about 1e-2
    bounds them near 1e-6."""
    import scipy.linalg as sla
    from gcge_amd import checks
    prob = P.stencil_problem(24, 2, 10)
    r = S.solve_problem(prob, shift=1 if mode == "fp64" else 0, mode=mode, rand="device", seed=4,
                        want_vectors=True)
    w, V = sla.eigh(prob.A.toarray(), prob.B.toarray())
    k = checks.cluster_cut(w, min(r.nev_conv, 40))
    assert k >= prob.nev
    s = checks.max_subspace_sine(r.evec[:, :k], V[:, :k], prob.B)
    assert s <= 1e-6, s

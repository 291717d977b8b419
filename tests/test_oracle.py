"""Pins the oracle (oracle/gcge_oracle.py) before it is trusted.

* slots: against explicit loops, and against the compiled reference's
  own slots called through its table (oracle/_ref);
* BlockPCG: known answers on SPD systems;
* MGS: Q^T B Q = I, span preservation, dependent-column dropping;
* the driver: the reference's recorded runs (tests/golden/ref_runs.json,
  written by tests/golden/make_golden.py from oracle/_ref), the closed-form
  stencil spectra, and dense LAPACK (dsygv) on a variable-coefficient
  problem with no closed form.
"""
import ctypes as C
import json
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import gcge_oracle as O
from gcge_amd import abi, problems as P

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_runs.json")))


def ia(*v):
    return (C.c_int * len(v))(*v)


# ---------------------------------------------------------------- slots
def test_inner_prod_loops():
    rng = np.random.default_rng(0)
    X, Y = rng.standard_normal((50, 4)), rng.standard_normal((50, 4))
    G = np.array([[sum(X[r, i] * Y[r, j] for r in range(50)) for j in range(4)]
                  for i in range(4)])
    assert np.allclose(O.inner_prod("N", X, Y), G, atol=1e-13)
    S = O.inner_prod("S", X, Y)
    assert np.allclose(np.tril(S), np.tril(G), atol=1e-13)     # lower computed
    assert np.allclose(S, S.T)                                  # mirrored
    assert not np.allclose(S, G)                                # G itself is not symmetric
    assert np.allclose(O.inner_prod("D", X, Y), np.diag(G), atol=1e-13)


def test_linear_comb_and_axpby_loops():
    rng = np.random.default_rng(1)
    X, Y0, Cf = rng.standard_normal((30, 3)), rng.standard_normal((30, 2)), \
        rng.standard_normal((3, 2))
    b = np.array([0.5, -2.0])
    Y = Y0.copy()
    O.linear_comb(X, Y, Cf, b)
    ref = np.empty_like(Y0)
    for k in range(2):
        for r in range(30):
            ref[r, k] = sum(X[r, i] * Cf[i, k] for i in range(3)) + b[k] * Y0[r, k]
    assert np.allclose(Y, ref, atol=1e-13)
    Y = Y0.copy()
    O.linear_comb(None, Y, None, None)      # no-op, as the reference (gamma only feeds dgemm)
    assert np.array_equal(Y, Y0)
    Y = Y0.copy()
    Y[0, 0] = np.nan
    O.axpby(2.0, X[:, :2], 0.0, Y)          # beta = 0 overwrites
    assert np.array_equal(Y, 2.0 * X[:, :2])


def test_spmm_is_ccs_product():
    A, _ = P.stencil_pair(9, 2)
    X = np.random.default_rng(2).standard_normal((81, 3))
    ref = np.zeros_like(X)
    for j in range(81):              # column scatter of app/app_ccs.c:126-129
        for e in range(A.indptr[j], A.indptr[j + 1]):
            ref[A.indices[e]] += A.data[e] * X[j]
    assert np.allclose(O.spmm(A, X), ref, atol=1e-13)


def test_qtap_side_effect():
    A, B = P.stencil_pair(6, 2)
    rng = np.random.default_rng(3)
    Q, Pm = rng.standard_normal((36, 5)), rng.standard_normal((36, 3))
    ws = np.zeros((36, 4))
    G = O.qtap("S", "N", Q, B, Pm, ws)
    assert np.allclose(ws[:, :3], B @ Pm) and np.all(ws[:, 3] == 0)
    assert np.allclose(G, Q.T @ (B @ Pm))
    assert np.allclose(O.qtap("S", "T", Q, B, Pm, None), (Q.T @ (B @ Pm)).T)


class RefDense:
    """The compiled reference's dense LAPACK table, called through ctypes."""

    def __init__(self, reflib):
        self.r = reflib
        self.ops = C.cast(C.c_void_p(reflib.lapack_table()), abi.OPS_p)
        self.keep = []

    def mv(self, M):
        h, keep = self.r.wrap_dense(M)
        self.keep.append(keep)
        return C.c_void_p(h), keep

    def slot(self, name):
        return abi.slot(self.ops, name)


def test_slots_vs_compiled_reference(reflib):
    """The oracle's slot semantics equal the reference's own slots (dense
    LAPACK app, app/app_lapack.c) on the same inputs."""
    R = RefDense(reflib)
    rng = np.random.default_rng(4)
    n = 40
    X = np.asfortranarray(rng.standard_normal((n, 9)))
    Y = np.asfortranarray(rng.standard_normal((n, 9)))
    hx, kx = R.mv(X)
    hy, ky = R.mv(Y)
    # LinearComb with per-column beta and ldc > rows (app/app_lapack.c:463-534)
    coef = np.asfortranarray(rng.standard_normal((6, 3)))
    beta = rng.standard_normal(6)
    ref_y = ky.copy()
    R.slot("MultiVecLinearComb")(hx, hy, 0, ia(1, 4), ia(5, 7), coef.ctypes.data_as(abi.c_dbl_p),
                                 6, beta.ctypes.data_as(abi.c_dbl_p), 2, R.ops)
    O.linear_comb(X[:, 1:5], ref_y[:, 4:7], coef[:4], beta[::2][:3])
    assert np.allclose(ky, ref_y, atol=1e-13)
    # InnerProd 'S' and 'D' with ld > rows
    for nsd, ld in (("S", 7), ("D", 3), ("N", 6)):
        ip = np.zeros(64)
        R.slot("MultiVecInnerProd")(nsd.encode(), hx, hx, 0, ia(2, 2 if nsd != "N" else 5),
                                    ia(6, 6 if nsd != "N" else 8), ip.ctypes.data_as(abi.c_dbl_p),
                                    ld, R.ops)
        if nsd == "D":
            got = ip[0:ld * 4:ld]
            want = O.inner_prod("D", X[:, 2:6], X[:, 2:6])
        elif nsd == "S":
            got = ip[:ld * 4].reshape((ld, 4), order="F")[:4]
            want = O.inner_prod("S", X[:, 2:6], X[:, 2:6])
        else:
            got = ip[:ld * 3].reshape((ld, 3), order="F")[:4]
            want = O.inner_prod("N", X[:, 2:6], X[:, 5:8])
        assert np.allclose(got, want, atol=1e-12), nsd
    # Axpby with beta 0 overwriting
    Z = np.asfortranarray(np.full((n, 2), np.nan))
    hz, kz = R.mv(Z)
    R.slot("MultiVecAxpby")(1.5, hx, 0.0, hz, ia(0, 0), ia(2, 2), R.ops)
    assert np.allclose(kz, 1.5 * X[:, :2])


def test_libc_random_matches_reference(reflib):
    R = RefDense(reflib)
    M = np.asfortranarray(np.zeros((25, 3)))
    hm, km = R.mv(M)
    reflib.lib.ref_srand(5)
    R.slot("MultiVecSetRandomValue")(hm, 0, 3, R.ops)
    O.libc_srand(5)
    assert np.array_equal(km, O.libc_uniform(25, 3))


# ------------------------------------------------------------- BlockPCG
def test_block_pcg_exact_in_n_steps():
    """CG on an SPD system reaches the solution in at most N steps."""
    A, _ = P.stencil_pair(5, 2)   # N = 25
    n = 25
    rng = np.random.default_rng(5)
    b = rng.standard_normal((n, 3))
    x = np.zeros((n, 3))
    ws = [np.zeros((n, 3)) for _ in range(3)]

    def op(src, dst, z):
        dst[...] = A @ src

    it = O.block_pcg(op, b.copy(), x, 200, 1e-14, 1e-30, *ws)
    assert it <= n + 5
    assert np.allclose(A @ x, b, atol=1e-10)


def test_block_pcg_rate_and_column_stop():
    A, _ = P.stencil_pair(20, 2)
    n = 400
    rng = np.random.default_rng(6)
    b = rng.standard_normal((n, 4))
    b[:, 2] = 0.0             # converged from the start: never iterated
    x = np.zeros((n, 4))
    ws = [np.zeros((n, 4)) for _ in range(3)]

    def op(src, dst, z):
        dst[...] = A @ src

    O.block_pcg(op, b.copy(), x, 500, 1e-3, 1e-14, *ws)
    r = np.linalg.norm(b - A @ x, axis=0)
    r0 = np.linalg.norm(b, axis=0)
    assert np.all(r[[0, 1, 3]] <= 1e-3 * r0[[0, 1, 3]])
    assert np.all(x[:, 2] == 0.0)


# ------------------------------------------------------------------ MGS
def test_mgs_b_orthonormal_and_span():
    A, B = P.stencil_pair(8, 2)
    n = 64
    rng = np.random.default_rng(7)
    V = rng.standard_normal((n, 10))
    V0 = V.copy()
    ws = np.zeros((n, 10))
    end = O.mgs_orth(V, 0, 4, B, -1, 2, 2 * O.DBL_EPSILON, 50 * O.DBL_EPSILON, ws)
    end = O.mgs_orth(V, 4, 10, B, -1, 2, 2 * O.DBL_EPSILON, 50 * O.DBL_EPSILON, ws)
    assert end == 10
    G = V.T @ (B @ V)
    assert np.allclose(G, np.eye(10), atol=1e-12)
    # span of the first k columns is preserved (Gram-Schmidt is triangular)
    for k in (4, 10):
        coef = np.linalg.lstsq(V0[:, :k], V[:, :k], rcond=None)[0]
        assert np.allclose(V0[:, :k] @ coef, V[:, :k], atol=1e-10)


def test_mgs_drops_dependent_column():
    A, B = P.stencil_pair(6, 2)
    rng = np.random.default_rng(8)
    V = rng.standard_normal((36, 5))
    V[:, 3] = V[:, 0] - 2 * V[:, 1]      # dependent on earlier columns
    end = O.mgs_orth(V, 0, 5, B, -1, 2, 1e-10, 50 * O.DBL_EPSILON, np.zeros((36, 5)))
    assert end == 4
    assert np.allclose(V[:, :4].T @ (B @ V[:, :4]), np.eye(4), atol=1e-12)


# --------------------------------------------------------------- driver
def golden(name, shift, app="ccs"):
    for r in GOLD["runs"]:
        if r["name"] == name and r["shift"] == shift and r["app"] == app:
            return r
    raise KeyError((name, shift, app))


@pytest.mark.parametrize("name,n,dim,nev,shift", [("2d_32", 32, 2, 10, 1),
                                                  ("2d_32", 32, 2, 10, 0),
                                                  ("2d_24", 24, 2, 20, 1),
                                                  ("3d_12", 12, 3, 30, 1)])
def test_driver_matches_reference_runs(name, n, dim, nev, shift):
    prob = P.stencil_problem(n, dim, nev)
    g = golden(name, shift)
    r = O.gcg_solve(prob.A, prob.B, nev, O.GCGParams(prob.nev_init, prob.nev_max,
                                                     prob.block_size, shift=shift))
    assert r.nev_conv == g["nev_conv"]
    assert abs(r.num_iter - g["num_iter"]) <= 3
    k = r.nev_conv
    ref = np.array(g["eval"][:k])
    assert np.max(np.abs(r.eval[:k] - ref) / ref) <= 1e-9
    exact = P.analytic_eigenvalues(n, dim, k)
    assert np.max(np.abs(r.eval[:k] - exact) / exact) <= 1e-10
    rr = O.relative_residuals(prob.A, prob.B, r.eval[:k], r.evec[:, :k])
    assert rr.max() <= 1.01e-8


def test_driver_vs_dense_lapack_variable_coefficients():
    """No closed form: lognormal-coefficient 5-point stencil (the batched
    workload's family) against dense dsygv."""
    n = 16
    N = n * n
    rng = np.random.default_rng(20261202)
    kappa = np.exp(0.5 * rng.standard_normal(N))
    rows, cols, vals = [], [], []
    idx = lambda i, j: i * n + j
    diag = np.zeros(N)
    for i in range(n):
        for j in range(n):
            for di, dj in ((1, 0), (0, 1)):
                ii, jj = i + di, j + dj
                if ii < n and jj < n:
                    a, b = idx(i, j), idx(ii, jj)
                    w = 2 * kappa[a] * kappa[b] / (kappa[a] + kappa[b])
                    rows += [a, b]
                    cols += [b, a]
                    vals += [-w, -w]
                    diag[a] += w
                    diag[b] += w
    diag += 1.0  # Dirichlet ghost terms
    A = sp.csc_matrix((vals + list(diag), (rows + list(range(N)), cols + list(range(N)))),
                      shape=(N, N))
    _, B = P.stencil_pair(n, 2)
    r = O.gcg_solve(A, B, 8, O.GCGParams(6, 14, 2, shift=1))
    assert r.nev_conv >= 8
    ex = O.dense_generalized_eigenvalues(A, B, r.nev_conv)
    assert np.max(np.abs(r.eval[:r.nev_conv] - ex) / ex) <= 1e-10


@pytest.mark.parametrize("name,n,kind,nev,shift", [("vc2d_40", 40, "vc2", 16, 1),
                                                   ("vc27_16", 16, "vc27", 20, 1)])
def test_driver_matches_reference_runs_variable_coefficients(name, n, kind, nev, shift):
    """The generators of the batched (5-point) and config-3 (27-point)
    families, against the compiled reference's runs and dense dsygv."""
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    from make_golden import problem
    prob = problem(name, n, kind, nev)
    g = golden(name, shift)
    r = O.gcg_solve(prob.A, prob.B, nev, O.GCGParams(prob.nev_init, prob.nev_max,
                                                     prob.block_size, shift=shift))
    # these spectra are harder than the stencils': numpy's and MKL's
    # rounding let the last iterations differ (one more pair may converge
    # in the last step), so the common converged pairs are compared
    assert r.nev_conv >= nev
    assert abs(r.num_iter - g["num_iter"]) <= 6
    k = min(r.nev_conv, g["nev_conv"])
    ref = np.array(g["eval"][:k])
    assert np.max(np.abs(r.eval[:k] - ref) / ref) <= 1e-9
    ex = O.dense_generalized_eigenvalues(prob.A, prob.B, k)
    assert np.max(np.abs(ref - ex) / ex) <= 1e-10
    rr = O.relative_residuals(prob.A, prob.B, r.eval[:k], r.evec[:, :k])
    assert rr.max() <= 1.01e-8


def test_analytic_spot_values():
    """The closed forms reproduce the quoted anchor values."""
    l2 = P.analytic_eigenvalues(128, 2, 21)
    assert abs(l2[0] - 1.18624e-3) < 1e-8
    assert abs(l2[19] - 1.899388087554585e-2) < 1e-15
    assert abs(l2[20] - 2.017823482722744e-2) < 1e-15
    l3 = P.analytic_eigenvalues(128, 3, 102)
    assert abs(l3[0] - 1.779356806813896e-3) < 1e-15
    assert np.allclose(l3[96:102], 2.491836969232519e-2, rtol=1e-14, atol=0)
    # dense check of the closed form on a small grid
    for n, dim in ((7, 2), (4, 3)):
        A, B = P.stencil_pair(n, dim)
        ex = O.dense_generalized_eigenvalues(A, B, 10)
        assert np.allclose(ex, P.analytic_eigenvalues(n, dim, 10), rtol=1e-12)


def _sine_mode(n, i, j):
    """The (i, j) sine mode of the n x n Dirichlet grid, point x + n y: an
    eigenvector of the 5-point Laplacian (4 - mu) and of the mass matrix
    ((8 + mu) / 12) alike, mu = 2 cos(i pi/(n+1)) + 2 cos(j pi/(n+1))."""
    t = np.arange(1, n + 1) * np.pi / (n + 1)
    v = np.outer(np.sin(j * t), np.sin(i * t)).ravel()  # row y, column x
    mu = 2 * np.cos(i * np.pi / (n + 1)) + 2 * np.cos(j * np.pi / (n + 1))
    return v, mu


def test_relative_residuals_known_values():
    """The residual checker against residuals fixed by the closed-form
    eigenvectors: an exact eigenpair has residual at rounding level, and
    x = v + eps u (u another mode) with the eigenvalue of v has
    A x - lam B x = eps (lam_u - lam) b_u u, so its relative residual is
    eps |lam_u - lam| b_u |u| / |lam| with b_u = (8 + mu_u) / 12.  A dropped
    B, a wrong sign or a missing division by |lam| each fails."""
    n = 24
    A, B = P.stencil_pair(n, 2)
    v, mv = _sine_mode(n, 1, 2)
    u, mu_ = _sine_mode(n, 3, 1)
    lam_v = (4 - mv) * 12 / (8 + mv)
    lam_u = (4 - mu_) * 12 / (8 + mu_)
    b_u = (8 + mu_) / 12
    eps = 1e-3
    X = np.stack([v, v + eps * u], axis=1)
    lam = np.array([lam_v, lam_v])
    rr = O.relative_residuals(A, B, lam, X)
    assert rr[0] < 1e-13
    expect = eps * abs(lam_u - lam_v) * b_u * np.linalg.norm(u) / lam_v
    assert abs(rr[1] - expect) <= 1e-10 * expect, (rr[1], expect)

"""This library's GCG driver, run on the host over its own dense table (no
device use), against the reference driver's recorded runs over the
reference dense app (tests/golden/ref_runs.json, app "lapack"), and the
host small-dense operations (DenseMatQtAP, DenseMatOrth) against numpy.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import gcge_oracle as O
from gcge_amd import abi, problems as P
from gcge_amd import solver as S

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_runs.json")))


def golden(name, shift, app):
    for r in GOLD["runs"]:
        if r["name"] == name and r["shift"] == shift and r["app"] == app:
            return r
    raise KeyError((name, shift, app))


def b_orth_defect(B, X):
    """max |X^T B X - I|: the eigenvectors leave the driver B-orthonormal
    (the B inner product of src/ops_orth.c, kept through Rayleigh-Ritz)."""
    G = X.T @ (B @ X)
    return np.abs(G - np.eye(G.shape[0])).max()


@pytest.mark.parametrize("name,n,dim,nev,shift", [("2d_32", 32, 2, 10, 1),
                                                  ("2d_32", 32, 2, 10, 0),
                                                  ("2d_24", 24, 2, 20, 1)])
def test_host_driver_matches_reference(hip, name, n, dim, nev, shift):
    prob = P.stencil_problem(n, dim, nev)
    g = golden(name, shift, "lapack")
    r = S.solve_host_dense(prob.A.toarray(), prob.B.toarray(), nev, prob.block_size,
                           prob.nev_init, prob.nev_max, shift=shift)
    assert r.nev_conv == g["nev_conv"]
    assert abs(r.num_iter - g["num_iter"]) <= 3
    k = r.nev_conv
    ref = np.array(g["eval"][:k])
    assert np.max(np.abs(r.eval[:k] - ref) / ref) <= 1e-9
    rr = O.relative_residuals(prob.A, prob.B, r.eval[:k], r.evec[:, :k])
    assert rr.max() <= 1.01e-8
    assert b_orth_defect(prob.B, r.evec[:, :r.nev_conv]) <= 1e-12


@pytest.mark.parametrize("name,n,dim,nev,shift", [("2d_32", 32, 2, 10, 1),
                                                  ("2d_24", 24, 2, 20, 1)])
def test_host_driver_cholqr(hip, name, n, dim, nev, shift):
    """Block Cholesky QR in place of the column MGS: the same pairs as the
    reference (a different rounding path, so the iteration count may move
    by a few)."""
    prob = P.stencil_problem(n, dim, nev)
    g = golden(name, shift, "lapack")
    r = S.solve_host_dense(prob.A.toarray(), prob.B.toarray(), nev, prob.block_size,
                           prob.nev_init, prob.nev_max, shift=shift, orth="cholqr")
    assert r.nev_conv >= nev
    # every W block went through CholQR2 (unit-diagonal scaling keeps the
    # tiny corrections of nearly converged pairs from looking dependent)
    assert r.cholqr_blocks >= r.num_iter and r.cholqr_fallbacks == 0
    assert abs(r.num_iter - g["num_iter"]) <= 5
    k = min(r.nev_conv, g["nev_conv"])
    ref = np.array(g["eval"][:k])
    assert np.max(np.abs(r.eval[:k] - ref) / ref) <= 1e-9
    rr = O.relative_residuals(prob.A, prob.B, r.eval[:k], r.evec[:, :k])
    assert rr.max() <= 1.01e-8
    assert b_orth_defect(prob.B, r.evec[:, :r.nev_conv]) <= 1e-12


def test_host_driver_max_iter_reports_nonconvergence(hip):
    """numIterMax = 1: a reported non-convergence, not an error
    (gcge.c:525-531 maps it to EPS_DIVERGED_ITS)."""
    prob = P.stencil_problem(16, 2, 10)
    r = S.solve_host_dense(prob.A.toarray(), prob.B.toarray(), 10, prob.block_size,
                           prob.nev_init, prob.nev_max, shift=1, max_iter=1)
    assert r.num_iter == 1 and r.nev_conv < 10


@pytest.mark.parametrize("size", [(12, 5), (460, 200)])
@pytest.mark.parametrize("ntlu", ["L", "U", "N", "T"])
@pytest.mark.parametrize("nsd", ["N", "S", "D"])
def test_dense_qtap(hip, ntlu, nsd, size):
    """Small sizes run serially; 460 x 200 passes the 8 M multiply-add
    threshold in every product (op(A) P, Q^T W and the symmetric result),
    which then split their output columns over threads."""
    rng = np.random.default_rng(ord(ntlu) + ord(nsd))
    n, m = size
    A = rng.standard_normal((n, n))
    if ntlu in "LU":
        A = A + A.T
        Ad = np.tril(A) if ntlu == "L" else np.triu(A)  # only that triangle is read
        Ad = Ad + np.triu(rng.standard_normal((n, n)), 1) if ntlu == "L" else \
            Ad + np.tril(rng.standard_normal((n, n)), -1)
        opA = A
    else:
        Ad = A
        opA = A if ntlu == "N" else A.T
    Q = rng.standard_normal((n, m))
    Pm = Q + 0.1 * rng.standard_normal((n, m)) if nsd == "S" else rng.standard_normal((n, m))
    ld = m + 2
    Cm = np.asfortranarray(rng.standard_normal((ld, m)))
    C0 = Cm.copy()
    Af, Qf, Pf = (np.asfortranarray(x) for x in (Ad, Q, Pm))
    ws = np.zeros(n * m)
    hip.gcge_host_DenseMatQtAP(ntlu.encode(), nsd.encode(), n, n, m, m, 0.5,
                               Qf.ctypes.data_as(abi.c_dbl_p), n,
                               Af.ctypes.data_as(abi.c_dbl_p), n,
                               Pf.ctypes.data_as(abi.c_dbl_p), n, 2.0,
                               Cm.ctypes.data_as(abi.c_dbl_p), ld,
                               ws.ctypes.data_as(abi.c_dbl_p))
    G = Q.T @ opA @ Pm
    # FP64 rounding of sums of n^2 products, relative to their size
    tol = 1e-12 * max(1.0, float(np.abs(G).max()))
    if nsd == "N":
        want = 0.5 * G + 2.0 * C0[:m]
        assert np.allclose(Cm[:m], want, rtol=0, atol=tol)
    elif nsd == "S":
        low = np.tril(0.5 * G + 2.0 * C0[:m])
        want = low + np.tril(low, -1).T
        assert np.allclose(Cm[:m], want, rtol=0, atol=tol)
    else:
        # 'D' writes dot c at matC[ldC*c] (app/app_lapack.c:75-113)
        flat, flat0 = Cm.ravel(order="F"), C0.ravel(order="F")
        want = 0.5 * np.diag(G) + 2.0 * flat0[0:ld * m:ld]
        assert np.allclose(flat[0:ld * m:ld], want, rtol=0, atol=tol)


def test_dense_orth(hip):
    rng = np.random.default_rng(11)
    m = 20
    M = np.asfortranarray(rng.standard_normal((m, 9)))
    Q0, _ = np.linalg.qr(M[:, :3])
    M[:, :3] = Q0
    M[:, 6] = M[:, 4] + M[:, 1]           # dependent on earlier columns
    orig = M.copy()
    end = C.c_int(9)
    hip.gcge_host_DenseMatOrth(M.ctypes.data_as(abi.c_dbl_p), m, m, 3, C.byref(end), 1e-10,
                               None, 0, None)
    assert end.value == 8
    Q = M[:, :end.value]
    assert np.allclose(Q.T @ Q, np.eye(end.value), atol=1e-12)
    # the new columns span the projected originals
    proj = orig[:, 3:9] - Q0 @ (Q0.T @ orig[:, 3:9])
    coef = np.linalg.lstsq(Q[:, 3:], proj, rcond=None)[0]
    assert np.allclose(Q[:, 3:] @ coef, proj, atol=1e-10)


@pytest.mark.parametrize("n", [1, 2, 5, 40, 130])
def test_host_dense_eigh(hip, n):
    """The host projected eigensolver against numpy, on graded spectra and
    with repeated and clustered eigenvalues (the Ritz values of converged
    pairs)."""
    rng = np.random.default_rng(n)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.sort(np.concatenate([np.full(n // 3, 0.25), 0.25 + 1e-13 * np.arange(n // 5),
                                  rng.uniform(-3, 40, n - n // 3 - n // 5)]))
    M = (Q * lam) @ Q.T
    M = 0.5 * (M + M.T)
    A = np.asfortranarray(np.triu(M) + np.tril(rng.standard_normal((n, n)), -1))  # upper read
    w = np.zeros(n)
    Z = np.zeros((n, n), order="F")
    info = hip.gcge_host_dense_eigh(n, A.ctypes.data_as(abi.c_dbl_p), n,
                                    w.ctypes.data_as(abi.c_dbl_p),
                                    Z.ctypes.data_as(abi.c_dbl_p), n)
    assert info == 0
    scale = np.abs(lam).max()
    assert np.all(np.diff(w) >= 0)
    assert np.max(np.abs(w - np.linalg.eigvalsh(M))) <= 1e-13 * scale
    assert np.max(np.abs(Z.T @ Z - np.eye(n))) <= 1e-13
    assert np.max(np.abs(M @ Z - Z * w)) <= 1e-13 * scale

"""The block Cholesky QR pass with its small dense step on the device
(HIP_MultiVecCholQR; the GCGE_ORTH_CHOLQR orthogonalization of the
driver).

The coefficients are checked against numpy's Cholesky factorization of the
same Gram matrix (X <- X D^{-1} R^{-1}, D^{-1} X^T B X D^{-1} = R^T R), and
the result against B-orthonormality.  Widths cover the LDS factorization
(up to 128 columns) and the global-memory one (129 to 1024), and both
layouts of the coefficients: transposed for the matrix-core linear
combination, column-major for the library GEMM (from 448 Ki rows and 100
columns), and both forms of the factorization kernel (LDS up to 128
columns, global memory above).  A numerically dependent block must leave X
bit for bit as it
was, so the caller's fallback to Gram-Schmidt starts from the same state.
"""

import numpy as np
import pytest
import scipy.sparse as sp

from gcge_amd import problems as P
from gcge_amd import solver as S
from gcge_amd.solver import DeviceMatrix, HipTable

pytestmark = pytest.mark.gpu

ZERO_TOL = 1e-14
REL_TOL = 1e-10


@pytest.fixture(scope="module")
def table(hip):
    return HipTable()


def b_matrix(n, seed):
    """SPD B: a diagonal with entries in [0.5, 1.5] (config 3's B)."""
    rng = np.random.default_rng(seed)
    return sp.diags(rng.uniform(0.5, 1.5, n)).tocsc()


def reference(X, B):
    BX = X if B is None else B @ X
    G = X.T @ BX
    d = np.sqrt(np.diag(G))
    R = np.linalg.cholesky((G / d[:, None]) / d[None, :]).T
    return X @ (np.linalg.inv(R) / d[:, None])


def run_pass(table, hip, X, B, dB, start, extra=2):
    """Uploads X into columns [start, start + n) of a block with `extra`
    more columns on each side, runs the pass, returns (rc, block before,
    block after)."""
    N, n = X.shape
    rng = np.random.default_rng(n)
    full = rng.standard_normal((N, start + n + extra))
    full[:, start:start + n] = X
    mat = dB if dB is not None else DeviceMatrix(sp.identity(N, format="csc"))
    hx = table.create_multivec(full.shape[1], mat)
    hw = table.create_multivec(n, mat)
    table.upload(hx, full)
    rc = hip.HIP_MultiVecCholQR(hx, start, start + n, dB.handle if dB is not None else None,
                                hw, ZERO_TOL, REL_TOL)
    out = table.download(hx, 0, full.shape[1], N)
    table.destroy_multivec(hx, full.shape[1])
    table.destroy_multivec(hw, n)
    return rc, full, out


@pytest.mark.parametrize("N,n,withB", [
    (1000, 1, True), (5000, 7, True), (20000, 40, False), (20000, 64, True),
    (20000, 65, False), (20000, 120, True),
    (20000, 128, True), (20000, 129, True), (30000, 300, False),
    (460000, 100, True), (460000, 130, True)])
def test_cholqr_pass_matches_numpy(table, hip, N, n, withB):
    rng = np.random.default_rng(N + n)
    # columns of very different norms (the diagonal scaling's case)
    X = rng.standard_normal((N, n)) * np.logspace(-3, 3, n)
    B = b_matrix(N, n) if withB else None
    dB = DeviceMatrix(B) if withB else None
    rc, before, after = run_pass(table, hip, X, B, dB, start=3)
    assert rc == 1
    Q = after[:, 3:3 + n]
    ref = reference(X, B)
    assert np.abs(Q - ref).max() <= 1e-10 * np.abs(ref).max(), np.abs(Q - ref).max()
    G = Q.T @ (Q if B is None else B @ Q)
    assert np.abs(G - np.eye(n)).max() <= 1e-12, np.abs(G - np.eye(n)).max()
    # the other columns are untouched
    assert np.array_equal(after[:, :3], before[:, :3])
    assert np.array_equal(after[:, 3 + n:], before[:, 3 + n:])


@pytest.mark.parametrize("case", ["duplicate", "zero", "dependent_wide"])
def test_cholqr_pass_dependent_block_unchanged(table, hip, case):
    N = 4000
    n = 200 if case == "dependent_wide" else 12
    rng = np.random.default_rng(5)
    X = rng.standard_normal((N, n))
    if case == "zero":
        X[:, 5] = 0.0
    else:
        X[:, n - 2] = X[:, 1] * 3.0 + X[:, 0] * 1e-13
    B = b_matrix(N, 9)
    rc, before, after = run_pass(table, hip, X, B, DeviceMatrix(B), start=1)
    assert rc == 0
    assert np.array_equal(after, before)


def test_cholqr_pass_empty(table, hip):
    N = 100
    mat = DeviceMatrix(sp.identity(N, format="csc"))
    hx = table.create_multivec(4, mat)
    hw = table.create_multivec(4, mat)
    assert hip.HIP_MultiVecCholQR(hx, 2, 2, None, hw, ZERO_TOL, REL_TOL) == -1
    table.destroy_multivec(hx, 4)
    table.destroy_multivec(hw, 4)


@pytest.mark.parametrize("mode,n,nev", [("mixed", 16, 40), ("fp64", 12, 30)])
def test_solve_device_and_host_factorization_agree(hip, mode, n, nev):
    """The same solve with the CholQR factorization on the device and on
    the host: eigenvalues within 1e-10 of each other and of the analytic
    spectrum, B-orthonormal eigenvectors."""
    prob = P.stencil_problem(n, 3, nev)
    shift = 0 if mode == "mixed" else 1
    res = {o: S.solve_problem(prob, shift=shift, mode=mode, want_vectors=True, orth=o)
           for o in ("cholqr", "cholqr_host")}
    exact = P.analytic_eigenvalues(n, 3, prob.nev)
    for o, r in res.items():
        assert r.nev_conv >= prob.nev and r.cholqr_blocks >= r.num_iter, o
        lam = r.eval[:prob.nev]
        assert (np.abs(lam - exact) / exact).max() <= 1e-10, o
        X = r.evec[:, :r.nev_conv]
        G = X.T @ (prob.B @ X)
        assert np.abs(G - np.eye(r.nev_conv)).max() <= 1e-10, o
    a, b = res["cholqr"].eval[:prob.nev], res["cholqr_host"].eval[:prob.nev]
    assert (np.abs(a - b) / np.abs(b)).max() <= 1e-10


def test_solve_device_compute_p(hip):
    """A 32^3 solve with config 2's block sizes (nev 100, block 120, 220
    initial vectors), whose P coefficient blocks (order ~460 by 120) are
    orthonormalized on the device by Cholesky QR: eigenvalues within 1e-10
    of the analytic ones and of the run that orthonormalizes them on the
    host by MGS, B-orthonormal eigenvectors."""
    prob = P.stencil_problem(32, 3, 100, 120, 220, 220, name="3d_32")
    res = {o: S.solve_problem(prob, shift=0, mode="mixed", want_vectors=True, orth=o)
           for o in ("cholqr", "cholqr_host")}
    exact = P.analytic_eigenvalues(32, 3, prob.nev)
    for o, r in res.items():
        assert r.nev_conv >= prob.nev, o
        lam = r.eval[:prob.nev]
        assert (np.abs(lam - exact) / exact).max() <= 1e-10, o
        X = r.evec[:, :r.nev_conv]
        G = X.T @ (prob.B @ X)
        assert np.abs(G - np.eye(r.nev_conv)).max() <= 1e-10, o
    assert res["cholqr"].phases["compP"] < res["cholqr_host"].phases["compP"]
    a, b = res["cholqr"].eval[:prob.nev], res["cholqr_host"].eval[:prob.nev]
    assert (np.abs(a - b) / np.abs(b)).max() <= 1e-10

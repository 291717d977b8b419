"""Matrix-free products with constant-coefficient 7-point operators
(stencil.hip): the 3D Laplacian and mass matrices are recognised, uploaded
or assembled on the device, and MatDotMultiVec gives the general SELL
kernel's results bit for bit (the same fused multiply-adds in the SELL
row's order) and the oracle's product (app/app_ccs.c:50-139 semantics)
within FP64 rounding; a matrix that differs in one value is not
recognised."""
import ctypes as C

import numpy as np
import pytest

import gcge_oracle as O
from gcge_amd import problems as P
from gcge_amd.solver import DeviceMatrix, HipTable

pytestmark = pytest.mark.gpu


def ia(*v):
    return (C.c_int * len(v))(*v)


def product(table, dM, X, s0, s1, k):
    n = X.shape[0]
    hx = table.create_multivec(X.shape[1], dM)
    hy = table.create_multivec(s1 + k + 1, dM)
    table.upload(hx, X)
    table.slot("MatDotMultiVec")(dM.handle, hx, hy, ia(s0, s1), ia(s0 + k, s1 + k), table.ops)
    Y = table.download(hy, 0, s1 + k + 1, n)
    table.destroy_multivec(hx, X.shape[1])
    table.destroy_multivec(hy, s1 + k + 1)
    return Y


@pytest.mark.parametrize("n,k", [(4, 5), (5, 3), (17, 9), (48, 4), (64, 2), (97, 2), (128, 2)])
def test_stencil_matches_sell_bit_for_bit(hip, n, k):
    table = HipTable()
    A, B = P.stencil_pair(n, 3)
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n ** 3, k + 3))
    for M, kind in ((A, 0), (B, 1)):
        dS = DeviceMatrix(M)                       # recognised on upload
        dG = DeviceMatrix(M, general=True)         # general SELL kernels
        dD = DeviceMatrix.grid(kind, 3, 7, n)      # assembled on the device
        assert dS.stencil() == n and dD.stencil() == n and dG.stencil() == 0
        # an odd source offset moves the columns off 16-byte alignment when
        # n is odd
        ys = product(table, dS, X, 1, 2, k)
        yg = product(table, dG, X, 1, 2, k)
        yd = product(table, dD, X, 1, 2, k)
        assert np.array_equal(ys, yg) and np.array_equal(yd, yg)
        ref = O.spmm(M, X[:, 1:1 + k])
        got = ys[:, 2:2 + k]
        assert np.max(np.abs(got - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))
        # columns outside the destination range are untouched (zero)
        assert not ys[:, :2].any() and not ys[:, 2 + k:].any()


def test_not_a_stencil_is_not_recognised(hip):
    n = 12
    A, _ = P.stencil_pair(n, 3)
    A = A.tocsc(copy=True)
    A.data[37] = np.nextafter(A.data[37], 0.0)     # one value one ulp off
    assert DeviceMatrix(A).stencil() == 0
    A2, _ = P.stencil_pair(n, 2)                   # 2D: not a 7-point operator
    assert DeviceMatrix(A2).stencil() == 0


def pcg_solve(table, lib, A, dA, b, rate, max_iter, fused=True):
    n, nc = b.shape
    hb = table.create_multivec(nc, dA)
    hx = table.create_multivec(nc, dA)
    table.upload(hb, b)
    lib.HIPGCG_MultiLinearSolverSetup_BlockPCG32(max_iter, rate, 0.0, None, 0.0, table.ops)
    lib.HIPGCG_BlockPCG32_SetFused(table.ops, 1 if fused else 0)
    table.slot("MultiLinearSolver")(dA.handle, hb, hx, ia(0, 0), ia(nc, nc), table.ops)
    x = table.download(hx, 0, nc, n)
    table.destroy_multivec(hb, nc)
    table.destroy_multivec(hx, nc)
    lib.HIPGCG_BlockPCG32_SetFused(table.ops, 1)
    return x


@pytest.mark.parametrize("n", [24, 48])
def test_blockpcg_with_matrix_free_product(hip, n):
    """The FP32 BlockPCG over a stencil operator (its products matrix-free)
    reaches FP32 accuracy against a direct FP64 solve, agrees with the same
    solve over the general SELL kernels to that accuracy, and its three-
    and five-launch iterations (n = 24: 13,824 rows) agree bit for bit."""
    import scipy.sparse.linalg as spla
    A, _ = P.stencil_pair(n, 3)
    table = HipTable()
    dS, dG = DeviceMatrix(A, fp32=True), DeviceMatrix(A, fp32=True, general=True)
    assert dS.stencil() == n
    b = np.random.default_rng(n).standard_normal((A.shape[0], 11))
    xs = pcg_solve(table, hip, A, dS, b, 1e-6, 3000)
    xg = pcg_solve(table, hip, A, dG, b, 1e-6, 3000)
    xe = spla.spsolve(A.tocsc(), b).reshape(b.shape)
    for x in (xs, xg):
        r = np.linalg.norm(b - A @ x, axis=0) / np.linalg.norm(b, axis=0)
        assert r.max() <= 1e-4, r.max()
        assert np.max(np.abs(x - xe)) <= 1e-3 * np.max(np.abs(xe))
    if n == 24:
        x5 = pcg_solve(table, hip, A, dS, b, 1e-6, 3000, fused=False)
        assert np.array_equal(x5, xs)


@pytest.mark.parametrize("orth", ["mgs", "cholqr"])
def test_solve_matrix_free_equals_general(hip, orth):
    """A whole FP64 GCG solve over recognised stencil operators -- products
    matrix-free and the convergence check's residuals A x - lambda B x in one
    fused pass -- runs the same trajectory as over the general SELL kernels
    with the reference's separate product, scaling and axpby steps
    (src/ops_eig_sol_gcg.c:210-222): the same iteration count and the same
    eigenvalues bit for bit."""
    from gcge_amd import solver as S
    n, nev = 20, 10          # 10: the end of a cluster of the cube's spectrum
    A, B = P.stencil_pair(n, 3)
    bs, ni, nm = P.slepc_block_policy(nev)
    ms = (DeviceMatrix(A), DeviceMatrix(B))
    mg = (DeviceMatrix(A, general=True), DeviceMatrix(B, general=True))
    assert ms[0].stencil() == n and ms[1].stencil() == n and mg[0].stencil() == 0
    rs = S.solve(A, B, nev, bs, ni, nm, mats=ms, orth=orth, seed=3)
    rg = S.solve(A, B, nev, bs, ni, nm, mats=mg, orth=orth, seed=3)
    print(f"# {orth}: {rs.num_iter} iterations, {rs.nev_conv} converged", flush=True)
    assert rs.nev_conv >= nev
    assert rs.num_iter == rg.num_iter and rs.nev_conv == rg.nev_conv
    assert np.array_equal(rs.eval, rg.eval)

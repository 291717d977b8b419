"""Known-answer tests of the single-precision device BlockPCG installed in
the MultiLinearSolver slot (HIPGCG_MultiLinearSolverSetup_BlockPCG32).

The stopping rule is the reference BlockPCG's (src/ops_lin_sol.c:383-394):
per column, residual below rate * initial residual or below tol, at most
max_iter iterations.  Residuals are measured in FP64 on the host.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from gcge_amd import abi, problems as P
from gcge_amd.solver import DeviceMatrix, HipTable

pytestmark = pytest.mark.gpu


def ia(*v):
    return (C.c_int * len(v))(*v)


def run(table, lib, A, dA, dB, sigma, X0, Bm, max_iter, rate, ncols, b_off=3, x_off=5):
    n = A.shape[0]
    rng = np.random.default_rng(ncols)
    b = rng.standard_normal((n, ncols + b_off))
    x = np.zeros((n, ncols + x_off)) if X0 is None else X0
    hb = table.create_multivec(b.shape[1], dA)
    hx = table.create_multivec(x.shape[1], dA)
    table.upload(hb, b)
    table.upload(hx, x)
    lib.HIPGCG_MultiLinearSolverSetup_BlockPCG32(max_iter, rate, 0.0,
                                                 dB.handle if dB else None, sigma, table.ops)
    abi.slot(table.ops, "MultiLinearSolver")(dA.handle, hb, hx, ia(b_off, x_off),
                                             ia(b_off + ncols, x_off + ncols), table.ops)
    xs = table.download(hx, 0, x.shape[1], n)
    op = A + sigma * Bm if Bm is not None else A
    bb = b[:, b_off:b_off + ncols]
    xx = xs[:, x_off:x_off + ncols]
    x0 = x[:, x_off:x_off + ncols]
    r0 = np.linalg.norm(bb - op @ x0, axis=0)
    r = np.linalg.norm(bb - op @ xx, axis=0)
    # columns outside the solved range are untouched
    assert np.array_equal(xs[:, :x_off], x[:, :x_off])
    return r / r0, xx


@pytest.fixture(scope="module")
def setup(hip):
    A, B = P.stencil_pair(40, 2)  # N = 1600
    return HipTable(), hip, A, B, DeviceMatrix(A, fp32=True), DeviceMatrix(B, fp32=True)


@pytest.mark.parametrize("ncols", [1, 5, 9, 33])
def test_converges_to_fp32_accuracy(setup, ncols):
    table, lib, A, B, dA, dB = setup
    rel, xx = run(table, lib, A, dA, None, 0.0, None, None, 2000, 1e-6, ncols)
    # FP32 recursion: the true residual trails the recursive one by about
    # cond(A) * eps32 (cond ~ 1e3 here)
    assert rel.max() <= 1e-4, rel.max()
    # agrees with a direct FP64 solve to the conditioning of the problem
    rng = np.random.default_rng(ncols)
    b = rng.standard_normal((A.shape[0], ncols + 3))[:, 3:3 + ncols]
    xe = spla.spsolve(A.tocsc(), b)
    xe = xe.reshape(b.shape)
    assert np.max(np.abs(xx - xe)) <= 1e-3 * np.max(np.abs(xe))


def test_rate_stop(setup):
    """With rate 1e-2 every column stops once its residual has dropped 100x
    (and not far beyond: it stops at the first iteration below)."""
    table, lib, A, B, dA, dB = setup
    rel, _ = run(table, lib, A, dA, None, 0.0, None, None, 500, 1e-2, 12)
    assert rel.max() <= 1e-2
    assert rel.min() > 1e-5


def test_max_iter_stop(setup):
    table, lib, A, B, dA, dB = setup
    rel1, _ = run(table, lib, A, dA, None, 0.0, None, None, 1, 1e-12, 4)
    rel5, _ = run(table, lib, A, dA, None, 0.0, None, None, 5, 1e-12, 4)
    assert np.all(rel1 < 1.0) and np.all(rel5 < rel1)


def test_shifted_operator(setup):
    """(A + sigma B) x = b with B's FP32 copy; nonzero initial guess."""
    table, lib, A, B, dA, dB = setup
    n = A.shape[0]
    X0 = np.random.default_rng(9).standard_normal((n, 11))
    rel, _ = run(table, lib, A, dA, dB, 0.37, X0, B, 2000, 1e-6, 6)
    # FP32 recursion: the true residual trails the recursive one by about
    # cond(A) * eps32 (cond ~ 1e3 here)
    assert rel.max() <= 1e-4, rel.max()


@pytest.mark.parametrize("rate", [0.0, 0.2])
def test_graph_replays_match_eager(setup, rate):
    """Inner iterations replayed from a captured graph give the eager
    launches' result bit for bit, call after call, with eager solves,
    uploads and synchronizations in between (rate 0.2: tiles stop part way,
    so the graph is captured again for the smaller active count)."""
    table, lib, A, B, dA, dB = setup
    n, nc = A.shape[0], 12
    rng = np.random.default_rng(42)
    b = rng.standard_normal((n, nc))
    hb = table.create_multivec(nc, dA)
    hx = table.create_multivec(nc, dA)
    table.upload(hb, b)
    lib.HIPGCG_MultiLinearSolverSetup_BlockPCG32(30, rate, 0.0, None, 0.0, table.ops)
    outs = []
    for graphs in (0, 1, 1, 0, 1, 1):
        lib.HIPGCG_BlockPCG32_SetGraphs(table.ops, graphs)
        table.upload(hx, np.zeros((n, nc)))
        abi.slot(table.ops, "MultiLinearSolver")(dA.handle, hb, hx, ia(0, 0), ia(nc, nc),
                                                 table.ops)
        outs.append(table.download(hx, 0, nc, n))
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    r = np.linalg.norm(b - A @ outs[0], axis=0) / np.linalg.norm(b, axis=0)
    assert r.max() < (1.0 if rate == 0.0 else rate * 1.01)
    table.destroy_multivec(hb, nc)
    table.destroy_multivec(hx, nc)


@pytest.mark.parametrize("shifted", [False, True])
def test_row_block_orders_give_identical_bits(setup, tmp_path, shifted):
    """Plain and XCD-contiguous row-block orders of the gathering kernels
    (tuner op 3) give the same bits: partial sums are stored by row block."""
    table, lib, A, B, dA, dB = setup
    n, nc = A.shape[0], 12
    keep = str(tmp_path / "kept.txt")
    lib.gcge_tune_save(keep.encode())
    rng = np.random.default_rng(43)
    b = rng.standard_normal((n, nc))
    hb = table.create_multivec(nc, dA)
    hx = table.create_multivec(nc, dA)
    table.upload(hb, b)
    sigma = 0.5 if shifted else 0.0
    # enough iterations for every column to reach the rate criterion
    lib.HIPGCG_MultiLinearSolverSetup_BlockPCG32(400, 1e-2, 0.0,
                                                 dB.handle if shifted else None, sigma,
                                                 table.ops)
    lg = int(np.floor(np.log2(n)))
    outs = []
    try:
        for ch in (0, 1, 0, 1):
            lib.gcge_tune_set(3, 8 + (100 if shifted else 0), lg, 0, ch)
            table.upload(hx, np.zeros((n, nc)))
            abi.slot(table.ops, "MultiLinearSolver")(dA.handle, hb, hx, ia(0, 0), ia(nc, nc),
                                                     table.ops)
            outs.append(table.download(hx, 0, nc, n))
    finally:
        lib.gcge_tune_clear()
        lib.gcge_tune_load(keep.encode())
        table.destroy_multivec(hb, nc)
        table.destroy_multivec(hx, nc)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    op = A + sigma * B
    r = np.linalg.norm(b - op @ outs[0], axis=0) / np.linalg.norm(b, axis=0)
    assert r.max() < 1.01e-2


def test_graph_follows_new_parameters(setup):
    """A graph captured under one stopping rule is not replayed under
    another: after a new setup, graph and eager runs agree and the new
    rate holds."""
    table, lib, A, B, dA, dB = setup
    n, nc = A.shape[0], 12
    b = np.random.default_rng(44).standard_normal((n, nc))
    hb = table.create_multivec(nc, dA)
    hx = table.create_multivec(nc, dA)
    table.upload(hb, b)
    outs = {}
    try:
        for rate, graphs in ((0.0, 1), (1e-2, 1), (1e-2, 0)):
            lib.HIPGCG_MultiLinearSolverSetup_BlockPCG32(300 if rate else 40, rate, 0.0, None,
                                                         0.0, table.ops)
            lib.HIPGCG_BlockPCG32_SetGraphs(table.ops, graphs)
            table.upload(hx, np.zeros((n, nc)))
            abi.slot(table.ops, "MultiLinearSolver")(dA.handle, hb, hx, ia(0, 0), ia(nc, nc),
                                                     table.ops)
            outs[(rate, graphs)] = table.download(hx, 0, nc, n)
    finally:
        lib.HIPGCG_BlockPCG32_SetGraphs(table.ops, 1)
        table.destroy_multivec(hb, nc)
        table.destroy_multivec(hx, nc)
    assert np.array_equal(outs[(1e-2, 1)], outs[(1e-2, 0)])
    r = np.linalg.norm(b - A @ outs[(1e-2, 1)], axis=0) / np.linalg.norm(b, axis=0)
    assert r.max() < 1.01e-2


@pytest.mark.parametrize("graphs", [0, 1])
@pytest.mark.parametrize("rate", [0.0, 0.05])
def test_three_launch_iteration_matches_five(setup, graphs, rate):
    """Up to 65,536 rows the reductions run inside the following kernels
    (state double-buffered by parity); the result is the five-launch
    iteration's bit for bit, eager or from graphs, with tiles stopping part
    way (rate 0.05) or never (rate 0)."""
    table, lib, A, B, dA, dB = setup
    n, nc = A.shape[0], 20
    b = np.random.default_rng(46).standard_normal((n, nc))
    hb = table.create_multivec(nc, dA)
    hx = table.create_multivec(nc, dA)
    table.upload(hb, b)
    outs = []
    try:
        for fused in (1, 0, 1):
            lib.HIPGCG_MultiLinearSolverSetup_BlockPCG32(81, rate, 0.0, None, 0.0, table.ops)
            lib.HIPGCG_BlockPCG32_SetGraphs(table.ops, graphs)
            lib.HIPGCG_BlockPCG32_SetFused(table.ops, fused)
            table.upload(hx, np.zeros((n, nc)))
            abi.slot(table.ops, "MultiLinearSolver")(dA.handle, hb, hx, ia(0, 0), ia(nc, nc),
                                                     table.ops)
            outs.append(table.download(hx, 0, nc, n))
    finally:
        lib.HIPGCG_BlockPCG32_SetFused(table.ops, 1)
        lib.HIPGCG_BlockPCG32_SetGraphs(table.ops, 1)
        table.destroy_multivec(hb, nc)
        table.destroy_multivec(hx, nc)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    r = np.linalg.norm(b - A @ outs[0], axis=0) / np.linalg.norm(b, axis=0)
    assert r.max() < (1.0 if rate == 0.0 else rate * 1.01)


def test_five_launch_iteration_above_threshold(hip):
    """Past 65,536 rows the iteration keeps its separate reductions."""
    A, B = P.stencil_pair(264, 2)  # 69,696 rows
    table = HipTable()
    dA = DeviceMatrix(A, fp32=True)
    rel, _ = run(table, hip, A, dA, None, 0.0, None, None, 400, 1e-2, 9)
    assert rel.max() <= 1.01e-2


@pytest.mark.parametrize("graphs", [0, 1])
def test_block_columns_equal_columns_solved_alone(hip, graphs):
    """Past 65,536 rows each column's steps depend on that column alone
    (its alpha, beta, stop flag and partial sums), so a column of a block
    solve must equal the same column solved by itself, bit for bit.  With
    17 columns (three tiles) stopping at different iterations, this covers
    the d step deferred into the next p update: a tile leaving the list,
    at a read of the active count or between reads, takes its last step
    once, and the loop's last step is applied after it."""
    A, _ = P.stencil_pair(264, 2)  # 69,696 rows
    n, nc = A.shape[0], 17
    table = HipTable()
    dA = DeviceMatrix(A, fp32=True)
    b = np.random.default_rng(91).standard_normal((n, nc))
    hb = table.create_multivec(nc, dA)
    hx = table.create_multivec(nc, dA)
    table.upload(hb, b)
    try:
        hip.HIPGCG_MultiLinearSolverSetup_BlockPCG32(400, 1e-2, 0.0, None, 0.0, table.ops)
        hip.HIPGCG_BlockPCG32_SetGraphs(table.ops, graphs)
        solve = abi.slot(table.ops, "MultiLinearSolver")
        table.upload(hx, np.zeros((n, nc)))
        solve(dA.handle, hb, hx, ia(0, 0), ia(nc, nc), table.ops)
        block = table.download(hx, 0, nc, n)
        for c in range(nc):
            table.upload(hx, np.zeros((n, nc)))
            solve(dA.handle, hb, hx, ia(c, c), ia(c + 1, c + 1), table.ops)
            alone = table.download(hx, 0, nc, n)
            assert np.array_equal(alone[:, c], block[:, c]), f"column {c}"
            r = np.linalg.norm(b[:, c] - A @ alone[:, c]) / np.linalg.norm(b[:, c])
            assert r <= 1.01e-2
    finally:
        hip.HIPGCG_BlockPCG32_SetGraphs(table.ops, 1)
        table.destroy_multivec(hb, nc)
        table.destroy_multivec(hx, nc)

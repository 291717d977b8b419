"""The launch-shape tuner on the GPU: every SpMM launch shape gives the
same bits (and so does the matrix-free 7-point kernel); a solve with tuning on records its choices, and the saved table
reproduces that solve bit for bit."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import gcge_oracle as O
from gcge_amd import abi, problems as P
from gcge_amd import solver as S
from gcge_amd.solver import DeviceMatrix, HipTable

pytestmark = pytest.mark.gpu


def ncols_class(k):
    return 0 if k <= 1 else 1 if k <= 2 else 2 if k <= 4 else 3 if k <= 8 else 4 if k <= 32 else 5


def entries(lib):
    out = []
    i = 0
    while True:
        v = [C.c_int() for _ in range(5)]
        d = [C.c_double(), C.c_double()]
        if not lib.gcge_tune_entry(i, *[C.byref(x) for x in v], *[C.byref(x) for x in d]):
            return out
        out.append(tuple(x.value for x in v) + (d[0].value, d[1].value))
        i += 1


@pytest.fixture
def lib(hip, tmp_path):
    """The tuner with an empty table; the process's table and mode are
    restored afterwards."""
    lib = abi.load()
    keep = str(tmp_path / "kept.txt")
    lib.gcge_tune_save(keep.encode())
    mode = lib.gcge_tune_get_mode()
    lib.gcge_tune_clear()
    yield lib
    lib.gcge_tune_clear()
    lib.gcge_tune_load(keep.encode())
    lib.gcge_tune_set_mode(mode)


def nine_point(n):
    T = sp.diags([np.ones(n - 1), np.ones(n), np.ones(n - 1)], [-1, 0, 1])
    M = sp.csc_matrix(sp.kron(T, T) + 8.0 * sp.identity(n * n))
    M.sort_indices()
    return M


@pytest.mark.parametrize("kind", ["7pt", "9pt", "27pt"])
def test_spmm_shapes_give_identical_bits(lib, kind):
    if kind == "7pt":
        A, W, ncand = P.stencil_pair(30, 3)[0], 8, 15
    elif kind == "9pt":
        A, W, ncand = nine_point(150), 16, 11
    else:
        A, W, ncand = P.config3_grid(18).A.to_ccs(), 32, 6
    n = A.shape[0]
    table = HipTable()
    dA = DeviceMatrix(A, general=True)  # the SELL shapes, not the matrix-free stencil
    lg = int(np.floor(np.log2(n)))
    for k in (1, 5, 24):
        X = np.random.default_rng(k).standard_normal((n, k))
        hx = table.create_multivec(k, dA)
        table.upload(hx, X)
        outs = []
        for ch in range(ncand):
            assert lib.gcge_tune_set(1, W, lg, ncols_class(k), ch) == 0
            hy = table.create_multivec(k, dA)
            table.slot("MatDotMultiVec")(dA.handle, hx, hy, (C.c_int * 2)(0, 0),
                                          (C.c_int * 2)(k, k), table.ops)
            outs.append(table.download(hy, 0, k, n))
            table.destroy_multivec(hy, k)
        if kind == "7pt":
            # and the matrix-free stencil kernel that the same matrix gets
            # when it is recognised (no tuned shapes of its own)
            dS = DeviceMatrix(A)
            assert dS.stencil() == 30
            hy = table.create_multivec(k, dS)
            table.slot("MatDotMultiVec")(dS.handle, hx, hy, (C.c_int * 2)(0, 0),
                                          (C.c_int * 2)(k, k), table.ops)
            outs.append(table.download(hy, 0, k, n))
            table.destroy_multivec(hy, k)
            dS.free()
        table.destroy_multivec(hx, k)
        print(f"spmm {kind} N={n} k={k}: {ncand} shapes", flush=True)
        for o in outs[1:]:
            assert np.array_equal(o, outs[0])
        ref = O.spmm(A, X)
        assert np.max(np.abs(outs[0] - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_tuned_solve_is_recorded_and_reproduced(lib, tmp_path):
    prob = P.config4_grid(3)
    kw = dict(shift=0, mode="mixed", rand="device", seed=3)
    print(f"tuned solve {prob.name}", flush=True)
    lib.gcge_tune_set_mode(1)
    r1 = S.solve_problem(prob, **kw)
    lib.gcge_tune_set_mode(0)
    got = entries(lib)
    ops = {e[0] for e in got}
    # op 3 only when the BlockPCG runs one launch per step (larger systems)
    assert {1, 2} <= ops <= {1, 2, 3}, got
    for op, a, b, c, ch, best, dflt in got:
        assert 0 <= ch < {1: 15, 2: 10, 3: 2}[op]
        assert best <= dflt          # the default is among the candidates
    path = str(tmp_path / "table.txt")
    assert lib.gcge_tune_save(path.encode()) == len(got)
    # heuristics only: same spectrum, not necessarily the same bits
    lib.gcge_tune_clear()
    r0 = S.solve_problem(prob, **kw)
    k = min(r0.nev_conv, r1.nev_conv)
    assert k >= prob.nev
    assert np.max(np.abs(r0.eval[:k] - r1.eval[:k]) / r1.eval[:k]) <= 1e-10
    # the saved table reproduces the tuned solve exactly
    assert lib.gcge_tune_load(path.encode()) == len(got)
    r2 = S.solve_problem(prob, **kw)
    assert r2.num_iter == r1.num_iter and r2.nev_conv == r1.nev_conv
    assert np.array_equal(r2.eval[:r2.nev_conv], r1.eval[:r1.nev_conv])

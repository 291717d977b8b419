"""Slot-by-slot parity of the HIP table (OPS_HIP_Set) with the oracle.

Each test calls a slot through the C ABI (the function pointer in the
table) on seeded inputs and compares with the oracle's restatement of the
reference semantics (oracle/gcge_oracle.py), within FP64 rounding.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import gcge_oracle as O
from gcge_amd import abi, problems as P
from gcge_amd.solver import DeviceMatrix, HipTable

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def ia(*v):
    return (C.c_int * len(v))(*v)


def rand_sym(n, density, seed, extra_wide_row=False):
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=density, random_state=rng, format="csr")
    M = M + M.T + sp.identity(n) * 4.0
    if extra_wide_row:
        # one row/column with 40 entries exercises the generic wide-row kernel
        cols = rng.choice(n, 40, replace=False)
        W = sp.csr_matrix((np.ones(40), (np.zeros(40, int), cols)), shape=(n, n))
        M = M + W + W.T
    M = sp.csc_matrix(M)
    M.sort_indices()
    return M


@pytest.fixture(scope="module")
def table(hip):
    return HipTable()


@pytest.fixture(scope="module")
def mats(hip):
    A, B = P.stencil_pair(24, 2)          # N = 576
    R = rand_sym(3001, 0.002, 7)          # N = 3001: a partial last SELL slice
    Rw = rand_sym(1500, 0.003, 8, True)   # width > 32
    return {"A": (A, DeviceMatrix(A)), "B": (B, DeviceMatrix(B)),
            "R": (R, DeviceMatrix(R)), "Rw": (Rw, DeviceMatrix(Rw))}


def upload(table, X, mat):
    h = table.create_multivec(X.shape[1], mat)
    table.upload(h, X)
    return h


def close(a, b, tol=RTOL):
    scale = max(1.0, float(np.max(np.abs(b))))
    assert np.max(np.abs(a - b)) <= tol * scale, np.max(np.abs(a - b))


@pytest.mark.parametrize("key", ["A", "R", "Rw"])
@pytest.mark.parametrize("ncols", [1, 7, 8, 13, 40])
def test_matdotmultivec(table, mats, key, ncols):
    M, dM = mats[key]
    n = M.shape[0]
    rng = np.random.default_rng(ncols)
    X = rng.standard_normal((n, ncols + 3))
    Y0 = rng.standard_normal((n, ncols + 5))
    hx, hy = upload(table, X, dM), upload(table, Y0, dM)
    table.slot("MatDotMultiVec")(dM.handle, hx, hy, ia(2, 4), ia(2 + ncols, 4 + ncols), table.ops)
    Y = table.download(hy, 0, ncols + 5, n)
    ref = Y0.copy()
    ref[:, 4:4 + ncols] = O.spmm(M, X[:, 2:2 + ncols])
    close(Y, ref)
    # mat == NULL copies (app/app_ccs.c:134-137)
    table.slot("MatDotMultiVec")(None, hx, hy, ia(0, 1), ia(1, 2), table.ops)
    Y = table.download(hy, 1, 2, n)
    close(Y[:, 0], X[:, 0], 0.0)
    table.destroy_multivec(hx, ncols + 3)
    table.destroy_multivec(hy, ncols + 5)


@pytest.mark.parametrize("alpha,beta,use_x", [(2.0, 0.5, True), (1.0, 0.0, True),
                                              (0.0, 3.0, False), (-1.5, 1.0, True)])
def test_axpby(table, mats, alpha, beta, use_x):
    M, dM = mats["R"]
    n = M.shape[0]
    rng = np.random.default_rng(1)
    X = rng.standard_normal((n, 6))
    Y0 = rng.standard_normal((n, 6))
    if beta == 0.0:
        Y0[:, 3] = np.nan  # beta == 0 overwrites without reading y
    hx, hy = upload(table, X, dM), upload(table, Y0, dM)
    table.slot("MultiVecAxpby")(alpha, hx if use_x else None, beta, hy, ia(1, 2), ia(5, 6),
                                table.ops)
    Y = table.download(hy, 0, 6, n)
    ref = Y0.copy()
    O.axpby(alpha, X[:, 1:5] if use_x else None, beta, ref[:, 2:6])
    close(Y, ref)


@pytest.mark.parametrize("mode", ["overwrite", "scalar", "percol", "stride2", "scale_only",
                                  "none_none"])
def test_linear_comb(table, mats, mode):
    M, dM = mats["R"]
    n = M.shape[0]
    rng = np.random.default_rng(2)
    X = rng.standard_normal((n, 70))
    Y0 = rng.standard_normal((n, 90))
    m, k = 67, 75
    ldc = m + 3
    coef = rng.standard_normal((ldc, k))
    hx, hy = upload(table, X, dM), upload(table, Y0, dM)
    cf = np.asfortranarray(coef)
    beta_arr = rng.standard_normal(2 * k)
    s, e = ia(2, 10), ia(2 + m, 10 + k)
    ref = Y0.copy()
    fn = table.slot("MultiVecLinearComb")
    cptr = cf.ctypes.data_as(abi.c_dbl_p)
    if mode == "overwrite":
        fn(hx, hy, 0, s, e, cptr, ldc, None, 0, table.ops)
        O.linear_comb(X[:, 2:2 + m], ref[:, 10:10 + k], coef[:m], None)
    elif mode == "scalar":
        b = C.c_double(0.25)
        fn(hx, hy, 0, s, e, cptr, ldc, C.byref(b), 0, table.ops)
        O.linear_comb(X[:, 2:2 + m], ref[:, 10:10 + k], coef[:m], 0.25)
    elif mode in ("percol", "stride2"):
        inc = 1 if mode == "percol" else 2
        fn(hx, hy, 0, s, e, cptr, ldc, beta_arr.ctypes.data_as(abi.c_dbl_p), inc, table.ops)
        O.linear_comb(X[:, 2:2 + m], ref[:, 10:10 + k], coef[:m], beta_arr[::inc][:k])
    elif mode == "scale_only":
        fn(None, hy, 0, s, e, None, 0, beta_arr.ctypes.data_as(abi.c_dbl_p), 1, table.ops)
        O.linear_comb(None, ref[:, 10:10 + k], None, beta_arr[:k])
    else:
        fn(None, hy, 0, s, e, None, 0, None, 0, table.ops)
    Y = table.download(hy, 0, 90, n)
    close(Y, ref)


def test_linear_comb_aliased(table, mats):
    """x == y with disjoint column ranges (src/ops_orth.c:90, 254)."""
    M, dM = mats["A"]
    n = M.shape[0]
    rng = np.random.default_rng(3)
    V0 = rng.standard_normal((n, 30))
    hv = upload(table, V0, dM)
    coef = rng.standard_normal((10, 12))
    one = C.c_double(1.0)
    table.slot("MultiVecLinearComb")(hv, hv, 0, ia(0, 15), ia(10, 27),
                                     np.asfortranarray(coef).ctypes.data_as(abi.c_dbl_p), 10,
                                     C.byref(one), 0, table.ops)
    ref = V0.copy()
    ref[:, 15:27] += V0[:, 0:10] @ coef
    close(table.download(hv, 0, 30, n), ref)


@pytest.mark.parametrize("nsd", ["N", "S", "D"])
@pytest.mark.parametrize("shape", [(5, 5), (64, 64), (130, 70), (1, 1), (40, 40), (130, 40),
                                   (200, 17), (100, 12), (65, 64), (40, 1), (130, 2), (17, 4),
                                   (3, 100)])
def test_inner_prod(table, mats, nsd, shape):
    M, dM = mats["R"]
    n = M.shape[0]
    m, k = shape
    if nsd in "SD":
        k = m
    rng = np.random.default_rng(m * 7 + k)
    X = rng.standard_normal((n, m + 2))
    Y = X.copy() if nsd == "S" else rng.standard_normal((n, k + 1))
    if nsd == "S":
        Y[:, 2:2 + k] = X[:, 2:2 + k] + 1e-3 * rng.standard_normal((n, k))
    hx, hy = upload(table, X, dM), upload(table, Y, dM)
    ld = m + 4 if nsd != "D" else 3
    ip = np.full(ld * k + 8, 99.0)
    table.slot("MultiVecInnerProd")(nsd.encode(), hx, hy, 0, ia(2, 0 if nsd != "S" else 2),
                                    ia(2 + m, k if nsd != "S" else 2 + k),
                                    ip.ctypes.data_as(abi.c_dbl_p), ld, table.ops)
    ys = Y[:, 0:k] if nsd != "S" else Y[:, 2:2 + k]
    G = O.inner_prod(nsd, X[:, 2:2 + m], ys)
    if nsd == "D":
        got = ip[0:ld * k:ld]
        close(got, G)
        # entries between the strided diagonal slots are untouched
        assert np.all(ip[1:ld * k:ld] == 99.0)
    else:
        got = ip[: ld * k].reshape((ld, k), order="F")
        close(got[:m], G)
        assert np.all(got[m:] == 99.0)


@pytest.mark.parametrize("ntsd", ["N", "S", "T"])
def test_qtap(table, mats, ntsd):
    (A, dA), (B, dB) = mats["A"], mats["B"]
    n = A.shape[0]
    rng = np.random.default_rng(5)
    V = rng.standard_normal((n, 40))
    hv = upload(table, V, dA)
    ws = table.create_multivec(20, dA)
    if ntsd == "S":
        s, e = ia(10, 10), ia(22, 22)
        nr, nc = 12, 12
    else:
        s, e = ia(0, 25), ia(18, 37)
        nr, nc = 18, 12
    ld = (nc if ntsd == "T" else nr) + 1
    out = np.zeros(ld * max(nr, nc))
    table.slot("MultiVecQtAP")(b"S", ntsd.encode(), hv, dB.handle, hv, 0, s, e,
                               out.ctypes.data_as(abi.c_dbl_p), ld, ws, table.ops)
    Q, Pm = V[:, s[0]:e[0]], V[:, s[1]:e[1]]
    wsref = np.zeros((n, 20))
    G = O.qtap("S", ntsd, Q, B, Pm, wsref)
    rows = G.shape[0]
    got = out[: ld * G.shape[1]].reshape((ld, G.shape[1]), order="F")[:rows]
    close(got, G)
    # side effect: mv_ws[:, :nc] holds B P (src/ops_orth.c:316-323)
    close(table.download(ws, 0, nc, n), wsref[:, :nc])
    # A == NULL: plain inner product, mv_ws untouched
    table.slot("MultiVecQtAP")(b"S", b"N", hv, None, hv, 0, ia(0, 30), ia(5, 33),
                               out.ctypes.data_as(abi.c_dbl_p), 5, ws, table.ops)
    close(out[:15].reshape((5, 3), order="F"), V[:, 0:5].T @ V[:, 30:33])


def test_random_libc_sequence(table, mats, hip):
    """The default random fill reproduces libc rand() after srand(s),
    column by column (app/app_lapack.c:322-333)."""
    M, dM = mats["A"]
    n = M.shape[0]
    hv = table.create_multivec(4, dM)
    hip.gcge_hip_set_random_mode(abi.GCGE_HIP_RAND_LIBC, 0)
    O.libc_srand(11)
    abi.slot(table.ops, "MultiVecSetRandomValue")(hv, 1, 3, table.ops)
    got = table.download(hv, 0, 4, n)
    O.libc_srand(11)
    ref = O.libc_uniform(n, 2)
    assert np.array_equal(got[:, 1:3], ref)
    assert np.all(got[:, [0, 3]] == 0.0)


def test_random_device(table, mats, hip):
    M, dM = mats["R"]
    n = M.shape[0]
    hv = table.create_multivec(8, dM)
    hip.gcge_hip_set_random_mode(abi.GCGE_HIP_RAND_DEVICE, 123)
    abi.slot(table.ops, "MultiVecSetRandomValue")(hv, 0, 8, table.ops)
    got = table.download(hv, 0, 8, n)
    hip.gcge_hip_set_random_mode(abi.GCGE_HIP_RAND_LIBC, 0)
    assert got.min() >= 0.0 and got.max() < 1.0
    assert abs(got.mean() - 0.5) < 0.01
    assert np.linalg.matrix_rank(got) == 8


def test_info(mats, hip):
    M, dM = mats["A"]
    info = dM.info()
    assert info["nrows"] == M.shape[0] and info["nnz"] == M.nnz
    assert info["stored"] >= M.nnz and info["stored"] % 64 == 0


@pytest.fixture(scope="module")
def tall(hip):
    """A matrix handle of 300,000 rows, for multivectors past the row count
    (262,144) where LinearComb's operand ring is 5 steps deep."""
    return DeviceMatrix(sp.identity(300_000, format="csc"))


@pytest.mark.parametrize("mode", ["overwrite", "scalar", "percol"])
def test_linear_comb_tall(table, tall, mode):
    """Y = X C + Y diag(beta) at 300,000 rows, 230 -> 130 columns (two
    64-column output tiles plus a partial one), every beta mode."""
    n, m, k = 300_000, 230, 130
    rng = np.random.default_rng(12)
    X = rng.standard_normal((n, m + 1))
    Y0 = rng.standard_normal((n, k + 2))
    ldc = m + 1
    coef = np.asfortranarray(rng.standard_normal((ldc, k)))
    beta_arr = rng.standard_normal(k)
    hx, hy = upload(table, X, tall), upload(table, Y0, tall)
    fn = table.slot("MultiVecLinearComb")
    s, e = ia(1, 2), ia(1 + m, 2 + k)
    cptr = coef.ctypes.data_as(abi.c_dbl_p)
    ref = Y0.copy()
    if mode == "overwrite":
        fn(hx, hy, 0, s, e, cptr, ldc, None, 0, table.ops)
        O.linear_comb(X[:, 1:1 + m], ref[:, 2:2 + k], coef[:m], None)
    elif mode == "scalar":
        b = C.c_double(-0.5)
        fn(hx, hy, 0, s, e, cptr, ldc, C.byref(b), 0, table.ops)
        O.linear_comb(X[:, 1:1 + m], ref[:, 2:2 + k], coef[:m], -0.5)
    else:
        fn(hx, hy, 0, s, e, cptr, ldc, beta_arr.ctypes.data_as(abi.c_dbl_p), 1, table.ops)
        O.linear_comb(X[:, 1:1 + m], ref[:, 2:2 + k], coef[:m], beta_arr)
    Y = table.download(hy, 0, k + 2, n)
    table.destroy_multivec(hx, m + 1)
    table.destroy_multivec(hy, k + 2)
    close(Y, ref)


@pytest.mark.parametrize("nsd", ["N", "S"])
def test_inner_prod_tall(table, tall, nsd):
    """Matrix-core Gram at 300,000 rows: 160 x 130 ('N') and 130 x 130
    ('S'), with ldIP larger than the block."""
    n = 300_000
    m, k = (160, 130) if nsd == "N" else (130, 130)
    rng = np.random.default_rng(13)
    X = rng.standard_normal((n, m + 2))
    Y = X.copy() if nsd == "S" else rng.standard_normal((n, k))
    hx = upload(table, X, tall)
    hy = hx if nsd == "S" else upload(table, Y, tall)
    ld = m + 3
    ip = np.full(ld * k, 99.0)
    sy = 2 if nsd == "S" else 0
    table.slot("MultiVecInnerProd")(nsd.encode(), hx, hy, 0, ia(2, sy), ia(2 + m, sy + k),
                                    ip.ctypes.data_as(abi.c_dbl_p), ld, table.ops)
    G = O.inner_prod(nsd, X[:, 2:2 + m], (X if nsd == "S" else Y)[:, sy:sy + k])
    got = ip.reshape((ld, k), order="F")
    close(got[:m], G)
    assert np.all(got[m:] == 99.0)
    table.destroy_multivec(hx, m + 2)
    if nsd == "N":
        table.destroy_multivec(hy, k)


@pytest.mark.parametrize("nsd,m,k", [("N", 200, 40), ("N", 90, 12), ("S", 40, 40), ("S", 64, 64),
                                     ("N", 40, 1), ("N", 130, 2), ("N", 33, 8)])
def test_inner_prod_tall_narrow(table, tall, nsd, m, k):
    """The Gram's narrow tile shapes at 300,000 rows: a side of at most 64
    columns (256 x 16 NB tiles) and both sides at most 64 (64 x 64), and
    the thin stream for a side of at most 8 (16 X columns per workgroup
    for one or two Y columns, several column tiles)."""
    n = 300_000
    rng = np.random.default_rng(m + k)
    X = rng.standard_normal((n, m + 2))
    Y = X.copy() if nsd == "S" else rng.standard_normal((n, k))
    hx = upload(table, X, tall)
    hy = hx if nsd == "S" else upload(table, Y, tall)
    ld = m + 1
    ip = np.full(ld * k, 99.0)
    sy = 2 if nsd == "S" else 0
    table.slot("MultiVecInnerProd")(nsd.encode(), hx, hy, 0, ia(2, sy), ia(2 + m, sy + k),
                                    ip.ctypes.data_as(abi.c_dbl_p), ld, table.ops)
    G = O.inner_prod(nsd, X[:, 2:2 + m], (X if nsd == "S" else Y)[:, sy:sy + k])
    got = ip.reshape((ld, k), order="F")
    close(got[:m], G)
    assert np.all(got[m:] == 99.0)
    table.destroy_multivec(hx, m + 2)
    if nsd == "N":
        table.destroy_multivec(hy, k)


@pytest.mark.parametrize("mode", ["overwrite", "scalar", "percol"])
def test_linear_comb_tall_narrow(table, tall, mode):
    """Y = X C + Y diag(beta) at 300,000 rows into 40 columns (48-wide wave
    tiles), every beta mode."""
    n, m, k = 300_000, 150, 40
    rng = np.random.default_rng(14)
    X = rng.standard_normal((n, m))
    Y0 = rng.standard_normal((n, k))
    coef = np.asfortranarray(rng.standard_normal((m, k)))
    beta_arr = rng.standard_normal(k)
    hx, hy = upload(table, X, tall), upload(table, Y0, tall)
    fn = table.slot("MultiVecLinearComb")
    s, e = ia(0, 0), ia(m, k)
    cptr = coef.ctypes.data_as(abi.c_dbl_p)
    ref = Y0.copy()
    if mode == "overwrite":
        fn(hx, hy, 0, s, e, cptr, m, None, 0, table.ops)
        O.linear_comb(X, ref, coef, None)
    elif mode == "scalar":
        b = C.c_double(0.25)
        fn(hx, hy, 0, s, e, cptr, m, C.byref(b), 0, table.ops)
        O.linear_comb(X, ref, coef, 0.25)
    else:
        fn(hx, hy, 0, s, e, cptr, m, beta_arr.ctypes.data_as(abi.c_dbl_p), 1, table.ops)
        O.linear_comb(X, ref, coef, beta_arr)
    close(table.download(hy, 0, k, n), ref)
    table.destroy_multivec(hx, m)
    table.destroy_multivec(hy, k)


@pytest.mark.parametrize("mode", ["overwrite", "scalar"])
def test_linear_comb_library_gemm(table, hip, mode):
    """From 448 Ki rows with at least 100 source and destination columns
    and a scalar beta, LinearComb is rocBLAS's dgemm: the same results as
    the oracle at 1,100,000 rows, 110 -> 100 columns."""
    n, m, k = 1_100_000, 110, 100
    big = DeviceMatrix(sp.identity(n, format="csc"))
    rng = np.random.default_rng(15)
    X = rng.standard_normal((n, m))
    Y0 = rng.standard_normal((n, k))
    coef = np.asfortranarray(rng.standard_normal((m + 3, k)))
    hx, hy = upload(table, X, big), upload(table, Y0, big)
    fn = table.slot("MultiVecLinearComb")
    cptr = coef.ctypes.data_as(abi.c_dbl_p)
    ref = Y0.copy()
    if mode == "overwrite":
        fn(hx, hy, 0, ia(0, 0), ia(m, k), cptr, m + 3, None, 0, table.ops)
        O.linear_comb(X, ref, coef[:m], None)
    else:
        b = C.c_double(-1.5)
        fn(hx, hy, 0, ia(0, 0), ia(m, k), cptr, m + 3, C.byref(b), 0, table.ops)
        O.linear_comb(X, ref, coef[:m], -1.5)
    close(table.download(hy, 0, k, n), ref)
    table.destroy_multivec(hx, m)
    table.destroy_multivec(hy, k)
    big.free()

"""The allocation-free inner products (HIP_MultiVecInnerProdNoAlloc)
against the table's MultiVecInnerProd slot they are a variant of, on two
calls in a row, the second with smaller inputs, for the 'N', 'S' and 'D'
forms, both checked against the oracle.  The first call's 130 x 70 block
over 70,000 rows needs more Gram partial sums in the slot's default
chunking than the fixed array holds, so the variant groups the rows in
fewer chunks and agrees to FP64 rounding; the second call's plan fits and
the results are the slot's bit for bit.  A block too large for the array
is refused with an error, not run."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import gcge_oracle as O
from gcge_amd import abi
from gcge_amd.solver import DeviceMatrix, HipTable

pytestmark = pytest.mark.gpu


def ia(*v):
    return (C.c_int * len(v))(*v)


def rows(n):
    """A device matrix of order n, for multivectors of n rows."""
    return DeviceMatrix(sp.identity(n, format="csc"))


def both(table, lib, nsd, hx, hy, s, e, ld, size):
    a = np.full(size, 99.0)
    b = np.full(size, 99.0)
    table.slot("MultiVecInnerProd")(nsd.encode(), hx, hy, 0, ia(*s), ia(*e),
                                    a.ctypes.data_as(abi.c_dbl_p), ld, table.ops)
    rc = lib.HIP_MultiVecInnerProdNoAlloc(nsd.encode(), hx, hy, 0, ia(*s), ia(*e),
                                          b.ctypes.data_as(abi.c_dbl_p), ld)
    assert rc == 0
    return a, b


@pytest.mark.parametrize("nsd", ["N", "S", "D"])
def test_noalloc_matches_slot_twice(hip, nsd):
    table = HipTable()
    # first call: 70,000 rows (several row chunks), then 3,001 rows and
    # fewer columns (a partial last slice and the streaming kernel for 'N')
    for n, m, k in ((70000, 130, 70), (3001, 9, 5)):
        dM = rows(n)
        rng = np.random.default_rng(n + m)
        if nsd in "SD":
            k = m
        X = rng.standard_normal((n, m + 2))
        Y = rng.standard_normal((n, k + 1))
        hx = table.create_multivec(X.shape[1], dM)
        hy = table.create_multivec(Y.shape[1], dM)
        table.upload(hx, X)
        table.upload(hy, Y)
        if nsd == "S":
            s, e, ys = (2, 2), (2 + m, 2 + m), X[:, 2:2 + m]
            hy_use = hx
        else:
            s, e, ys = (2, 0), (2 + m, k), Y[:, 0:k]
            hy_use = hy
        ld = m + 4 if nsd != "D" else 3
        size = ld * k + 8
        a, b = both(table, hip, nsd, hx, hy_use, s, e, ld, size)
        G = O.inner_prod(nsd, X[:, 2:2 + m], ys)
        scale = max(1.0, float(np.max(np.abs(G))))
        if n == 3001 or nsd == "D":
            assert np.array_equal(a, b)
        else:
            assert np.max(np.abs(a - b)) <= 1e-13 * scale
        if nsd == "D":
            got = b[0:ld * k:ld]
            assert np.all(b[1:ld * k:ld] == 99.0)
        else:
            got = b[:ld * k].reshape((ld, k), order="F")
            assert np.all(got[m:] == 99.0)
            got = got[:m]
        assert np.max(np.abs(got - G)) <= 1e-12 * scale
        table.destroy_multivec(hx, X.shape[1])
        table.destroy_multivec(hy, Y.shape[1])


def test_noalloc_refuses_oversize(hip):
    table = HipTable()
    n, m = 64, 2100  # a 2100 x 2100 result exceeds the 4 Mi-double array
    hx = table.create_multivec(m, rows(n))
    out = np.zeros(8)
    rc = hip.HIP_MultiVecInnerProdNoAlloc(b"N", hx, hx, 0, ia(0, 0), ia(m, m),
                                          out.ctypes.data_as(abi.c_dbl_p), m)
    assert rc == -1
    buf = C.create_string_buffer(512)
    hip.gcge_hip_last_error(buf, 512)
    assert b"scratch" in buf.value
    table.destroy_multivec(hx, m)

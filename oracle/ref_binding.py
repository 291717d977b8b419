"""ctypes binding of oracle/_ref (the reference GCGE compiled from its own
sources by oracle/Makefile) and of oracle/ref_harness.c.

TEST INFRASTRUCTURE: used by tests/ and bench.py's cpu_baseline leg only,
as the checker / the baseline; never by the product.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT = os.path.join(HERE, "_ref", "libgcge_ref.so")
DEFAULT_OMP = os.path.join(HERE, "_ref", "libgcge_ref_omp.so")


class RefLib:
    def __init__(self, path: str = DEFAULT):
        self.path = path
        lib = C.CDLL(path, mode=os.RTLD_LOCAL if hasattr(os, "RTLD_LOCAL") else 0)
        for f in ("ref_ops_create_ccs", "ref_ops_create_lapack", "ref_ops_create_empty"):
            getattr(lib, f).restype = C.c_void_p
            getattr(lib, f).argtypes = []
        lib.ref_ops_setup.argtypes = [C.c_void_p]
        lib.ref_ops_destroy.argtypes = [C.c_void_p]
        lib.ref_ccs_wrap.restype = C.c_void_p
        lib.ref_ccs_wrap.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ref_dense_wrap.restype = C.c_void_p
        lib.ref_dense_wrap.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p]
        lib.ref_free.argtypes = [C.c_void_p]
        lib.ref_srand.argtypes = [C.c_uint]
        lib.ref_solve.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                  C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                  C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                  C.c_void_p]
        lib.ref_phase_times.argtypes = [C.c_void_p]
        self.lib = lib

    # ---- tables -------------------------------------------------------
    def ccs_table(self) -> int:
        return self.lib.ref_ops_create_ccs()

    def lapack_table(self) -> int:
        return self.lib.ref_ops_create_lapack()

    def empty_table(self) -> int:
        return self.lib.ref_ops_create_empty()

    def setup(self, ops: int) -> None:
        self.lib.ref_ops_setup(ops)

    # ---- objects ------------------------------------------------------
    def wrap_ccs(self, A):
        """CCSMAT view of a scipy CSC matrix; returns (handle, keepalive)."""
        A = A.tocsc()
        indptr = np.ascontiguousarray(A.indptr, dtype=np.int32)
        indices = np.ascontiguousarray(A.indices, dtype=np.int32)
        data = np.ascontiguousarray(A.data, dtype=np.float64)
        h = self.lib.ref_ccs_wrap(A.shape[0], indptr.ctypes.data, indices.ctypes.data,
                                  data.ctypes.data)
        return h, (indptr, indices, data)

    def wrap_dense(self, M: np.ndarray):
        M = np.asfortranarray(M, dtype=np.float64)
        h = self.lib.ref_dense_wrap(M.shape[0], M.shape[1], M.shape[0], M.ctypes.data)
        return h, M

    # ---- the solve ----------------------------------------------------
    def solve(self, ops: int, A: int, B: int, nrows: int, nev: int, block_size: int,
              nev_init: int, nev_max: int, tol=(1e-1, 1e-8), gap_min: float = 1e-5,
              shift: int = 1, max_iter: int = 1000, print_level: int = 0,
              host_vectors: bool = True, want_vectors: bool = False):
        ev = np.zeros(nev_max)
        st = np.zeros(2, np.int32)
        tm = np.zeros(1)
        vec = np.zeros((nrows, nev_max), order="F") if (want_vectors and host_vectors) else None
        self.lib.ref_solve(ops, A, B or None, nrows, nev, block_size, nev_init, nev_max,
                           tol[0], tol[1], gap_min, shift, max_iter, print_level,
                           ev.ctypes.data, vec.ctypes.data if vec is not None else None,
                           1 if host_vectors else 0, st.ctypes.data, tm.ctypes.data)
        return {"eval": ev, "num_iter": int(st[0]), "nev_conv": int(st[1]),
                "seconds": float(tm[0]), "evec": vec}

    def solve_ccs(self, prob, shift: int = 1, max_iter: int = 1000, want_vectors=False,
                  print_level: int = 0):
        ops = self.ccs_table()
        hA, ka = self.wrap_ccs(prob.A)
        hB, kb = self.wrap_ccs(prob.B) if prob.B is not None else (None, None)
        out = self.solve(ops, hA, hB, prob.n, prob.nev, prob.block_size, prob.nev_init,
                         prob.nev_max, (prob.tol_abs, prob.tol_rel), prob.gap_min, shift,
                         max_iter, print_level=print_level, want_vectors=want_vectors)
        self.lib.ref_free(hA)
        if hB:
            self.lib.ref_free(hB)
        self.lib.ref_ops_destroy(ops)
        del ka, kb
        return out

    def phase_times(self):
        out = np.zeros(11)
        self.lib.ref_phase_times(out.ctypes.data)
        return dict(zip(("initX", "checkconv", "compP", "compRR", "rr_matW", "dsyevx",
                         "compRV", "compW", "compX", "linsol", "total"), out))

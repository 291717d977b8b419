"""Concurrent solves on one GPU through the library's solve service.

The service (gcge_service_* in libgcge_hip.so) runs a pool of native
worker threads, each on its own device stream; requests are taken in
arrival order as soon as a worker is free and their device footprint fits
in the HBM budget.  This module keeps each request's host arrays alive
until its result is collected.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import abi


@dataclass
class ServedResult:
    eval: np.ndarray
    nev_conv: int
    num_iter: int
    concurrent: int
    queue_seconds: float
    setup_seconds: float
    solve_seconds: float
    total_seconds: float
    evec: np.ndarray | None = None
    pool_bytes: int = 0


def _ccs(M):
    M = M.tocsc()
    keep = (np.ascontiguousarray(M.indptr, dtype=np.int32),
            np.ascontiguousarray(M.indices, dtype=np.int32),
            np.ascontiguousarray(M.data, dtype=np.float64))
    m = abi.CCSMat(keep[2].ctypes.data_as(abi.c_dbl_p), keep[1].ctypes.data_as(abi.c_int_p),
                   keep[0].ctypes.data_as(abi.c_int_p), M.shape[0], M.shape[1])
    return m, keep


class Service:
    def __init__(self, workers: int = 8, hbm_budget: float = 200e9, batching: bool = True,
                 orth: str = "cholqr_host"):
        self.lib = abi.load()
        self.h = self.lib.gcge_service_create(int(workers), int(hbm_budget))
        if not self.h:
            raise ValueError("gcge_service_create rejected its arguments")
        self.lib.gcge_service_set_batching(self.h, 1 if batching else 0)
        # "cholqr_host" (factorizations on the host, the default) or "cholqr"
        if self.lib.gcge_service_set_orth(
                self.h, {"cholqr": abi.GCGE_ORTH_CHOLQR,
                         "cholqr_host": abi.GCGE_ORTH_CHOLQR_HOST}[orth]) != 0:
            raise ValueError(orth)
        self._live = {}

    @staticmethod
    def request(prob, mode: str = "mixed", seed: int = 0, shift: int = 1,
                max_iter: int = 1000, x0=None):
        """The gcge_request of a problems.CCSProblem (with the host arrays
        it points to).  x0: N x k initial vectors (k <= nev_init) for a warm
        start, or None."""
        A, ka = _ccs(prob.A)
        keep = [A, ka]
        px0, ngiven = None, 0
        if x0 is not None:
            x0 = np.asfortranarray(x0, dtype=np.float64)
            if x0.shape[0] != prob.n or x0.shape[1] > prob.nev_init:
                raise ValueError("x0 must be N x k with k <= nev_init")
            keep.append(x0)
            px0, ngiven = x0.ctypes.data_as(abi.c_dbl_p), x0.shape[1]
        pB = None
        if prob.B is not None:
            B, kb = _ccs(prob.B)
            keep += [B, kb]
            pB = C.pointer(B)
        req = abi.Request(C.pointer(A), pB, prob.nev, prob.block_size, prob.nev_init,
                          prob.nev_max, max_iter,
                          abi.GCGE_MODE_MIXED if mode == "mixed" else abi.GCGE_MODE_FP64,
                          shift, prob.tol_abs, prob.tol_rel, prob.gap_min, seed, px0, ngiven)
        return req, keep

    @staticmethod
    def footprint(prob, mode: str = "mixed") -> int:
        """Device bytes the request holds while it runs (host-side formula)."""
        req, _keep = Service.request(prob, mode)
        return int(abi.load().gcge_request_footprint(C.byref(req)))

    def submit(self, prob, mode: str = "mixed", seed: int = 0, want_vectors: bool = False,
               shift: int = 1, x0=None) -> int:
        req, keep = self.request(prob, mode, seed, shift, x0=x0)
        ev = np.zeros(prob.nev_max)
        vec = np.zeros((prob.n, prob.nev_max), order="F") if want_vectors else None
        res = abi.Result()
        res.eval = ev.ctypes.data_as(abi.c_dbl_p)
        res.evec = vec.ctypes.data_as(abi.c_dbl_p) if vec is not None else None
        t = self.lib.gcge_service_submit(self.h, C.byref(req), C.byref(res))
        if t < 0:
            raise RuntimeError(f"request refused (status {res.status}): footprint "
                               f"{self.lib.gcge_request_footprint(C.byref(req))} bytes")
        self._live[t] = (req, keep, res, ev, vec)
        return t

    def wait(self, ticket: int) -> ServedResult:
        st = self.lib.gcge_service_wait(self.h, ticket)
        req, keep, res, ev, vec = self._live.pop(ticket)
        if st != 0:
            raise RuntimeError(f"request {ticket} failed with status {st}")
        k = res.nev_conv
        return ServedResult(ev[:k].copy(), k, res.num_iter, res.concurrent, res.queue_seconds,
                            res.setup_seconds, res.solve_seconds, res.total_seconds,
                            vec[:, :k].copy() if vec is not None else None, res.pool_bytes)

    def batch_stats(self):
        """Batched inner solves: rounds, request iterations they carried and
        the most requests in one round."""
        r, it, mx = C.c_int64(), C.c_int64(), C.c_int()
        self.lib.gcge_service_batch_stats(self.h, C.byref(r), C.byref(it), C.byref(mx))
        return {"rounds": r.value, "solve_iterations": it.value, "max_solves": mx.value}

    def stats(self):
        done, conc, used = C.c_int64(), C.c_int(), C.c_int64()
        self.lib.gcge_service_stats(self.h, C.byref(done), C.byref(conc), C.byref(used))
        return {"completed": done.value, "max_concurrent": conc.value,
                "hbm_in_use": used.value}

    def close(self):
        if self.h:
            self.lib.gcge_service_destroy(self.h)
            self.h = None
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

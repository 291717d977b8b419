"""Python host side of the GCG eigensolver on MI355X.

Mirrors the reference call sequence of the SLEPc plugin (gcge.c:442-535):
an evec block, the driver workspaces (GCGE_Create), srand(0), the setup
(EigenSolverSetup_GCG + GCGE_Setparameters) and ops->EigenSolver -- here
with the HIP operation table (OPS_HIP_Set) and this library's driver
(HIPGCG_*).  Everything below the table runs in libgcge_hip.so; a missing
library raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import abi, problems

ORTH_METHODS = {"mgs": abi.GCGE_ORTH_MGS, "cholqr": abi.GCGE_ORTH_CHOLQR,
                "cholqr_host": abi.GCGE_ORTH_CHOLQR_HOST}

PHASES = ("initX", "checkconv", "compP", "compRR", "rr_matW", "dsyevx", "compRV",
          "compW", "compX", "linsol")


@dataclass
class SolveResult:
    eval: np.ndarray
    nev_conv: int
    num_iter: int
    seconds: float          # first slot call after the upload -> EigenSolver returns
    bpcg_iters: int = 0
    linsolves: int = 0
    phases: dict = field(default_factory=dict)
    evec: np.ndarray | None = None
    cholqr_blocks: int = 0
    cholqr_fallbacks: int = 0
    solve_seconds: float = 0.0  # EigenSolver alone (workspace creation excluded)


class DeviceMatrix:
    """A symmetric CCS matrix uploaded to the device (HIPCCS_Create).
    general=True keeps a constant-coefficient 7-point grid operator on the
    general sparse kernels instead of the matrix-free stencil ones."""

    def __init__(self, A, fp32: bool = False, general: bool = False):
        lib = abi.load()
        A = A.tocsc()
        self.n = A.shape[0]
        self._indptr = np.ascontiguousarray(A.indptr, dtype=np.int32)
        self._indices = np.ascontiguousarray(A.indices, dtype=np.int32)
        self._data = np.ascontiguousarray(A.data, dtype=np.float64)
        m = abi.CCSMat(self._data.ctypes.data_as(abi.c_dbl_p),
                       self._indices.ctypes.data_as(abi.c_int_p),
                       self._indptr.ctypes.data_as(abi.c_int_p), self.n, self.n)
        h = lib.HIPCCS_Create(C.byref(m), (abi.GCGE_HIP_MAT_FP32 if fp32 else 0) |
                              (abi.GCGE_HIP_MAT_GENERAL if general else 0))
        if not h:
            raise ValueError("HIPCCS_Create rejected the matrix")
        self.handle = C.c_void_p(h)

    @classmethod
    def grid(cls, kind: int, dim: int, points: int, n: int, coef=None,
             delta_rel: float = 0.0, fp32: bool = False,
             general: bool = False) -> "DeviceMatrix":
        """A grid operator assembled on the device (HIPCCS_CreateGrid):
        the same handle HIPCCS_Create gives for the host CCS of the
        operator, from the grid size and the node coefficients only."""
        lib = abi.load()
        c = None
        if coef is not None:
            c = np.ascontiguousarray(coef, dtype=np.float64)
        op = abi.GridOp(kind, dim, points, n,
                        c.ctypes.data_as(abi.c_dbl_p) if c is not None else None, delta_rel)
        h = lib.HIPCCS_CreateGrid(C.byref(op), (abi.GCGE_HIP_MAT_FP32 if fp32 else 0) |
                                  (abi.GCGE_HIP_MAT_GENERAL if general else 0))
        if not h:
            raise ValueError("HIPCCS_CreateGrid rejected the operator")
        self = cls.__new__(cls)
        self.n = n ** dim
        self.handle = C.c_void_p(h)
        return self

    def equal(self, other: "DeviceMatrix") -> bool:
        """True when both handles hold the same device arrays."""
        return bool(abi.load().HIPCCS_Equal(self.handle, other.handle))

    def stencil(self) -> int:
        """Grid side n when products run matrix-free as a 7-point stencil, else 0."""
        return int(abi.load().HIPCCS_Stencil(self.handle))

    def info(self):
        lib = abi.load()
        v = [C.c_int64() for _ in range(4)]
        lib.HIPCCS_Info(self.handle, *[C.byref(x) for x in v])
        return {"nrows": v[0].value, "nnz": v[1].value, "stored": v[2].value,
                "bytes": v[3].value}

    def free(self):
        if self.handle:
            abi.load().HIPCCS_Destroy(C.byref(self.handle))
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HipTable:
    """An operation table filled by OPS_HIP_Set (host dense lapack_ops)."""

    def __init__(self):
        self.lib = abi.load()
        self.ops = abi.new_ops()
        self.lib.OPS_HIP_Set(self.ops)

    def slot(self, name):
        return abi.slot(self.ops, name)

    def create_multivec(self, ncols: int, mat: DeviceMatrix) -> C.c_void_p:
        h = C.c_void_p()
        self.slot("MultiVecCreateByMat")(C.byref(h), ncols, mat.handle, self.ops)
        return h

    def destroy_multivec(self, h: C.c_void_p, ncols: int):
        self.slot("MultiVecDestroy")(C.byref(h), ncols, self.ops)

    def upload(self, h, X: np.ndarray, start: int = 0):
        X = np.asfortranarray(X, dtype=np.float64)
        self.lib.HIP_MultiVecSetFromHost(h, start, start + X.shape[1],
                                         X.ctypes.data_as(abi.c_dbl_p), X.shape[0])

    def download(self, h, start: int, end: int, nrows: int) -> np.ndarray:
        out = np.zeros((nrows, end - start), order="F")
        self.lib.HIP_MultiVecGetToHost(h, start, end, out.ctypes.data_as(abi.c_dbl_p), nrows)
        return out


def solve(A, B, nev: int, block_size: int, nev_init: int, nev_max: int,
          tol=(1e-1, 1e-8), gap_min: float = 1e-5, shift: int = 1,
          max_iter: int = 1000, mode: str = "fp64", rand: str = "libc",
          seed: int = 0, eig: str = "auto", want_vectors: bool = False,
          print_level: int = 0, mats=None, orth: str | None = None,
          graphs: bool = True) -> SolveResult:
    """Smallest `nev` eigenpairs of A x = lambda B x on the GPU.

    mode "fp64": the reference-equivalent path, every operation in FP64.
    mode "mixed": BlockPCG in FP32 on the device (MultiLinearSolver slot);
    requires shift = 0 (the reference forbids autoshift with a user
    solver, src/ops_eig_sol_gcg.c:503).
    rand "libc": initial block from srand(seed)/rand() as the reference;
    "device": counter-based generator on the device.
    orth "mgs" (the reference's orthogonalization), "cholqr" (block
    Cholesky QR with its factorization on the device) or "cholqr_host"
    (the factorization on the host); default "mgs" for fp64 and "cholqr"
    for mixed.
    graphs: replay the mixed path's BlockPCG iteration groups as captured
    HIP graphs (False launches every kernel eagerly).
    `mats` may pass pre-uploaded (DeviceMatrix A, DeviceMatrix B).

    `seconds` runs from the first slot call after the matrix upload (the
    creation of the eigenvector block) to the return of EigenSolver, the
    solve window of the reference's harness (gcge.c:496-517);
    `solve_seconds` is EigenSolver alone.
    """
    lib = abi.load()
    if mode not in ("fp64", "mixed"):
        raise ValueError("mode must be 'fp64' or 'mixed'")
    if mode == "mixed" and shift:
        raise ValueError("mixed precision needs shift=0 (no autoshift with a user solver)")
    t = HipTable()
    ops = t.ops
    if mats is None:
        dA = DeviceMatrix(A, fp32=(mode == "mixed"))
        dB = DeviceMatrix(B, fp32=(mode == "mixed")) if B is not None else None
    else:
        dA, dB = mats
    hB = dB.handle if dB is not None else C.c_void_p()
    n = dA.n
    libc = C.CDLL(None)
    if rand == "libc":
        lib.gcge_hip_set_random_mode(abi.GCGE_HIP_RAND_LIBC, 0)
    else:
        lib.gcge_hip_set_random_mode(abi.GCGE_HIP_RAND_DEVICE, seed)
    lib.gcge_hip_synchronize()
    t_first = time.perf_counter()
    evec = t.create_multivec(nev_max, dA)
    mv_ws = (C.c_void_p * 4)()
    lib.HIPGCG_Create(dA.handle, nev_max, block_size, nev_init, mv_ws, None, None, ops)
    tolarr = (C.c_double * 2)(*tol)
    libc.srand(C.c_uint(seed))
    user_ls = 1 if mode == "mixed" else 0
    lib.HIPGCG_EigenSolverSetup(gap_min, nev_init, nev_max, block_size, tolarr, max_iter,
                                print_level, 0, user_ls, mv_ws, None, None, ops)
    lib.HIPGCG_Setparameters(gap_min, shift, ops)
    if eig != "auto":
        lib.HIPGCG_SetDenseEigensolver(ops, abi.GCGE_EIG_HOST if eig == "host"
                                       else abi.GCGE_EIG_DEVICE)
    if orth is None:
        orth = "cholqr" if mode == "mixed" else "mgs"
    lib.HIPGCG_SetOrthMethod(ops, ORTH_METHODS[orth])
    if mode == "mixed":
        lib.HIPGCG_MultiLinearSolverSetup_BlockPCG32(30, 1e-2, 1e-14, hB, 0.0, ops)
        if not graphs:
            lib.HIPGCG_BlockPCG32_SetGraphs(ops, 0)
    ev = np.zeros(nev_max)
    nconv = C.c_int(nev)
    lib.gcge_hip_synchronize()
    t0 = time.perf_counter()
    abi.slot(ops, "EigenSolver")(dA.handle, hB, ev.ctypes.data_as(abi.c_dbl_p), evec, 0,
                                 C.byref(nconv), ops)
    lib.gcge_hip_synchronize()
    t1 = time.perf_counter()
    secs = t1 - t_first
    stats = (C.c_int64 * 4)()
    times = (C.c_double * 10)()
    lib.HIPGCG_GetStats(ops, stats, times)
    cq = (C.c_int64(), C.c_int64())
    lib.HIPGCG_GetOrthStats(ops, C.byref(cq[0]), C.byref(cq[1]))
    vecs = t.download(evec, 0, nev_max, n) if want_vectors else None
    lib.HIPGCG_Destroy(mv_ws, None, None, nev_max, block_size, ops)
    t.destroy_multivec(evec, nev_max)
    lib.HIPGCG_Free(ops)
    lib.OPS_HIP_Clear(ops)
    return SolveResult(ev, nconv.value, int(stats[0]), secs, int(stats[2]), int(stats[3]),
                       dict(zip(PHASES, list(times))), vecs, cq[0].value, cq[1].value,
                       t1 - t0)


def device_operator(op, fp32: bool = False) -> DeviceMatrix:
    """A problems.GridOperator assembled on the device."""
    return DeviceMatrix.grid(problems.GRID_KINDS[op.kind], op.dim, op.points, op.n, op.coef,
                             op.delta_rel, fp32)


def solve_problem(prob, **kw) -> SolveResult:
    """solve() with the sizes and tolerances of a problems.CCSProblem, or of
    a problems.GridProblem, whose matrices are then assembled on the
    device."""
    if isinstance(prob, problems.GridProblem) and kw.get("mats") is None:
        fp32 = kw.get("mode", "fp64") == "mixed"
        kw["mats"] = (device_operator(prob.A, fp32),
                      device_operator(prob.B, fp32) if prob.B is not None else None)
    A = getattr(prob, "A", None) if not isinstance(prob, problems.GridProblem) else None
    B = getattr(prob, "B", None) if not isinstance(prob, problems.GridProblem) else None
    return solve(A, B, prob.nev, prob.block_size, prob.nev_init, prob.nev_max,
                 tol=(prob.tol_abs, prob.tol_rel), gap_min=prob.gap_min, **kw)


def solve_host_dense(A: np.ndarray, B: np.ndarray, nev: int, block_size: int, nev_init: int,
                     nev_max: int, tol=(1e-1, 1e-8), gap_min: float = 1e-5, shift: int = 1,
                     max_iter: int = 1000, seed: int = 0, orth: str = "mgs") -> SolveResult:
    """This library's driver over its host dense table (no device use).

    Dense A, B in the LAPACKMAT layout; runs the same control flow as the
    GPU path, so it can be compared with the reference driver over the
    reference's dense app on the CPU."""
    lib = abi.load()
    ops = abi.new_ops()
    lib.OPS_HOSTDENSE_Set(ops)
    n = A.shape[0]
    Af = np.asfortranarray(A, dtype=np.float64)
    Bf = np.asfortranarray(B, dtype=np.float64) if B is not None else None
    mA = abi.DenseMV(Af.ctypes.data_as(abi.c_dbl_p), n, n, n)
    mB = abi.DenseMV(Bf.ctypes.data_as(abi.c_dbl_p), n, n, n) if Bf is not None else None
    pA = C.cast(C.pointer(mA), C.c_void_p)
    pB = C.cast(C.pointer(mB), C.c_void_p) if mB is not None else C.c_void_p()
    evec = C.c_void_p()
    abi.slot(ops, "MultiVecCreateByMat")(C.byref(evec), nev_max, pA, ops)
    mv_ws = (C.c_void_p * 4)()
    lib.HIPGCG_Create(pA, nev_max, block_size, nev_init, mv_ws, None, None, ops)
    libc = C.CDLL(None)
    libc.srand(C.c_uint(seed))
    tolarr = (C.c_double * 2)(*tol)
    lib.HIPGCG_EigenSolverSetup(gap_min, nev_init, nev_max, block_size, tolarr, max_iter, 0,
                                0, 0, mv_ws, None, None, ops)
    lib.HIPGCG_Setparameters(gap_min, shift, ops)
    lib.HIPGCG_SetDenseEigensolver(ops, abi.GCGE_EIG_HOST)
    lib.HIPGCG_SetOrthMethod(ops, ORTH_METHODS[orth])
    ev = np.zeros(nev_max)
    nconv = C.c_int(nev)
    t0 = time.perf_counter()
    abi.slot(ops, "EigenSolver")(pA, pB, ev.ctypes.data_as(abi.c_dbl_p), evec, 0,
                                 C.byref(nconv), ops)
    secs = time.perf_counter() - t0
    stats = (C.c_int64 * 4)()
    times = (C.c_double * 10)()
    lib.HIPGCG_GetStats(ops, stats, times)
    cq = (C.c_int64(), C.c_int64())
    lib.HIPGCG_GetOrthStats(ops, C.byref(cq[0]), C.byref(cq[1]))
    ev_mv = C.cast(evec, C.POINTER(abi.DenseMV)).contents
    vecs = np.ctypeslib.as_array(ev_mv.data, shape=(nev_max * n,)).reshape(
        (n, nev_max), order="F").copy()
    lib.HIPGCG_Destroy(mv_ws, None, None, nev_max, block_size, ops)
    abi.slot(ops, "MultiVecDestroy")(C.byref(evec), nev_max, ops)
    lib.HIPGCG_Free(ops)
    return SolveResult(ev, nconv.value, int(stats[0]), secs, int(stats[2]), int(stats[3]),
                       dict(zip(PHASES, list(times))), vecs, cq[0].value, cq[1].value)


@dataclass
class EpsResult:
    eigr: np.ndarray
    nconv: int
    its: int
    reason: int
    evec: np.ndarray | None = None


def eps_solve(A, B, nev: int, ncv: int = 0, tol: float = 0.0, absolute: bool = False,
              max_it: int = 0, autoshift: int = -1, mode: str = "fp64",
              device_random: bool = False, seed: int = 0,
              want_vectors: bool = False) -> EpsResult:
    """gcge_eps_solve: the SLEPc plug-in's policy and call sequence
    (gcge.c:364-535) in one call."""
    lib = abi.load()
    keep = []

    def ccs(M):
        M = M.tocsc()
        arrs = (np.ascontiguousarray(M.indptr, dtype=np.int32),
                np.ascontiguousarray(M.indices, dtype=np.int32),
                np.ascontiguousarray(M.data, dtype=np.float64))
        keep.append(arrs)
        return abi.CCSMat(arrs[2].ctypes.data_as(abi.c_dbl_p), arrs[1].ctypes.data_as(abi.c_int_p),
                          arrs[0].ctypes.data_as(abi.c_int_p), M.shape[0], M.shape[1])
    mA = ccs(A)
    mB = ccs(B) if B is not None else None
    opt = abi.EpsOptions(nev, ncv, tol, abi.GCGE_EPS_ABSOLUTE if absolute else abi.GCGE_EPS_RELATIVE,
                         max_it, autoshift,
                         abi.GCGE_MODE_MIXED if mode == "mixed" else abi.GCGE_MODE_FP64, 0,
                         1 if device_random else 0, seed)
    eigr = np.zeros(nev)
    vec = np.zeros((A.shape[0], nev), order="F") if want_vectors else None
    nconv, its, reason = C.c_int(), C.c_int(), C.c_int()
    st = lib.gcge_eps_solve(C.byref(mA), C.byref(mB) if mB is not None else None, C.byref(opt),
                            eigr.ctypes.data_as(abi.c_dbl_p),
                            vec.ctypes.data_as(abi.c_dbl_p) if vec is not None else None,
                            C.byref(nconv), C.byref(its), C.byref(reason))
    if st != 0:
        raise ValueError(f"gcge_eps_solve failed with {st}")
    k = nconv.value
    return EpsResult(eigr[:k].copy(), k, its.value, reason.value,
                     vec[:, :k].copy() if vec is not None else None)

"""ctypes mirror of the GCGE operation table and the HIP backend's C ABI.

`OPS` has the field order of src/ops.h:43-152 (struct OPS_), so a pointer
to a table created by any library with that layout -- this backend or
the reference -- can be read and its slots called from Python.  Function
pointers are kept as raw addresses and cast to their C signatures on use.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgcge_hip.so")

c_int_p = C.POINTER(C.c_int)
c_dbl_p = C.POINTER(C.c_double)

# (name, kind) in declaration order; kind "f" = function pointer, "p" = data pointer
OPS_FIELDS = [
    ("Printf", "f"), ("GetWtime", "f"), ("GetOptionFromCommandLine", "f"),
    ("MatView", "f"), ("MatAxpby", "f"),
    ("VecCreateByMat", "f"), ("VecCreateByVec", "f"), ("VecDestroy", "f"),
    ("VecView", "f"), ("VecInnerProd", "f"), ("VecLocalInnerProd", "f"),
    ("VecSetRandomValue", "f"), ("VecAxpby", "f"), ("MatDotVec", "f"),
    ("MatTransDotVec", "f"),
    ("MultiVecCreateByMat", "f"), ("MultiVecCreateByVec", "f"),
    ("MultiVecCreateByMultiVec", "f"), ("MultiVecDestroy", "f"),
    ("GetVecFromMultiVec", "f"), ("RestoreVecForMultiVec", "f"),
    ("MultiVecView", "f"), ("MultiVecLocalInnerProd", "f"),
    ("MultiVecInnerProd", "f"), ("MultiVecSetRandomValue", "f"),
    ("MultiVecAxpby", "f"), ("MultiVecLinearComb", "f"), ("MatDotMultiVec", "f"),
    ("MatTransDotMultiVec", "f"), ("MultiVecQtAP", "f"),
    ("lapack_ops", "p"), ("DenseMatQtAP", "f"), ("DenseMatOrth", "f"),
    ("LinearSolver", "f"), ("linear_solver_workspace", "p"),
    ("MultiLinearSolver", "f"), ("multi_linear_solver_workspace", "p"),
    ("MultiVecOrth", "f"), ("orth_workspace", "p"),
    ("MultiGridCreate", "f"), ("MultiGridDestroy", "f"), ("VecFromItoJ", "f"),
    ("MultiVecFromItoJ", "f"),
    ("EigenSolver", "f"), ("eigen_solver_workspace", "p"),
    ("app_ops", "p"),
]


class OPS(C.Structure):
    _fields_ = [(n, C.c_void_p) for n, _ in OPS_FIELDS]


OPS_p = C.POINTER(OPS)

# C signatures of the slots called from Python
SIG = {
    "MultiVecCreateByMat": C.CFUNCTYPE(None, C.POINTER(C.c_void_p), C.c_int, C.c_void_p, OPS_p),
    "MultiVecDestroy": C.CFUNCTYPE(None, C.POINTER(C.c_void_p), C.c_int, OPS_p),
    "MultiVecSetRandomValue": C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, OPS_p),
    "MultiVecAxpby": C.CFUNCTYPE(None, C.c_double, C.c_void_p, C.c_double, C.c_void_p,
                                 c_int_p, c_int_p, OPS_p),
    "MultiVecLinearComb": C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_int, c_int_p,
                                      c_int_p, c_dbl_p, C.c_int, c_dbl_p, C.c_int, OPS_p),
    "MultiVecInnerProd": C.CFUNCTYPE(None, C.c_char, C.c_void_p, C.c_void_p, C.c_int,
                                     c_int_p, c_int_p, c_dbl_p, C.c_int, OPS_p),
    "MultiVecLocalInnerProd": C.CFUNCTYPE(None, C.c_char, C.c_void_p, C.c_void_p, C.c_int,
                                          c_int_p, c_int_p, c_dbl_p, C.c_int, OPS_p),
    "MatDotMultiVec": C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p, c_int_p, c_int_p,
                                  OPS_p),
    "MatTransDotMultiVec": C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p, c_int_p,
                                       c_int_p, OPS_p),
    "MultiVecQtAP": C.CFUNCTYPE(None, C.c_char, C.c_char, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_int, c_int_p, c_int_p, c_dbl_p, C.c_int, C.c_void_p,
                                OPS_p),
    "EigenSolver": C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, c_dbl_p, C.c_void_p, C.c_int,
                               c_int_p, OPS_p),
    "MultiLinearSolver": C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p, c_int_p,
                                     c_int_p, OPS_p),
}


def slot(ops: OPS_p, name: str):
    """The C function in slot `name` of the table `ops`, callable from Python."""
    addr = getattr(ops.contents, name)
    if not addr:
        raise RuntimeError(f"operation table slot {name} is NULL")
    return SIG[name](addr)


class CCSMat(C.Structure):
    """Layout of CCSMAT (app/app_ccs.h:20-24)."""
    _fields_ = [("data", c_dbl_p), ("i_row", c_int_p), ("j_col", c_int_p),
                ("nrows", C.c_int), ("ncols", C.c_int)]


class DenseMV(C.Structure):
    """Layout of LAPACKVEC (app/app_lapack.h:17-20)."""
    _fields_ = [("data", c_dbl_p), ("nrows", C.c_int), ("ncols", C.c_int), ("ldd", C.c_int)]


class Request(C.Structure):
    """gcge_request (include/gcge_hip.h)."""
    _fields_ = [("A", C.POINTER(CCSMat)), ("B", C.POINTER(CCSMat)), ("nev", C.c_int),
                ("block_size", C.c_int), ("nev_init", C.c_int), ("nev_max", C.c_int),
                ("max_iter", C.c_int), ("mode", C.c_int), ("shift", C.c_int),
                ("tol_abs", C.c_double), ("tol_rel", C.c_double), ("gap_min", C.c_double),
                ("seed", C.c_uint64), ("x0", c_dbl_p), ("nev_given", C.c_int)]


class Result(C.Structure):
    """gcge_result (include/gcge_hip.h)."""
    _fields_ = [("status", C.c_int), ("nev_conv", C.c_int), ("num_iter", C.c_int),
                ("concurrent", C.c_int), ("queue_seconds", C.c_double),
                ("setup_seconds", C.c_double), ("solve_seconds", C.c_double),
                ("total_seconds", C.c_double), ("eval", c_dbl_p), ("evec", c_dbl_p),
                ("pool_bytes", C.c_int64)]


class EpsOptions(C.Structure):
    """gcge_eps_options (include/gcge_hip.h)."""
    _fields_ = [("nev", C.c_int), ("ncv", C.c_int), ("tol", C.c_double), ("tol_type", C.c_int),
                ("max_it", C.c_int), ("autoshift", C.c_int), ("mode", C.c_int),
                ("print", C.c_int), ("device_random", C.c_int), ("seed", C.c_uint64)]


class GridOp(C.Structure):
    """gcge_grid_op (include/gcge_hip.h)."""
    _fields_ = [("kind", C.c_int), ("dim", C.c_int), ("points", C.c_int), ("n", C.c_int),
                ("coef", c_dbl_p), ("delta_rel", C.c_double)]


# Exported functions of libgcge_hip.so and their signatures.
EXPORTS = {
    "OPS_HIP_Set": (None, [OPS_p]),
    "OPS_HOSTDENSE_Set": (None, [OPS_p]),
    "OPS_HIP_Clear": (None, [OPS_p]),
    "gcge_host_DenseMatQtAP": (None, [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_double, c_dbl_p, C.c_int, c_dbl_p, C.c_int, c_dbl_p,
                                      C.c_int, C.c_double, c_dbl_p, C.c_int, c_dbl_p]),
    "gcge_host_DenseMatOrth": (None, [c_dbl_p, C.c_int, C.c_int, C.c_int, c_int_p, C.c_double,
                                      c_dbl_p, C.c_int, c_int_p]),
    "gcge_host_Printf": (None, [C.c_char_p]),
    "gcge_host_GetWtime": (C.c_double, []),
    "gcge_host_set_print_target": (None, [C.c_int]),
    "gcge_host_GetOptionFromCommandLine": (C.c_int, [C.c_char_p, C.c_char, C.c_void_p, C.c_int,
                                                     C.c_void_p, OPS_p]),
    "HIPCCS_Create": (C.c_void_p, [C.POINTER(CCSMat), C.c_int]),
    "HIPCCS_Destroy": (None, [C.POINTER(C.c_void_p)]),
    "HIPCCS_CreateGrid": (C.c_void_p, [C.POINTER(GridOp), C.c_int]),
    "HIPGCG_BlockPCG32_SetFused": (None, [OPS_p, C.c_int]),
    "gcge_tune_set_mode": (None, [C.c_int]),
    "gcge_tune_get_mode": (C.c_int, []),
    "gcge_tune_load": (C.c_int, [C.c_char_p]),
    "gcge_tune_save": (C.c_int, [C.c_char_p]),
    "gcge_tune_clear": (None, []),
    "gcge_tune_set": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "gcge_tune_entry": (C.c_int, [C.c_int, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p,
                                  c_dbl_p, c_dbl_p]),
    "HIPCCS_Equal": (C.c_int, [C.c_void_p, C.c_void_p]),
    "HIPCCS_Stencil": (C.c_int, [C.c_void_p]),
    "gcge_hip_set_memory_pool": (None, [C.c_int]),
    "gcge_hip_memory_pool_stats": (None, [C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "HIPCCS_Info": (None, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                           C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "HIP_MultiVecSetFromHost": (None, [C.c_void_p, C.c_int, C.c_int, c_dbl_p, C.c_int64]),
    "HIP_MultiVecGetToHost": (None, [C.c_void_p, C.c_int, C.c_int, c_dbl_p, C.c_int64]),
    "HIP_MultiVecInnerProdNoAlloc": (C.c_int, [C.c_char, C.c_void_p, C.c_void_p, C.c_int, c_int_p,
                                               c_int_p, c_dbl_p, C.c_int]),
    "HIP_MultiVecInfo": (None, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                C.POINTER(C.c_int64)]),
    "HIP_MultiVecCholQR": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                     C.c_double, C.c_double]),
    "gcge_hip_set_random_mode": (None, [C.c_int, C.c_uint64]),
    "gcge_hip_synchronize": (None, []),
    "gcge_hip_thread_release": (None, []),
    "gcge_hip_last_error": (C.c_int, [C.c_char_p, C.c_int]),
    "gcge_hip_device": (C.c_int, []),
    "gcge_hip_profile_enable": (None, [C.c_int]),
    "gcge_hip_profile_reset": (None, []),
    "gcge_hip_profile_all": (None, [C.c_int]),
    "gcge_hip_profile_read_all": (None, [C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                         C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "gcge_hip_profile_read": (None, [C.c_int, C.POINTER(C.c_int64), c_dbl_p, c_dbl_p,
                                     c_dbl_p]),
    "HIPGCG_EigenSolverSetup": (None, [C.c_double, C.c_int, C.c_int, C.c_int, c_dbl_p, C.c_int,
                                       C.c_int, C.c_int, C.c_int, C.c_void_p, c_dbl_p, c_int_p,
                                       OPS_p]),
    "HIPGCG_Setparameters": (None, [C.c_double, C.c_int, OPS_p]),
    "HIPGCG_EigenSolverSetParameters": (None, [
        C.c_int, C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_char_p, C.c_int, C.c_int,
        C.c_double, C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double,
        C.c_double, C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_double, OPS_p]),
    "HIPGCG_Create": (None, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                             C.POINTER(c_dbl_p), C.POINTER(c_int_p), OPS_p]),
    "HIPGCG_Destroy": (None, [C.c_void_p, c_dbl_p, c_int_p, C.c_int, C.c_int, OPS_p]),
    "HIPGCG_Free": (None, [OPS_p]),
    "HIPGCG_GetStats": (None, [OPS_p, C.POINTER(C.c_int64), c_dbl_p]),
    "HIPGCG_SetDenseEigensolver": (None, [OPS_p, C.c_int]),
    "HIPGCG_SetOrthMethod": (None, [OPS_p, C.c_int]),
    "gcge_host_dense_eigh": (C.c_int, [C.c_int, c_dbl_p, C.c_int, c_dbl_p, c_dbl_p, C.c_int]),
    "HIPGCG_GetOrthStats": (None, [OPS_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "HIPGCG_MultiLinearSolverSetup_BlockPCG32": (None, [C.c_int, C.c_double, C.c_double,
                                                        C.c_void_p, C.c_double, OPS_p]),
    "HIPGCG_MultiLinearSolverFree_BlockPCG32": (None, [OPS_p]),
    "HIPGCG_BlockPCG32_SetGraphs": (None, [OPS_p, C.c_int]),
    "gcge_request_footprint": (C.c_int64, [C.POINTER(Request)]),
    "gcge_service_create": (C.c_void_p, [C.c_int, C.c_int64]),
    "gcge_service_submit": (C.c_int64, [C.c_void_p, C.POINTER(Request), C.POINTER(Result)]),
    "gcge_service_wait": (C.c_int, [C.c_void_p, C.c_int64]),
    "gcge_service_stats": (None, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int),
                                  C.POINTER(C.c_int64)]),
    "gcge_service_set_batching": (None, [C.c_void_p, C.c_int]),
    "gcge_service_set_orth": (C.c_int, [C.c_void_p, C.c_int]),
    "gcge_service_batch_stats": (None, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                        c_int_p]),
    "gcge_service_destroy": (None, [C.c_void_p]),
    "gcge_eps_parameters": (C.c_int, [C.c_int, C.c_int, c_int_p, c_int_p, c_int_p]),
    "gcge_eps_solve": (C.c_int, [C.POINTER(CCSMat), C.POINTER(CCSMat), C.POINTER(EpsOptions),
                                 c_dbl_p, c_dbl_p, c_int_p, c_int_p, c_int_p]),
}

GCGE_HIP_MAT_FP32 = 1
GCGE_HIP_MAT_GENERAL = 2
GCGE_GRID_LAPLACE = 0
GCGE_GRID_MASS = 1
GCGE_GRID_DIFFUSION = 2
GCGE_GRID_DIAGONAL = 3
GCGE_HIP_RAND_LIBC = 0
GCGE_HIP_RAND_DEVICE = 1
GCGE_EIG_HOST = 0
GCGE_EIG_DEVICE = 1
GCGE_ORTH_MGS = 0
GCGE_ORTH_CHOLQR = 1
GCGE_ORTH_CHOLQR_HOST = 2
GCGE_MODE_FP64 = 0
GCGE_MODE_MIXED = 1
GCGE_EPS_RELATIVE = 0
GCGE_EPS_ABSOLUTE = 1
GCGE_EPS_CONVERGED_TOL = 1
GCGE_EPS_DIVERGED_ITS = -1
GCGE_EPS_ERR_ARG = -10
GCGE_EPS_ERR_NCV = -11
GCGE_EPS_ERR_SIZE = -12

_lib = None


def load(path: str | None = None) -> C.CDLL:
    """Loads libgcge_hip.so (built in-tree) and declares its signatures.
    Raises if the library has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build it with __graft_entry__.build() "
                           "or make -C gcge_amd/csrc")
    lib = C.CDLL(p)
    for name, (res, args) in EXPORTS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def new_ops() -> OPS_p:
    """A zeroed table (the state OPS_Create leaves, src/ops.c:26-85)."""
    buf = (C.c_byte * C.sizeof(OPS))()
    ops = C.cast(buf, OPS_p)
    ops._keep = buf  # keep the storage alive with the pointer
    return ops

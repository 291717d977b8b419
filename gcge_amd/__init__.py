"""GCG eigensolver hot path for MI355X (gfx950).

The product is libgcge_hip.so (gcge_amd/csrc): a HIP app that fills the
GCGE operation table (OPS_HIP_Set), hand-written gfx950 kernels behind it,
a reentrant GCG driver over the table, and a single-precision device
BlockPCG for the mixed-precision path.  This package is its Python host
side: ctypes bindings (abi), the solve sequence (solver) and the
synthetic problem generators (problems).
"""
from . import abi, problems  # noqa: F401

__all__ = ["abi", "problems", "solver"]

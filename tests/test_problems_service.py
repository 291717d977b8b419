"""Problem generators of the batched and served configurations, and the
service's admission footprint (host-only: no device calls)."""
import ctypes as C

import numpy as np
import scipy.linalg as sla

from gcge_amd import abi, problems as P
from gcge_amd.service import Service


def test_varcoef_reduces_to_constant_coefficients():
    A, B = P.varcoef_pair_2d(12, 0.0, np.random.default_rng(1))
    A0, B0 = P.stencil_pair(12, 2)
    assert abs(A - A0).max() == 0.0 and abs(B - B0).max() == 0.0


def test_varcoef_is_spd_with_harmonic_edges():
    rng = np.random.default_rng(7)
    A, B = P.varcoef_pair_2d(9, 0.5, rng)
    D = A.toarray()
    assert np.array_equal(D, D.T)
    assert np.linalg.eigvalsh(D).min() > 0
    # an interior edge carries the harmonic mean of its nodes' coefficients
    kappa = np.exp(0.5 * np.random.default_rng(7).standard_normal(81))
    i, j = 4 + 9 * 4, 5 + 9 * 4
    assert np.isclose(-D[i, j], 2 * kappa[i] * kappa[j] / (kappa[i] + kappa[j]), rtol=1e-14)
    # row sums: only boundary nodes keep a positive excess (the ghost terms)
    rs = D.sum(axis=1).reshape(9, 9)
    assert np.allclose(rs[1:-1, 1:-1], 0.0, atol=1e-12)
    assert np.all(rs[0, :] > 0)


def test_config4_distribution():
    for i in range(6):
        p = P.config4_problem(i)
        n = int(round(np.sqrt(p.n)))
        assert n * n == p.n and 45 <= n <= 256
        assert 8 <= p.nev <= 64
        bs, ni, nm = P.slepc_block_policy(p.nev)
        assert (p.block_size, p.nev_init, p.nev_max) == (bs, ni, nm)
    # deterministic in (seed, index)
    a, b = P.config4_problem(3), P.config4_problem(3)
    assert a.n == b.n and abs(a.A - b.A).max() == 0.0


def test_config5_stream_mix():
    s = P.config5_stream(200, 4.0)
    t = np.array([x[0] for x in s])
    assert np.all(np.diff(t) > 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert 0.15 < gaps.mean() < 0.35               # Poisson at 4 per second
    n3 = sum(1 for _, p in s if p.name.startswith("cfg5_3d"))
    assert 20 <= n3 <= 60                          # about 20 % 3D problems


def test_footprint_matches_the_workspace_formula():
    """Multivectors 8 N (2 nevMax + 5 bs) plus the matrices, within the
    stated slack (matrix padding, partials, projected matrices)."""
    prob = P.config2()
    req, _keep = Service.request(prob, mode="fp64")
    lib = abi.load()
    fp = lib.gcge_request_footprint(C.byref(req))
    mv = 8 * prob.n * (2 * prob.nev_max + 5 * prob.block_size)
    mats = 2 * (12 * prob.A.nnz + 4 * (prob.n + 1))
    assert mv + mats <= fp <= 1.05 * (mv + mats) + (64 << 20)
    req, _keep = Service.request(prob, mode="mixed")
    fpm = lib.gcge_request_footprint(C.byref(req))
    assert fpm - fp >= 4 * 4 * prob.n * prob.block_size   # FP32 BlockPCG blocks


def test_dense_oracle_for_varcoef():
    """The dense generalized eigensolver the GPU tests compare with agrees
    with the constant-coefficient closed form (a check of the harness)."""
    A, B = P.varcoef_pair_2d(10, 0.0, np.random.default_rng(0))
    w = sla.eigh(A.toarray(), B.toarray(), eigvals_only=True, subset_by_index=[0, 9])
    assert np.allclose(w, P.analytic_eigenvalues(10, 2, 10), rtol=1e-12)


def test_varcoef27_structure():
    A, B = P.varcoef27_pair(7, 1.0, np.random.default_rng(5))
    D = A.toarray()
    assert np.array_equal(D, D.T)                       # bit for bit
    assert A.nnz == (3 * 7 - 2) ** 3 and A.has_sorted_indices
    off = D - np.diag(np.diag(D))
    assert np.all(np.diag(D) > -off.sum(axis=1))        # strictly dominant
    # an interior face edge: harmonic mean, weight 1; a corner edge: 1/3
    kappa = np.exp(np.random.default_rng(5).standard_normal(343))
    c = 3 + 7 * 3 + 49 * 3
    for d, g in ((1, 1.0), (1 + 7 + 49, 1.0 / 3.0)):
        h = 2 * kappa[c] * kappa[c + d] / (kappa[c] + kappa[c + d])
        assert np.isclose(-D[c, c + d], g * h, rtol=1e-14)
    bd = B.diagonal()
    assert B.nnz == 343 and bd.min() >= 0.5 and bd.max() <= 1.5


def test_eps_parameter_policy():
    """gcge_eps_parameters follows the plug-in's policy (gcge.c:394-431),
    the one problems.slepc_block_policy restates, and its ncv checks."""
    lib = abi.load()
    out = [C.c_int() for _ in range(3)]
    for nev in (1, 6, 7, 10, 19, 20, 100, 999, 1000, 3999, 4000):
        assert lib.gcge_eps_parameters(nev, 0, *[C.byref(x) for x in out]) == 0
        assert tuple(x.value for x in out) == P.slepc_block_policy(nev)
    # ncv: nev <= 6 needs ncv == 2 nev; otherwise nevInit <= ncv <= nev + nevInit
    assert lib.gcge_eps_parameters(4, 8, *[C.byref(x) for x in out]) == 0
    assert lib.gcge_eps_parameters(4, 9, *[C.byref(x) for x in out]) == abi.GCGE_EPS_ERR_NCV
    assert lib.gcge_eps_parameters(50, 45, *[C.byref(x) for x in out]) == 0
    assert out[2].value == 45
    assert lib.gcge_eps_parameters(50, 20, *[C.byref(x) for x in out]) == abi.GCGE_EPS_ERR_NCV
    assert lib.gcge_eps_parameters(50, 81, *[C.byref(x) for x in out]) == abi.GCGE_EPS_ERR_NCV
    assert lib.gcge_eps_parameters(0, 0, *[C.byref(x) for x in out]) == abi.GCGE_EPS_ERR_ARG


def _same_ccs(X, Y):
    X, Y = X.tocsc(), Y.tocsc()
    return (np.array_equal(X.indptr, Y.indptr) and np.array_equal(X.indices, Y.indices)
            and np.array_equal(X.data.view(np.int64), Y.data.view(np.int64)))


def test_grid_forms_reproduce_the_generators():
    """The grid descriptions draw the same coefficients and, through
    to_ccs(), give the generators' matrices bit for bit (the device
    assembly is checked against these)."""
    for i in (0, 9):
        a, g = P.config4_problem(i), P.config4_grid(i)
        c = g.to_ccs()
        assert _same_ccs(a.A, c.A) and _same_ccs(a.B, c.B)
        assert (a.nev, a.block_size, a.nev_init, a.nev_max) == \
            (g.nev, g.block_size, g.nev_init, g.nev_max)
    a, g = P.config3(10), P.config3_grid(10)
    c = g.to_ccs()
    assert _same_ccs(a.A, c.A) and _same_ccs(a.B, c.B) and g.n == 1000
    a = P.stencil_problem(9, 3, 100, 120, 220, 220)
    c = P.stencil_grid_problem(9, 3, 100, 120, 220, 220, "s").to_ccs()
    assert _same_ccs(a.A, c.A) and _same_ccs(a.B, c.B)


def test_create_grid_rejects_unsupported_operators():
    """Rejected before any device call: unsupported stencils, missing
    coefficients, empty grids and orders beyond int32 column indices."""
    lib = abi.load()
    kappa = np.ones(27)
    p = kappa.ctypes.data_as(abi.c_dbl_p)
    bad = [abi.GridOp(abi.GCGE_GRID_LAPLACE, 3, 27, 3, None, 0.0),
           abi.GridOp(abi.GCGE_GRID_MASS, 2, 7, 3, None, 0.0),
           abi.GridOp(abi.GCGE_GRID_DIFFUSION, 3, 7, 3, p, 0.0),
           abi.GridOp(abi.GCGE_GRID_DIFFUSION, 3, 27, 3, None, 0.0),
           abi.GridOp(abi.GCGE_GRID_DIAGONAL, 3, 0, 3, None, 0.0),
           abi.GridOp(abi.GCGE_GRID_LAPLACE, 2, 5, 0, None, 0.0),
           abi.GridOp(abi.GCGE_GRID_LAPLACE, 4, 5, 3, None, 0.0),
           abi.GridOp(7, 2, 5, 3, None, 0.0),
           abi.GridOp(abi.GCGE_GRID_LAPLACE, 3, 7, 1291, None, 0.0)]   # 1291^3 > 2^31
    for op in bad:
        assert not lib.HIPCCS_CreateGrid(C.byref(op), 0), (op.kind, op.dim, op.points, op.n)
    assert not lib.HIPCCS_CreateGrid(None, 0)

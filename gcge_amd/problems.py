"""Synthetic generalized eigenproblems A x = lambda B x, stored CCS.

The constant-coefficient stencil problems have closed-form spectra; they
give the eigensolver's results an exact answer to meet.  All matrices are symmetric,
so their CCS arrays are also their CSR arrays (the HIP backend reads them
as a row gather).

Index arrays are int32 and values float64, the CCSMAT layout of
app/app_ccs.h:20-24.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp


@dataclass
class CCSProblem:
    """A pair of CCS matrices plus solver settings for one problem."""
    name: str
    A: sp.csc_matrix
    B: sp.csc_matrix | None
    nev: int
    block_size: int
    nev_init: int
    nev_max: int
    tol_abs: float = 1e-1
    tol_rel: float = 1e-8
    gap_min: float = 1e-5

    @property
    def n(self) -> int:
        return self.A.shape[0]


def _path_adjacency(n: int) -> sp.csr_matrix:
    e = np.ones(n - 1)
    return sp.diags([e, e], [-1, 1], shape=(n, n), format="csr")


def grid_adjacency(n: int, dim: int) -> sp.csr_matrix:
    """Adjacency of the n^dim interior grid with Dirichlet boundary."""
    T = _path_adjacency(n)
    I = sp.identity(n, format="csr")
    if dim == 2:
        return (sp.kron(I, T) + sp.kron(T, I)).tocsr()
    if dim == 3:
        II = sp.identity(n * n, format="csr")
        return (sp.kron(II, T) + sp.kron(I, sp.kron(T, I)) + sp.kron(T, II)).tocsr()
    raise ValueError("dim must be 2 or 3")


def _canonical_csc(M: sp.spmatrix) -> sp.csc_matrix:
    M = sp.csc_matrix(M)
    M.sum_duplicates()
    M.sort_indices()
    M.indptr = M.indptr.astype(np.int32)
    M.indices = M.indices.astype(np.int32)
    M.data = M.data.astype(np.float64)
    return M


def stencil_pair(n: int, dim: int):
    """A = 2d I - Adj, B = (4d I + Adj) / (6d) ... for d = 2: (8I+Adj)/12,
    for d = 3: (12I+Adj)/18."""
    Adj = grid_adjacency(n, dim)
    N = n ** dim
    I = sp.identity(N, format="csr")
    if dim == 2:
        A = 4.0 * I - Adj
        B = (8.0 * I + Adj) / 12.0
    else:
        A = 6.0 * I - Adj
        B = (12.0 * I + Adj) / 18.0
    return _canonical_csc(A), _canonical_csc(B)


def analytic_eigenvalues(n: int, dim: int, count: int) -> np.ndarray:
    """Smallest `count` eigenvalues of stencil_pair(n, dim), ascending.

    With mu = sum_k 2 cos(i_k pi / (n+1)) the eigenvalues of Adj, A and B
    share eigenvectors, so lambda = (2d - mu) * s / (s' + mu) with
    (s, s') = (12, 8) in 2D and (18, 12) in 3D.
    """
    c = 2.0 * np.cos(np.arange(1, n + 1) * math.pi / (n + 1))
    if dim == 2:
        mu = (c[:, None] + c[None, :]).ravel()
        lam = (4.0 - mu) * 12.0 / (8.0 + mu)
    else:
        mu = (c[:, None, None] + c[None, :, None] + c[None, None, :]).ravel()
        lam = (6.0 - mu) * 18.0 / (12.0 + mu)
    lam.sort()
    return lam[:count]


def slepc_block_policy(nev: int):
    """Block size / nevInit / nevMax chosen as the SLEPc plugin does
    (gcge.c:394-431)."""
    if nev > 6:
        if nev < 20:
            bs = nev // 3
        elif nev < 1000:
            bs = nev // 5
        elif nev < 4000:
            bs = 100
        else:
            bs = 200
        nev_init = 3 * bs
        nev_max = nev_init + nev
    else:
        bs = nev
        nev_init = 2 * nev
        nev_max = nev_init
    return bs, nev_init, nev_max


def stencil_problem(n: int, dim: int, nev: int, block_size: int | None = None,
                    nev_init: int | None = None, nev_max: int | None = None,
                    name: str | None = None) -> CCSProblem:
    A, B = stencil_pair(n, dim)
    if block_size is None:
        block_size, nev_init, nev_max = slepc_block_policy(nev)
    return CCSProblem(name or f"stencil{dim}d_n{n}", A, B, nev, block_size,
                      nev_init, nev_max)


def config1() -> CCSProblem:
    """2D 128x128 interior, nev = 20, SLEPc block policy (bs 4, 12, 32)."""
    return stencil_problem(128, 2, 20, name="cfg1_2d_128")


def config2(block_size: int = 120) -> CCSProblem:
    """3D 128^3 interior, nev = 100, block size 120 with nevInit = nevMax
    = 220 (the largest nevMax the driver's workspace asserts allow at
    nevInit < 3 bs, src/ops_eig_sol_gcg.c:1280-1284)."""
    if block_size == 120:
        return stencil_problem(128, 3, 100, 120, 220, 220, name="cfg2_3d_128")
    bs, ni, nm = slepc_block_policy(100)
    return stencil_problem(128, 3, 100, bs, ni, nm, name="cfg2_3d_128_slepc")


SEED_BASE = 20261202


def varcoef_pair_2d(n: int, sigma: float, rng: np.random.Generator):
    """Variable-coefficient diffusion on an n x n interior grid, Dirichlet
    boundary, 5-point stencil, index x + n*y.

    Node coefficients kappa = exp(sigma * z), z ~ N(0, 1); the weight of
    the edge between neighbours i, j is the harmonic mean
    2 kappa_i kappa_j / (kappa_i + kappa_j); A_ij = -w_ij and A_ii is the
    sum of the weights at i, where a missing (boundary) neighbour adds
    kappa_i.  B = (8 I + Adj) / 12 as in the constant-coefficient pair, to
    which kappa = 1 reduces.  A is symmetric and strictly diagonally
    dominant, hence SPD."""
    return varcoef_pair_2d_from(n, varcoef2d_coefficients(n, sigma, rng))


def varcoef2d_coefficients(n: int, sigma: float, rng: np.random.Generator) -> np.ndarray:
    """The node coefficients of varcoef_pair_2d: exp(sigma * z), z ~ N(0, 1)."""
    return np.exp(sigma * rng.standard_normal(n * n))


def varcoef_pair_2d_from(n: int, kappa: np.ndarray):
    """varcoef_pair_2d for given node coefficients."""
    N = n * n
    idx = np.arange(N).reshape(n, n)          # [y, x]
    rows, cols, vals = [], [], []
    diag = np.zeros(N)
    for a, b in ((idx[:, :-1].ravel(), idx[:, 1:].ravel()),    # x-neighbours
                 (idx[:-1, :].ravel(), idx[1:, :].ravel())):   # y-neighbours
        w = 2.0 * kappa[a] * kappa[b] / (kappa[a] + kappa[b])
        rows += [a, b]
        cols += [b, a]
        vals += [-w, -w]
        np.add.at(diag, a, w)
        np.add.at(diag, b, w)
    nb = np.zeros((n, n))
    nb[:, 0] += 1
    nb[:, -1] += 1
    nb[0, :] += 1
    nb[-1, :] += 1
    diag += nb.ravel() * kappa
    rows.append(np.arange(N))
    cols.append(np.arange(N))
    vals.append(diag)
    A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(N, N))
    B = (8.0 * sp.identity(N) + grid_adjacency(n, 2)) / 12.0
    return _canonical_csc(A), _canonical_csc(B)


def config4_problem(i: int, seed: int = SEED_BASE) -> CCSProblem:
    """Problem i of the batched set: N ~ U[2048, 65536], grid side
    round(sqrt(N)), variable coefficients lognormal(0, 0.5),
    nev ~ U{8..64}, SLEPc block policy, tol 1e-8."""
    rng = np.random.default_rng([seed, i])
    N = int(rng.integers(2048, 65536 + 1))
    n = int(round(math.sqrt(N)))
    nev = int(rng.integers(8, 64 + 1))
    A, B = varcoef_pair_2d(n, 0.5, rng)
    bs, ni, nm = slepc_block_policy(nev)
    return CCSProblem(f"cfg4_{i}_2d{n}", A, B, nev, bs, ni, nm)


def config4_batch(count: int = 512, seed: int = SEED_BASE) -> list:
    return [config4_problem(i, seed) for i in range(count)]


def config5_stream(count: int, qps: float, seed: int = SEED_BASE):
    """A served request stream: Poisson arrivals at `qps`; 80 % drawn as
    in config4_problem, 20 % 3D 7-point constant-coefficient problems with
    n ~ U[48, 96] and nev ~ U{32..128}.  Returns [(arrival_s, problem)]."""
    rng = np.random.default_rng([seed, 5])
    t = 0.0
    out = []
    for k in range(count):
        t += float(rng.exponential(1.0 / qps))
        if rng.random() < 0.8:
            prob = config4_problem(int(rng.integers(0, 1 << 30)), seed)
        else:
            n = int(rng.integers(48, 96 + 1))
            nev = int(rng.integers(32, 128 + 1))
            prob = stencil_problem(n, 3, nev, name=f"cfg5_3d{n}_nev{nev}")
        out.append((t, prob))
    return out


def varcoef27_pair(n: int, sigma: float, rng: np.random.Generator,
                   delta_rel: float = 1e-8):
    """Variable-coefficient diffusion on an n^3 interior grid with the
    27-point stencil, Dirichlet boundary, index x + n*y + n^2*z.

    Node coefficients kappa = exp(sigma * z), z ~ N(0, 1).  The neighbour
    at offset d (|d|_inf = 1) couples with weight g(d) * 2 kappa_i kappa_j /
    (kappa_i + kappa_j), g(d) = 1 / |d|^2 (faces 1, edges 1/2, corners
    1/3); A_ij = -weight, A_ii = the sum of the weights at i plus, for
    every neighbour outside the grid, g(d) * kappa_i, plus delta = delta_rel
    * max diagonal.  B = diag(U[0.5, 1.5]).  A is symmetric bit for bit and
    strictly diagonally dominant, hence SPD; nnz(A) = (3n - 2)^3.

    Built column by column in CCS directly: the 27 offsets in (dz, dy, dx)
    order are the columns of a row in ascending order."""
    kappa, bdiag = varcoef27_coefficients(n, sigma, rng)
    return varcoef27_from(n, kappa, delta_rel), diagonal_ccs(bdiag)


def varcoef27_coefficients(n: int, sigma: float, rng: np.random.Generator):
    """The node coefficients and the diagonal of B drawn by varcoef27_pair."""
    N = n ** 3
    kappa = np.exp(sigma * rng.standard_normal(N))
    bdiag = rng.uniform(0.5, 1.5, N)
    return kappa, bdiag


def diagonal_ccs(d: np.ndarray) -> sp.csc_matrix:
    B = sp.diags(d).tocsc()
    B.sort_indices()
    return B


def varcoef27_from(n: int, kappa: np.ndarray, delta_rel: float = 1e-8) -> sp.csc_matrix:
    """The 27-point operator of varcoef27_pair for given node coefficients."""
    N = n ** 3
    iz, iy, ix = (a.ravel() for a in np.indices((n, n, n), dtype=np.int32))
    offs = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]

    def valid(dz, dy, dx):
        m = np.ones(N, dtype=bool)
        for a, d in ((iz, dz), (iy, dy), (ix, dx)):
            if d < 0:
                m &= a > 0
            elif d > 0:
                m &= a < n - 1
        return m

    counts = np.zeros(N, dtype=np.int64)
    for o in offs:
        counts += valid(*o)
    indptr = np.zeros(N + 1, dtype=np.int64)
    np.cumsum(counts, out=indptr[1:])
    nnz = int(indptr[-1])
    indices = np.empty(nnz, dtype=np.int32)
    data = np.empty(nnz, dtype=np.float64)
    diag = np.zeros(N)
    pos = indptr[:-1].copy()
    pos_diag = None
    for dz, dy, dx in offs:
        m = valid(dz, dy, dx)
        r = np.flatnonzero(m)
        if (dz, dy, dx) == (0, 0, 0):
            pos_diag = pos.copy()
            pos += 1
            continue
        g = 1.0 / (dz * dz + dy * dy + dx * dx)
        c = r + (dz * n * n + dy * n + dx)
        w = g * (2.0 * kappa[r] * kappa[c] / (kappa[r] + kappa[c]))
        indices[pos[r]] = c
        data[pos[r]] = -w
        pos[r] += 1
        diag[r] += w
        out = ~m
        diag[out] += g * kappa[out]
    diag += delta_rel * diag.max()
    indices[pos_diag] = np.arange(N, dtype=np.int32)
    data[pos_diag] = diag
    return sp.csc_matrix((data, indices, indptr.astype(np.int32)), shape=(N, N))


def config3(n: int = 160, nev: int = 200, seed: int = SEED_BASE) -> CCSProblem:
    """27-point variable-coefficient diffusion (lognormal sigma = 1) on n^3,
    diagonal B, nev = 200 with the SLEPc block policy (bs 40, nevInit 120,
    nevMax 320), tol 1e-8."""
    rng = np.random.default_rng([seed, 3])
    A, B = varcoef27_pair(n, 1.0, rng)
    bs, ni, nm = slepc_block_policy(nev)
    return CCSProblem(f"cfg3_27pt_{n}", A, B, nev, bs, ni, nm)


# ---------------------------------------------------------------- grid forms
# The same problems described by their grid and node coefficients, for
# assembly on the device (HIPCCS_CreateGrid); to_ccs() gives the host CCS
# the generators above build, which the device assembly reproduces bit for
# bit.

GRID_KINDS = {"laplace": 0, "mass": 1, "diffusion": 2, "diagonal": 3}


@dataclass
class GridOperator:
    kind: str            # "laplace", "mass", "diffusion" or "diagonal"
    dim: int
    points: int          # 5 (2D), 7 or 27 (3D); ignored for "diagonal"
    n: int               # interior points per side
    coef: np.ndarray | None = None
    delta_rel: float = 0.0

    @property
    def order(self) -> int:
        return self.n ** self.dim

    def to_ccs(self) -> sp.csc_matrix:
        if self.kind in ("laplace", "mass"):
            return stencil_pair(self.n, self.dim)[0 if self.kind == "laplace" else 1]
        if self.kind == "diffusion" and self.dim == 2:
            return varcoef_pair_2d_from(self.n, self.coef)[0]
        if self.kind == "diffusion" and self.dim == 3 and self.points == 27:
            return varcoef27_from(self.n, self.coef, self.delta_rel)
        if self.kind == "diagonal":
            return diagonal_ccs(self.coef)
        raise ValueError(f"no host form for {self}")


@dataclass
class GridProblem:
    name: str
    A: GridOperator
    B: GridOperator | None
    nev: int
    block_size: int
    nev_init: int
    nev_max: int
    tol_abs: float = 1e-1
    tol_rel: float = 1e-8
    gap_min: float = 1e-5

    @property
    def n(self) -> int:
        return self.A.order

    def to_ccs(self) -> CCSProblem:
        return CCSProblem(self.name, self.A.to_ccs(),
                          self.B.to_ccs() if self.B is not None else None, self.nev,
                          self.block_size, self.nev_init, self.nev_max, self.tol_abs,
                          self.tol_rel, self.gap_min)


def stencil_grid_problem(n: int, dim: int, nev: int, block_size: int, nev_init: int,
                         nev_max: int, name: str) -> GridProblem:
    pts = 5 if dim == 2 else 7
    return GridProblem(name, GridOperator("laplace", dim, pts, n),
                       GridOperator("mass", dim, pts, n), nev, block_size, nev_init, nev_max)


def config1_grid() -> GridProblem:
    """config1() as a grid problem."""
    bs, ni, nm = slepc_block_policy(20)
    return stencil_grid_problem(128, 2, 20, bs, ni, nm, "cfg1_2d_128")


def config2_grid() -> GridProblem:
    """config2() as a grid problem."""
    return stencil_grid_problem(128, 3, 100, 120, 220, 220, "cfg2_3d_128")


def config3_grid(n: int = 160, nev: int = 200, seed: int = SEED_BASE) -> GridProblem:
    """config3() as a grid problem (the same random draws)."""
    rng = np.random.default_rng([seed, 3])
    kappa, bdiag = varcoef27_coefficients(n, 1.0, rng)
    bs, ni, nm = slepc_block_policy(nev)
    return GridProblem(f"cfg3_27pt_{n}", GridOperator("diffusion", 3, 27, n, kappa, 1e-8),
                       GridOperator("diagonal", 3, 0, n, bdiag), nev, bs, ni, nm)


def config4_grid(i: int, seed: int = SEED_BASE) -> GridProblem:
    """config4_problem(i) as a grid problem (the same random draws)."""
    rng = np.random.default_rng([seed, i])
    N = int(rng.integers(2048, 65536 + 1))
    n = int(round(math.sqrt(N)))
    nev = int(rng.integers(8, 64 + 1))
    kappa = varcoef2d_coefficients(n, 0.5, rng)
    bs, ni, nm = slepc_block_policy(nev)
    return GridProblem(f"cfg4_{i}_2d{n}", GridOperator("diffusion", 2, 5, n, kappa),
                       GridOperator("mass", 2, 5, n), nev, bs, ni, nm)

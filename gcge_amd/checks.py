"""Accuracy measures of computed eigenpairs beyond residuals: principal
angles between eigenvector sets in the B inner product, taken over whole
eigenvalue clusters (a degenerate cluster fixes a subspace, not its
vectors)."""
from __future__ import annotations

import numpy as np


def b_orthonormalize(X: np.ndarray, B) -> np.ndarray:
    """A basis of span(X) with Q^T B Q = I (Cholesky of X^T B X)."""
    BX = X if B is None else B @ X
    L = np.linalg.cholesky(X.T @ BX)
    return np.linalg.solve(L, X.T).T


def max_subspace_sine(X1: np.ndarray, X2: np.ndarray, B=None) -> float:
    """sin of the largest principal angle between span(X1) and span(X2)
    in the B inner product: the B-norm of the part of Q2 outside span(Q1),
    computed directly (accurate for small angles, unlike arccos)."""
    Q1, Q2 = b_orthonormalize(X1, B), b_orthonormalize(X2, B)
    BQ2 = Q2 if B is None else B @ Q2
    P = Q2 - Q1 @ (Q1.T @ BQ2)
    BP = P if B is None else B @ P
    w = np.linalg.eigvalsh(P.T @ BP)
    return float(np.sqrt(max(w.max(), 0.0)))


def cluster_cut(evals: np.ndarray, kmax: int, rel_gap: float = 1e-6) -> int:
    """The largest k <= kmax such that evals[k - 1] and evals[k] are
    separated by more than rel_gap (relative): the first k pairs are then
    whole clusters.  evals ascending; k = kmax when no later value is
    known."""
    evals = np.asarray(evals, dtype=float)
    for k in range(min(kmax, len(evals)), 0, -1):
        if k == len(evals) or evals[k] - evals[k - 1] > rel_gap * abs(evals[k - 1]):
            return k
    return 0

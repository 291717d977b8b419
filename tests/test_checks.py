"""Principal-angle and cluster helpers (host only)."""
import numpy as np
import scipy.sparse as sp

from gcge_amd import checks


def test_known_angle_in_b_inner_product():
    rng = np.random.default_rng(3)
    n = 60
    B = sp.diags(rng.uniform(0.5, 2.0, n)).tocsr()
    Q = checks.b_orthonormalize(rng.standard_normal((n, 4)), B)
    theta = 1e-5
    X2 = Q.copy()
    # rotate the last basis vector toward a B-orthogonal direction
    v = rng.standard_normal(n)
    v -= Q @ (Q.T @ (B @ v))
    v /= np.sqrt(v @ (B @ v))
    X2[:, 3] = np.cos(theta) * Q[:, 3] + np.sin(theta) * v
    s = checks.max_subspace_sine(Q, X2 @ np.array([[2.0, 0, 0, 0], [1, 1, 0, 0], [0, 0, 3, 0],
                                                   [0, 0, 1, 1]]), B)
    assert abs(s - np.sin(theta)) <= 1e-12
    assert checks.max_subspace_sine(Q, Q @ np.diag([1.0, 2, 3, 4]), B) <= 1e-14


def test_cluster_cut():
    ev = np.array([1.0, 2.0, 2.0, 2.0 + 1e-9, 3.0, 3.0])
    assert checks.cluster_cut(ev, 6) == 6
    assert checks.cluster_cut(ev, 5) == 4
    assert checks.cluster_cut(ev, 3) == 1
    assert checks.cluster_cut(ev, 1) == 1

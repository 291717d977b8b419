"""Grid operators assembled on the device (HIPCCS_CreateGrid) against the
same operators built on the host by the problem generators and uploaded
(HIPCCS_Create): the device arrays must be identical bit for bit, edge
sizes included, with and without the FP32 copy; solves through the
assembled handles give the same eigenvalues as through the uploaded ones."""
import numpy as np
import pytest

from gcge_amd import problems as P
from gcge_amd import solver as S

pytestmark = pytest.mark.gpu


def ops():
    rng = np.random.default_rng([P.SEED_BASE, 77])
    out = []
    for n in (1, 2, 3, 37):
        out.append(P.GridOperator("laplace", 2, 5, n))
        out.append(P.GridOperator("mass", 2, 5, n))
        out.append(P.GridOperator("diffusion", 2, 5, n, P.varcoef2d_coefficients(n, 0.5, rng)))
    for n in (1, 2, 3, 13):
        out.append(P.GridOperator("laplace", 3, 7, n))
        out.append(P.GridOperator("mass", 3, 7, n))
        kappa, bdiag = P.varcoef27_coefficients(n, 1.0, rng)
        out.append(P.GridOperator("diffusion", 3, 27, n, kappa, 1e-8))
        out.append(P.GridOperator("diagonal", 3, 0, n, bdiag))
    # lognormal spread of several decades and no diagonal shift
    kappa, _ = P.varcoef27_coefficients(9, 3.0, rng)
    out.append(P.GridOperator("diffusion", 3, 27, 9, kappa, 0.0))
    return out


OPS = ops()


def ident(op):
    return f"{op.kind}{op.dim}d{op.points}p_n{op.n}_d{op.delta_rel:g}"


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("op", OPS, ids=[ident(o) for o in OPS])
def test_assembled_equals_uploaded(op, fp32):
    print(f"assemble {ident(op)} fp32={fp32}", flush=True)
    dev = S.device_operator(op, fp32=fp32)
    host = S.DeviceMatrix(op.to_ccs(), fp32=fp32)
    assert dev.info() == host.info()
    assert dev.equal(host)


def test_assembled_differs_when_coefficients_differ():
    """HIPCCS_Equal is not vacuous: one coefficient changed by one ulp."""
    rng = np.random.default_rng([P.SEED_BASE, 78])
    kappa, _ = P.varcoef27_coefficients(6, 1.0, rng)
    op = P.GridOperator("diffusion", 3, 27, 6, kappa, 1e-8)
    k2 = kappa.copy()
    k2[100] = np.nextafter(k2[100], np.inf)
    a = S.device_operator(op)
    b = S.device_operator(P.GridOperator("diffusion", 3, 27, 6, k2, 1e-8))
    assert a.info() == b.info()
    assert not a.equal(b)


def test_config3_geometry_at_scale():
    """The 27-point operator at 64^3 (262,144 rows, 15.6 M nonzeros)."""
    g = P.config3_grid(64)
    print("assemble cfg3 64^3", flush=True)
    for op in (g.A, g.B):
        dev = S.device_operator(op, fp32=True)
        host = S.DeviceMatrix(op.to_ccs(), fp32=True)
        assert dev.equal(host)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_solve_through_assembled_handles(mode):
    """Same inputs, same kernels: the eigenvalues agree bit for bit."""
    g = P.config4_grid(3)
    ccs = g.to_ccs()
    print(f"solve {g.name} N={g.n} {mode}", flush=True)
    kw = dict(shift=1 if mode == "fp64" else 0, mode=mode, rand="device", seed=5)
    r1 = S.solve_problem(g, **kw)
    r2 = S.solve_problem(ccs, **kw)
    assert r1.nev_conv >= g.nev
    assert r1.nev_conv == r2.nev_conv and r1.num_iter == r2.num_iter
    k = r1.nev_conv
    assert np.array_equal(r1.eval[:k], r2.eval[:k])

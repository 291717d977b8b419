"""The solve service on the GPU: concurrent requests on separate worker
streams give the same results as the same requests solved alone, match
dense generalized eigensolves (scipy.linalg.eigh) and meet the residual
bound; admission respects the HBM budget."""
import numpy as np
import pytest
import scipy.linalg as sla

import gcge_oracle as O
from gcge_amd import problems as P
from gcge_amd.service import Service

pytestmark = pytest.mark.gpu


def small_varcoef(i):
    """Variable-coefficient 2D problems of uneven sizes (N 900..2500)."""
    rng = np.random.default_rng([P.SEED_BASE, 1000 + i])
    n = int(rng.integers(30, 51))
    nev = int(rng.integers(8, 21))
    A, B = P.varcoef_pair_2d(n, 0.5, rng)
    bs, ni, nm = P.slepc_block_policy(nev)
    return P.CCSProblem(f"svc{i}_2d{n}", A, B, nev, bs, ni, nm)


PROBS = [small_varcoef(i) for i in range(6)]
DENSE = {}


def dense_pairs(p):
    if p.name not in DENSE:
        DENSE[p.name] = sla.eigh(p.A.toarray(), p.B.toarray(), eigvals_only=True,
                                 subset_by_index=[0, p.nev_max - 1])
    return DENSE[p.name]


def check(p, r):
    k = r.nev_conv
    assert k >= p.nev
    ref = dense_pairs(p)[:k]
    assert np.max(np.abs(r.eval - ref) / ref) <= 1e-9
    rr = O.relative_residuals(p.A, p.B, r.eval, r.evec)
    assert rr.max() <= 1.01e-8


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_same_alone_and_in_the_mix(hip, mode):
    with Service(workers=1) as svc:
        alone = [svc.wait(svc.submit(p, mode=mode, seed=11, want_vectors=True)) for p in PROBS]
    with Service(workers=3) as svc:
        tickets = [svc.submit(p, mode=mode, seed=11, want_vectors=True) for p in PROBS]
        mix = [svc.wait(t) for t in tickets]
        st = svc.stats()
    assert st["completed"] == len(PROBS) and st["max_concurrent"] >= 2
    assert max(r.concurrent for r in mix) >= 2
    for p, a, m in zip(PROBS, alone, mix):
        check(p, a)
        check(p, m)
        # same request, same seed: the kernels reduce in a fixed order, so
        # concurrency does not change the arithmetic
        assert a.nev_conv == m.nev_conv and a.num_iter == m.num_iter
        assert np.array_equal(a.eval, m.eval), p.name


def test_device_factorization_option(hip):
    """Mixed requests with their Cholesky QR factorizations on the device
    (gcge_service_set_orth(GCGE_ORTH_CHOLQR)) instead of the default host:
    the same bounds, the same requests alone and in the mix bit for bit,
    and eigenvalues within 1e-10 of the host-factorized runs."""
    with Service(workers=1) as svc:
        host = [svc.wait(svc.submit(p, mode="mixed", seed=11)) for p in PROBS[:4]]
    with Service(workers=1, orth="cholqr") as svc:
        alone = [svc.wait(svc.submit(p, mode="mixed", seed=11, want_vectors=True))
                 for p in PROBS[:4]]
    with Service(workers=3, orth="cholqr") as svc:
        mix = [svc.wait(t) for t in [svc.submit(p, mode="mixed", seed=11, want_vectors=True)
                                     for p in PROBS[:4]]]
    for p, h, a, m in zip(PROBS, host, alone, mix):
        check(p, a)
        assert np.array_equal(a.eval, m.eval), p.name
        k = min(h.nev_conv, a.nev_conv)
        assert np.max(np.abs(a.eval[:k] - h.eval[:k]) / np.abs(h.eval[:k])) <= 1e-10
    with Service(workers=1) as svc:
        assert svc.lib.gcge_service_set_orth(svc.h, 0) == -1


def test_admission_by_footprint(hip):
    fps = [Service.footprint(p) for p in PROBS[:4]]
    budget = int(1.5 * max(fps))          # room for one request at a time
    with Service(workers=4, hbm_budget=budget) as svc:
        tickets = [svc.submit(p, seed=3) for p in PROBS[:4]]
        res = [svc.wait(t) for t in tickets]
        st = svc.stats()
        with pytest.raises(RuntimeError):
            svc.submit(P.config2())       # can never fit this budget
    assert st["max_concurrent"] == 1 and all(r.concurrent == 1 for r in res)
    assert st["hbm_in_use"] == 0


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_warm_start_from_a_nearby_problem(hip, mode):
    """A request warm-started (nevGiven) from the eigenvectors of a nearby
    problem -- the same grid with coefficients perturbed by 1 % -- converges
    in fewer iterations than cold, to the same eigenvalues and residual
    bound as a dense solve."""
    rng = np.random.default_rng([P.SEED_BASE, 2000])
    n = 40
    kappa = P.varcoef2d_coefficients(n, 0.5, rng)
    A0, B = P.varcoef_pair_2d_from(n, kappa)
    A1, _ = P.varcoef_pair_2d_from(n, kappa * (1.0 + 0.01 * rng.standard_normal(n * n)))
    bs, ni, nm = P.slepc_block_policy(15)
    p0 = P.CCSProblem("warm0", A0, B, 15, bs, ni, nm)
    p1 = P.CCSProblem("warm1", A1, B, 15, bs, ni, nm)
    with Service(workers=2) as svc:
        base = svc.wait(svc.submit(p0, mode=mode, seed=21, want_vectors=True))
        cold = svc.wait(svc.submit(p1, mode=mode, seed=21, want_vectors=True))
        warm = svc.wait(svc.submit(p1, mode=mode, seed=21, want_vectors=True,
                                   x0=base.evec[:, :ni]))
        with pytest.raises(ValueError):
            svc.submit(p1, x0=np.ones((p1.n, ni + 1)))   # more than nev_init
    check(p1, cold)
    check(p1, warm)
    assert warm.num_iter < cold.num_iter, (warm.num_iter, cold.num_iter)


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_worker_pool_stops_growing(hip, mode):
    """A worker's memory pool soon holds what its requests need: after a
    first pass over three requests, further passes allocate nothing more
    from HIP (pool bytes stay put), and the results stay the same bit for
    bit."""
    probs = PROBS[:3]
    passes = []
    with Service(workers=1) as svc:
        for _ in range(3):
            passes.append([svc.wait(svc.submit(p, mode=mode, seed=5)) for p in probs])
    sizes = [[r.pool_bytes for r in ps] for ps in passes]
    print(f"# pool bytes per request, three passes: {sizes}")
    assert sizes[0][-1] > 0
    assert sizes[1] == sizes[2] == [sizes[1][-1]] * 3
    for ps in passes[1:]:
        for a, b in zip(passes[0], ps):
            assert np.array_equal(a.eval, b.eval) and a.num_iter == b.num_iter


def test_batched_inner_solves_equal_solo(hip):
    """With batching on, the BlockPCG iterations of concurrent mixed requests
    run in shared launches (one per step for all of them); every request
    gets the same eigenvalues and iteration count as with batching off, bit
    for bit, and rounds do carry several requests.  A 3D request too large
    for the batched kernels (over 65,536 rows) runs its own loop in the
    same mix."""
    probs = [P.config4_problem(i) for i in range(8)] + [P.stencil_problem(42, 3, 10)]
    runs = {}
    for batching in (False, True):
        with Service(workers=6, batching=batching) as svc:
            tickets = [svc.submit(p, mode="mixed", seed=40 + i) for i, p in enumerate(probs)]
            runs[batching] = [svc.wait(t) for t in tickets]
            bst = svc.batch_stats()
        if batching:
            print(f"# batched rounds {bst}")
            assert bst["rounds"] > 0 and bst["max_solves"] >= 2
            assert bst["solve_iterations"] > bst["rounds"]
        else:
            assert bst["rounds"] == 0
    for p, a, b in zip(probs, runs[False], runs[True]):
        assert a.nev_conv == b.nev_conv and a.num_iter == b.num_iter, p.name
        assert np.array_equal(a.eval, b.eval), p.name
        assert a.nev_conv >= p.nev

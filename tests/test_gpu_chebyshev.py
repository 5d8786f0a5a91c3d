"""The cyclic Chebyshev over-relaxation schedule (qesmc_params.chebyshev_rho).

  (1) lambda after N scheduled iterations from zero equals the oracle's
      red-black loop with the same per-pass factors
      (oracle.chebyshev_omegas), bit for bit, on the split kernel (wall-mask
      and full-coefficient stencils), the two-colour march (a factor per
      colour within its launch) and the classic layout;
  (2) a solve split over calls that pass the running count
      (chebyshev_start) is the one-call solve, bit for bit, and a run with
      refine_every = K is K-iteration refinements with running starts;
  (3) qesmc_run_static gives qesmc_run's result, and a far running start
      runs at the recurrence's limit, Young's optimum;
  (4) out-of-range factors and batched runs are refused."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _city(O, Q, seed, ncx, ncy, ncz):
    dz = Q.dz_geometric(ncz, 1.0, 1.03)
    blds = Q.random_buildings(seed, ncx, ncy, 2.0, 8, 4.0, 12.0, 2.0, 12.0, margin=2)
    return O.build_city(ncx, ncy, 2.0, dz, blds, (1, 4.0, 10.0, 0.5, 250.0))


def _load(ctx, p):
    ctx.set_geometry(p)
    ctx.set_inflow(p.u0, p.v0, p.w0)


@pytest.mark.parametrize("full,path", [(False, 2), (True, 2), (False, 3), (False, 1)])
def test_schedule_matches_the_oracle(qesmc, oracle, full, path):
    p = _city(oracle, qesmc, 31, 40, 34, 13)
    if full:
        rng = np.random.default_rng(7)
        for a in p.coeffs:
            a *= rng.uniform(0.3, 1.0, a.size).astype(np.float32)
    rho, N = 0.995, 37
    R = oracle.divergence_f32(p)
    ref = oracle.sor_rb_sched_f32(p, R, oracle.chebyshev_omegas(rho, 2 * N))
    with qesmc.Context(0) as ctx:
        _load(ctx, p)
        ctx.set_lambda(None)
        r = ctx.run(itermax=N, tol=-1.0, chebyshev_rho=rho, path=path)
        got = ctx.get_lambda()
    assert r["path"] == path and r["iterations"] == N
    if path != qesmc.PATH_PLAIN:
        assert r["compressed"] == (0 if full else 1)
    assert np.array_equal(got, ref)
    # and not the fixed-omega run's
    assert not np.array_equal(got, oracle.sor_rb_f32(p, R, None, itermax=N, tol=-1.0)[0])


def test_running_start_continues_the_schedule(qesmc, oracle):
    p = _city(oracle, qesmc, 32, 44, 36, 12)
    rho = 0.998
    with qesmc.Context(0) as ctx:
        _load(ctx, p)
        ctx.set_lambda(None)
        ctx.run(itermax=40, tol=-1.0, chebyshev_rho=rho)
        one = ctx.get_lambda()
        ctx.set_lambda(None)
        ctx.run(itermax=17, tol=-1.0, chebyshev_rho=rho)
        ctx.run(itermax=23, tol=-1.0, chebyshev_rho=rho, chebyshev_start=17)
        split = ctx.get_lambda()
        ctx.set_lambda(None)
        ctx.run(itermax=17, tol=-1.0, chebyshev_rho=rho)
        ctx.run(itermax=23, tol=-1.0, chebyshev_rho=rho)  # restarted: a different trajectory
        restarted = ctx.get_lambda()
        K = 13
        ctx.set_lambda(None)
        for c in range(3):
            ctx.refine(itermax=K, tol=-1.0, chebyshev_rho=rho, chebyshev_start=c * K)
        chained = ctx.get_lambda(), ctx.get_velocity()
        ctx.set_lambda(None)
        ctx.run(itermax=3 * K, tol=-1.0, chebyshev_rho=rho, refine_every=K)
        refined = ctx.get_lambda(), ctx.get_velocity()
    assert np.array_equal(one, split)
    assert not np.array_equal(one, restarted)
    assert np.array_equal(chained[0], refined[0])
    for a, b in zip(chained[1], refined[1]):
        assert np.array_equal(a, b)


def test_far_start_runs_at_the_limit(qesmc, oracle):
    """Past the recurrence's fixed point every pass runs at Young's optimum
    2 / (1 + sqrt(1 - rho^2)): a far running start equals the fixed factor,
    and the host does not advance the schedule 2^31 passes to get there."""
    p = _city(oracle, qesmc, 35, 36, 30, 11)
    rho = 0.99
    w = float(np.float32(2.0 / (1.0 + np.sqrt(1.0 - rho * rho))))
    with qesmc.Context(0) as ctx:
        _load(ctx, p)
        ctx.set_lambda(None)
        ctx.run(itermax=20, tol=-1.0, omega=w)
        ref = ctx.get_lambda()
        ctx.set_lambda(None)
        t0 = time.perf_counter()
        ctx.run(itermax=20, tol=-1.0, chebyshev_rho=rho, chebyshev_start=2 ** 31 - 1)
        wall = time.perf_counter() - t0
        got = ctx.get_lambda()
    assert wall < 5.0
    assert np.array_equal(ref, got)


def test_static_entry_runs_the_schedule(qesmc, oracle):
    p = _city(oracle, qesmc, 33, 36, 30, 11)
    with qesmc.Context(0) as ctx:
        _load(ctx, p)
        ctx.set_lambda(None)
        a = ctx.run(itermax=50, tol=1e-6, chebyshev_rho=0.99)
        ref = ctx.get_lambda(), ctx.get_velocity()
        ctx.set_lambda(None)
        b = ctx.run_static(itermax=50, tol=1e-6, chebyshev_rho=0.99)
        got = ctx.get_lambda(), ctx.get_velocity()
    assert a["iterations"] == b["iterations"]
    assert np.array_equal(ref[0], got[0])
    for x, y in zip(ref[1], got[1]):
        assert np.array_equal(x, y)


def test_schedule_refusals(qesmc, oracle):
    p = _city(oracle, qesmc, 34, 28, 24, 9)
    with qesmc.Context(0) as ctx:
        _load(ctx, p)
        for bad in (dict(chebyshev_rho=1.0), dict(chebyshev_rho=-0.1), dict(chebyshev_rho=float("nan")),
                    dict(chebyshev_rho=0.9, chebyshev_start=-1)):
            with pytest.raises(qesmc.QesmcError) as e:
                ctx.run(itermax=5, **bad)
            assert e.value.code == qesmc.ERR_ARG, bad
        with pytest.raises(qesmc.QesmcError) as e:
            qesmc.run_batch([ctx], itermax=5, chebyshev_rho=0.9)
        assert e.value.code == qesmc.ERR_UNSUPPORTED

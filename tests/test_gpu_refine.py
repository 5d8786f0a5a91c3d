"""Mixed-precision refinement of the resident lambda (qesmc_refine).

In exact arithmetic refine(N) from lambda1 is N ordinary SOR iterations from
lambda1 (the correction delta obeys the same affine iteration with R replaced
by r' = R + D lambda1 - sum c lambda1), so:
  (1) from lambda1 = 0, r' = R exactly and lambda = 0 + delta: the result is
      qesmc_run from zero, bit for bit (wall-mask and full-coefficient
      stencils);
  (2) with itermax 0 nothing changes: lambda and the field equal qesmc_run's
      with itermax 0 from the same lambda;
  (3) from the FP64 oracle's lambda (rounded to fp32) the residual is the
      rounding alone: a refinement stays within 2e-6 m/s of the FP64 field,
      which a wrong sign or a dropped neighbour term in r' would not;
  (4) what it is for: after an fp32 solve at the estimated omega has reached
      its rounding-noise floor, a refinement takes the field at least 4x
      closer to the FP64 solution;
  (5) a run with refine_every = K is K-iteration refinements one after
      another, bit for bit, and with it the max|dlambda| stop rule ends at
      the estimated omega within 1e-4 m/s of the FP64 field."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _city(O, Q, seed, ncx, ncy, ncz, stretched):
    rng = np.random.default_rng(seed)
    dz = Q.dz_geometric(ncz, 1.0, 1.02) if stretched else Q.dz_uniform(ncz, 1.0)
    blds = Q.random_buildings(seed, ncx, ncy, 2.0, 10, 4.0, 14.0, 2.0, 16.0, margin=2)
    return O.build_city(ncx, ncy, 2.0, dz, blds, (1, 4.0, 10.0, 0.5, float(rng.uniform(0, 360))))


def _load(ctx, p):
    ctx.set_geometry(p)
    ctx.set_inflow(p.u0, p.v0, p.w0)


def _linf(a, b):
    return float(np.abs(a.astype(np.float64) - b).max())


@pytest.mark.parametrize("full,fused", [(False, False), (True, False), (False, True)])
def test_refine_from_zero_is_the_plain_run(qesmc, oracle, full, fused):
    p = _city(oracle, qesmc, 21, 48, 40, 14, True)
    if full:
        rng = np.random.default_rng(5)
        for a in p.coeffs:
            a *= rng.uniform(0.3, 1.0, a.size).astype(np.float32)
    path = qesmc.PATH_FUSED if fused else qesmc.PATH_SPLIT
    with qesmc.Context(0) as ctx:
        _load(ctx, p)
        ctx.set_lambda(None)
        a = ctx.run(itermax=37, tol=-1.0, omega=1.9, path=path)
        ref = ctx.get_lambda(), ctx.get_velocity()
        ctx.set_lambda(None)
        b = ctx.refine(itermax=37, tol=-1.0, omega=1.9, path=path)
        got = ctx.get_lambda(), ctx.get_velocity()
    assert a["path"] == b["path"] == path
    assert a["compressed"] == (0 if full else 1) and b["compressed"] == a["compressed"]
    assert a["iterations"] == b["iterations"] == 37
    assert np.array_equal(ref[0], got[0])
    for r, g in zip(ref[1], got[1]):
        assert np.array_equal(r, g)


def test_refine_without_iterations_changes_nothing(qesmc, oracle):
    p = _city(oracle, qesmc, 22, 40, 36, 12, False)
    with qesmc.Context(0) as ctx:
        _load(ctx, p)
        ctx.set_lambda(None)
        ctx.run(itermax=60, tol=-1.0)
        lam1 = ctx.get_lambda()
        ctx.run(itermax=0)
        ref = ctx.get_lambda(), ctx.get_velocity()
        assert np.array_equal(ref[0], lam1)
        r = ctx.refine(itermax=0)
        got = ctx.get_lambda(), ctx.get_velocity()
    assert r["iterations"] == 0
    assert np.array_equal(ref[0], got[0])
    for a, b in zip(ref[1], got[1]):
        assert np.array_equal(a, b)


def test_refine_from_the_fp64_solution_stays_there(qesmc, oracle):
    p = _city(oracle, qesmc, 23, 48, 40, 14, True)
    u64, v64, w64, lam64, _, _ = oracle.solve_rb_f64(p, tol=1e-10)
    with qesmc.Context(0) as ctx:
        _load(ctx, p)
        ctx.set_lambda(lam64.astype(np.float32))
        ctx.refine(itermax=40, tol=-1.0, omega=1.78)
        u, v, w = ctx.get_velocity()
    err = max(_linf(u, u64), _linf(v, v64), _linf(w, w64))
    print("refined from the FP64 lambda: L_inf", err)
    assert err <= 2e-6


@pytest.mark.timeout(300)
def test_refine_goes_below_the_fp32_noise_floor(qesmc, oracle):
    from qesmc import warmstart as W

    city = W.seeded_city(1_000_003)
    u64, v64, w64, _, _, _ = oracle.solve_rb_f64(oracle.Problem(**city.arrays()), tol=1e-10)
    with qesmc.Context(0) as ctx:
        W.load_problem(ctx, city)
        _, om = ctx.estimate_omega(steps=200)
        ctx.set_lambda(None)
        errs = []
        for _ in range(4):
            ctx.run(itermax=500, tol=-1.0, omega=om)
            u, v, w = ctx.get_velocity()
            errs.append(max(_linf(u, u64), _linf(v, v64), _linf(w, w64)))
        # the plain fp32 iteration has stopped improving
        floor = min(errs[1:])
        ref = []
        for _ in range(2):
            ctx.refine(itermax=250, tol=-1.0, omega=om)
            u, v, w = ctx.get_velocity()
            ref.append(max(_linf(u, u64), _linf(v, v64), _linf(w, w64)))
    print("omega", om, "plain", errs, "refined", ref)
    assert errs[0] <= 1e-4
    assert max(ref) * 4 <= floor


@pytest.mark.parametrize("full,fused", [(False, False), (True, False), (False, True)])
def test_refine_every_equals_chained_refinements(qesmc, oracle, full, fused):
    """A run that refreshes its base every K iterations equals K-iteration
    qesmc_refine calls one after another (each folds, mirrors k = 0 and forms
    the residual in between), bit for bit; the k = 0 ghosts of the base must
    be re-mirrored at each refresh for this to hold.  On the two-colour march
    (whose correction alternates between two buffer sets: after 23 and 69
    iterations it is in the second, after 46 in the first) the same, and the
    same values as on the split kernel."""
    p = _city(oracle, qesmc, 24, 44, 38, 13, True)
    if full:
        rng = np.random.default_rng(6)
        for a in p.coeffs:
            a *= rng.uniform(0.3, 1.0, a.size).astype(np.float32)
    K = 23
    path = qesmc.PATH_FUSED if fused else qesmc.PATH_SPLIT
    with qesmc.Context(0) as ctx:
        _load(ctx, p)
        ctx.set_lambda(None)
        for _ in range(3):
            ctx.refine(itermax=K, tol=-1.0, omega=1.93, path=path)
        ref = ctx.get_lambda(), ctx.get_velocity()
        ctx.set_lambda(None)
        r = ctx.run(itermax=3 * K, tol=-1.0, omega=1.93, refine_every=K, path=path)
        got = ctx.get_lambda(), ctx.get_velocity()
        assert r["path"] == path
        if fused:
            # the split kernel's refined run: the same relaxations, so the same values
            ctx.set_lambda(None)
            ctx.run(itermax=3 * K, tol=-1.0, omega=1.93, refine_every=K, path=qesmc.PATH_SPLIT)
            assert np.array_equal(ctx.get_lambda(), got[0])
            # a refresh after a stop: the device's stop rule ends the march
            # between refreshes, and the correction is folded from the set it
            # wrote last
            ctx.set_lambda(None)
            a = ctx.run(itermax=500, tol=2e-3, omega=1.93, refine_every=K, path=path)
            fl = ctx.get_lambda()
            ctx.set_lambda(None)
            b = ctx.run(itermax=500, tol=2e-3, omega=1.93, refine_every=K, path=qesmc.PATH_SPLIT)
            assert a["iterations"] == b["iterations"] < 500
            assert np.array_equal(fl, ctx.get_lambda())
        # and not the plain run's (the rounding differs once a base exists)
        ctx.set_lambda(None)
        ctx.run(itermax=3 * K, tol=-1.0, omega=1.93)
        plain = ctx.get_lambda()
    assert r["iterations"] == 3 * K
    assert np.array_equal(ref[0], got[0])
    for a, b in zip(ref[1], got[1]):
        assert np.array_equal(a, b)
    assert not np.array_equal(plain, got[0])
    assert float(np.abs(plain - got[0]).max()) <= 1e-3 * float(np.abs(plain).max())


def test_refine_every_lets_the_stop_rule_end_at_the_estimated_omega(qesmc, oracle):
    """At the estimated omega the plain fp32 iteration's max|dlambda| settles
    on its rounding noise (about 1e-4 on these cities) and the stop rule at
    tol 1e-4 may never end it; with refreshes every 100 iterations it ends,
    within 1e-4 m/s of the FP64 field (production rule: 30 % more after)."""
    from qesmc import warmstart as W

    for seed in (1_000_004, 1_000_005):
        city = W.seeded_city(seed)
        u64, v64, w64, _, _, _ = oracle.solve_rb_f64(oracle.Problem(**city.arrays()), tol=1e-10)
        with qesmc.Context(0) as ctx:
            W.load_problem(ctx, city)
            _, om = ctx.estimate_omega(steps=200)
            ctx.set_lambda(None)
            r = ctx.run(itermax=6000, tol=1e-4, stop_extend=0.3, omega=om, refine_every=100)
            u, v, w = ctx.get_velocity()
        err = max(_linf(u, u64), _linf(v, v64), _linf(w, w64))
        print(seed, "omega", om, "iterations", r["iterations"], "err", err)
        assert r["iterations"] < 2000
        assert err <= 1e-4


def test_refine_every_refusals(qesmc, oracle):
    p = _city(oracle, qesmc, 25, 30, 26, 10, False)
    with qesmc.Context(0) as ctx:
        _load(ctx, p)
        with pytest.raises(qesmc.QesmcError) as e:
            ctx.run(itermax=10, refine_every=5, path=qesmc.PATH_PLAIN)
        assert e.value.code == qesmc.ERR_UNSUPPORTED
        with pytest.raises(qesmc.QesmcError) as e:
            ctx.refine(itermax=10, path=qesmc.PATH_PLAIN)
        assert e.value.code == qesmc.ERR_UNSUPPORTED
        with pytest.raises(qesmc.QesmcError) as e:
            ctx.run(itermax=10, refine_every=-1)
        assert e.value.code == qesmc.ERR_ARG
        with pytest.raises(qesmc.QesmcError) as e:
            ctx.run_static(itermax=10, refine_every=5)
        assert e.value.code == qesmc.ERR_UNSUPPORTED
        with pytest.raises(qesmc.QesmcError) as e:
            qesmc.run_batch([ctx], itermax=10, refine_every=5)
        assert e.value.code == qesmc.ERR_UNSUPPORTED

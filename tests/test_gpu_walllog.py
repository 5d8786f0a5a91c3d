"""Log-law wall correction on the device (qesmc_wall_log_bc) against the
oracle's restatement of Wall::wallLogBC (Wall.cpp:659-797, applied as
qesWindsMain.cpp:157-165 does after a solve), on solved fields of seeded
cities.  The device evaluates atan2f, cosf and sinf with its own math
library, the oracle with the host's, so faces may differ by a few ulp of the
wall-parallel speed of that face: bar 2e-6 x (1 + speed) per component (four
orders of magnitude below the 1e-4 m/s field criterion), and the speed itself
(the 20-step fixed point, which amplifies any difference in its input up to
1e6-fold when it diverges) to 5e-7 relative; which faces are touched and the
zeroed components are exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def city(O, Q, seed, ncx, ncy, ncz, stretched, z0_inflow=0.5):
    rng = np.random.default_rng(seed)
    dz = Q.dz_geometric(ncz, 1.0, 1.02) if stretched else Q.dz_uniform(ncz, 1.0)
    blds = Q.random_buildings(seed, ncx, ncy, 2.0, 14, 4.0, 14.0, 2.0, 16.0, margin=2)
    return O.build_city(ncx, ncy, 2.0, dz, blds, (1, 4.0, 10.0, z0_inflow, float(rng.uniform(0, 360))))


def _compare(ref, got, touched, rel=2e-6, rel_speed=5e-7):
    # per face, the wall-parallel speed the components share (a one-ulp
    # difference of atan2f moves a component by that much of the speed)
    speed = np.sqrt(sum(r.astype(np.float64) ** 2 for r in ref))
    for a, (r, g) in enumerate(zip(ref, got)):
        fin = np.isfinite(r)
        assert np.array_equal(fin, np.isfinite(g)), a
        assert np.array_equal(np.signbit(r[~fin]), np.signbit(g[~fin])), a
        d = np.zeros(r.shape)
        d[fin] = np.abs(r[fin].astype(np.float64) - g[fin])
        bad = np.nonzero(d > rel * (1.0 + speed))[0]
        assert bad.size == 0, (a, [(int(f), [float(x[f]) for x in ref], [float(x[f]) for x in got]) for f in bad[:4]])
    gspeed = np.sqrt(sum(g.astype(np.float64) ** 2 for g in got))
    fin = np.isfinite(speed)
    assert np.all(np.abs(gspeed[fin] - speed[fin]) <= rel_speed * (1.0 + speed[fin]))
    same = np.mean([np.array_equal(r[touched].view(np.uint32), g[touched].view(np.uint32))
                    for r, g in zip(ref, got)])
    return same


@pytest.mark.parametrize("shape,stretched,z0", [((48, 40, 14), True, 0.1), ((37, 29, 11), False, 0.1),
                                                 ((48, 40, 14), True, 0.05), ((40, 24, 9), False, 0.15),
                                                 ((40, 24, 9), False, 0.3)])
def test_wall_log_bc_matches_oracle(qesmc, oracle, shape, stretched, z0):
    """z0 = 0.3 m with dz = 1 m makes the vertical passes' fixed point diverge
    (u*_{n+1} = a - b u*_n with b = ln 3 / ln(0.5/0.3) = 2.15): the reference
    writes speeds of order 1e6 m/s that depend on ulp-level differences of
    their inputs with a gain of ~1e6 (20 steps of b), so two math libraries
    cannot agree to a few ulp there.  That case is held to 1e-4 of the speed;
    the convergent ones (b < 1) to the bars above."""
    p = city(oracle, qesmc, sum(shape), *shape, stretched)
    with qesmc.Context(0) as ctx:
        u, v, w, lam, res = ctx.solve(p, itermax=60)
        ms = ctx.wall_log_bc(z0)
        got = ctx.get_velocity()
    ref_u, ref_v, ref_w, nfaces = oracle.wall_log_bc(p, u, v, w, z0)
    assert nfaces > 0 and ms >= 0
    touched = (ref_u != u) | (ref_v != v) | (ref_w != w)
    assert touched.sum() > 0
    # untouched faces keep the solved values exactly
    for a, b in zip(got, (u, v, w)):
        assert np.array_equal(a[~touched], b[~touched])
    if z0 < 0.25:
        _compare((ref_u, ref_v, ref_w), got, touched)
    else:
        _compare((ref_u, ref_v, ref_w), got, touched, rel=1e-4, rel_speed=1e-4)


def test_wall_log_bc_nonfinite_like_the_reference(qesmc, oracle):
    """dist1 = 0.5 dz = z0 makes log(dist1/z0) = 0: the reference divides by
    zero and writes +-inf; the device writes the same values."""
    p = city(oracle, qesmc, 5, 30, 26, 10, False)
    with qesmc.Context(0) as ctx:
        u, v, w, _, _ = ctx.solve(p, itermax=20)
        ctx.wall_log_bc(0.5)
        got = ctx.get_velocity()
    ref = oracle.wall_log_bc(p, u, v, w, 0.5)[:3]
    assert not np.all(np.isfinite(ref[0]))
    touched = (ref[0] != u) | (ref[1] != v) | (ref[2] != w)
    # the horizontal passes diverge here too (b = ln 3 / ln 2 = 1.58, a gain of
    # ~6000 over the 20 steps on their inputs' last-ulp differences): the
    # non-finite pattern and signs exactly, finite values to 1e-3 of the speed
    _compare(ref, got, touched, rel=1e-3, rel_speed=1e-3)


def test_wall_log_bc_state_and_cut_cells(qesmc, oracle):
    p = city(oracle, qesmc, 9, 24, 20, 8, True)
    with qesmc.Context(0) as ctx:
        ctx.set_geometry(p)
        ctx.set_inflow(p.u0, p.v0, p.w0)
        with pytest.raises(qesmc.QesmcError) as e:
            ctx.wall_log_bc(0.1)
        assert e.value.code == qesmc.ERR_STATE
        flag = p.icellflag.copy()
        flag[flag.size // 2] = 7
        q = oracle.Problem(p.nx, p.ny, p.nz, p.dx, p.dy, p.dz_array, *p.coeffs, flag, p.u0, p.v0, p.w0)
        ctx.set_geometry(q)
        ctx.run(itermax=5)
        before = ctx.get_velocity()
        with pytest.raises(qesmc.QesmcError) as e:
            ctx.wall_log_bc(0.1)
        assert e.value.code == qesmc.ERR_UNSUPPORTED
        for a, b in zip(ctx.get_velocity(), before):
            assert np.array_equal(a, b)

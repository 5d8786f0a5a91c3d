"""Velocity gradient tensor on the device (qesmc_velocity_gradient) against
the oracle's restatement of TURBGeneralData::derivativeVelocity, on solved
fields: bitwise, uniform and stretched grids, ragged shapes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def city(O, Q, seed, ncx, ncy, ncz, stretched):
    rng = np.random.default_rng(seed)
    dz = Q.dz_geometric(ncz, 1.0, 1.02) if stretched else Q.dz_uniform(ncz, 1.5)
    blds = Q.random_buildings(seed, ncx, ncy, 2.0, 6, 4.0, 14.0, 2.0, 12.0, margin=1) if ncx > 8 and ncy > 8 else []
    return O.build_city(ncx, ncy, 2.0, dz, blds, (1, 4.0, 10.0, 0.5, float(rng.uniform(0, 360))))


@pytest.mark.parametrize("shape", [(40, 32, 12), (3, 3, 1), (5, 4, 2), (65, 33, 9)])
@pytest.mark.parametrize("stretched", [False, True])
def test_velocity_gradient_bitwise(qesmc, oracle, shape, stretched):
    p = city(oracle, qesmc, sum(shape), *shape, stretched)
    with qesmc.Context(0) as ctx:
        u, v, w, lam, res = ctx.solve(p, itermax=80)
        G, ms = ctx.velocity_gradient(uniform_z=not stretched, dz=1.5)
    ref = oracle.velocity_gradient(p, u, v, w, uniform_z=not stretched, dz=1.5)
    for n in range(9):
        assert np.array_equal(G[n].view(np.uint32), ref[n].view(np.uint32)), n
    assert ms >= 0


def test_velocity_gradient_needs_a_solved_field(qesmc, oracle):
    p = city(oracle, qesmc, 7, 20, 16, 6, True)
    with qesmc.Context(0) as ctx:
        ctx.set_geometry(p)
        ctx.set_inflow(p.u0, p.v0, p.w0)
        with pytest.raises(qesmc.QesmcError) as e:
            ctx.velocity_gradient()
        assert e.value.code == qesmc.ERR_STATE
        ctx.run(itermax=5)
        G, _ = ctx.velocity_gradient(copy=False)
        assert G is None


def _walls_case(O, Q, seed, stretched):
    """A city whose flags are edited so that every wall pass has cells:
    terrain cells (flag 2) beside fluid and building cells, and a block in
    the air (walls above).  The solve uses the city's coefficients; the
    wall step reads only flags, u, v, w and the grid."""
    p = city(O, Q, seed, 40, 32, 12, stretched)
    NX, NY, NZ = p.nx - 1, p.ny - 1, p.nz - 1
    f = p.icellflag.reshape(NZ, NY, NX).copy()
    f[1:3, 2:5, 30:34] = 2
    f[7:9, 20:24, 5:9] = 0
    return O.Problem(p.nx, p.ny, p.nz, p.dx, p.dy, p.dz_array, *p.coeffs, f.ravel().astype(np.int32), p.u0, p.v0,
                     p.w0)


@pytest.mark.parametrize("modes", [(2, 2), (1, 1), (1, 2), (2, 1), (0, 2), (2, 0)])
@pytest.mark.parametrize("stretched", [False, True])
def test_gradient_walls_match_oracle(qesmc, oracle, modes, stretched):
    """The wall step on the device (qesmc_velocity_gradient_walls) against the
    oracle's restatement of TURBWallBuilding / TURBWallTerrain: bitwise with
    one-sided differences; with the log law, whose log() is the device's
    double log against the host's, within one float ulp (and equal in
    almost every cell), with the same cells changed."""
    bm, tm = modes
    q = _walls_case(oracle, qesmc, 5 + bm + 3 * tm, stretched)
    rng = np.random.default_rng(11)
    z0u = rng.uniform(0.05, 0.3, q.nx * q.ny).astype(np.float32)
    z0v = rng.uniform(0.05, 0.3, q.nx * q.ny).astype(np.float32)
    with qesmc.Context(0) as ctx:
        ctx.set_geometry(q)
        ctx.set_inflow(q.u0, q.v0, q.w0)
        ctx.run(itermax=60)
        u, v, w = ctx.get_velocity()
        G0, _ = ctx.velocity_gradient(uniform_z=False)
        G, ms = ctx.velocity_gradient_walls(bm, tm, 0.1, z0u, z0v)
    ref0 = oracle.velocity_gradient(q, u, v, w)
    for n in range(9):
        assert np.array_equal(G0[n].view(np.uint32), ref0[n].view(np.uint32)), n
    ref = oracle.turb_wall_stairstep(q, ref0, u, v, w, bm, tm, 0.1, z0u, z0v)
    changed = [int((ref[n] != ref0[n]).sum()) for n in range(9)]
    assert sum(changed) > 0 and ms >= 0
    for n in range(9):
        if bm != 1 and tm != 1:
            assert np.array_equal(G[n].view(np.uint32), ref[n].view(np.uint32)), n
        else:
            ulps = np.abs(G[n].view(np.int32).astype(np.int64) - ref[n].view(np.int32).astype(np.int64))
            assert ulps.max() <= 1, (n, int(ulps.max()))
            assert np.array_equal(G[n] != G0[n], ref[n] != ref0[n]), n
            assert (ulps == 0).mean() >= 0.99, n


def test_gradient_walls_refusals(qesmc, oracle):
    q = _walls_case(oracle, qesmc, 3, True)
    with qesmc.Context(0) as ctx:
        ctx.set_geometry(q)
        ctx.set_inflow(q.u0, q.v0, q.w0)
        ctx.run(itermax=5)
        with pytest.raises(qesmc.QesmcError) as e:   # no gradient yet
            ctx.velocity_gradient_walls(2, 2)
        assert e.value.code == qesmc.ERR_STATE
        ctx.velocity_gradient()
        for bad in ((3, 2), (2, -1)):
            with pytest.raises(qesmc.QesmcError) as e:
                ctx.velocity_gradient_walls(*bad)
            assert e.value.code == qesmc.ERR_ARG
        with pytest.raises(qesmc.QesmcError) as e:   # terrain log law without z0 arrays
            ctx.velocity_gradient_walls(2, 1)
        assert e.value.code == qesmc.ERR_ARG
        G0, _ = ctx.velocity_gradient()
        flag = q.icellflag.copy()
        flag[flag.size // 2] = 7
        r = oracle.Problem(q.nx, q.ny, q.nz, q.dx, q.dy, q.dz_array, *q.coeffs, flag, q.u0, q.v0, q.w0)
        ctx.set_geometry(r)
        ctx.run(itermax=5)
        ctx.velocity_gradient()
        with pytest.raises(qesmc.QesmcError) as e:
            ctx.velocity_gradient_walls(2, 2)
        assert e.value.code == qesmc.ERR_UNSUPPORTED

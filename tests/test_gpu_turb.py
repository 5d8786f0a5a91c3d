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

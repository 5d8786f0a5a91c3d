"""Device city builder (qesmc_set_city / qesmc_set_city_inflow) against the
oracle's WINDSGeneralData set-up: flags, wall coefficients, u0/v0/w0 must
be bitwise equal, and a solve on the device-built inputs must equal the
oracle's solve."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("icellflag", "e", "f", "g", "h", "m", "n", "u0", "v0", "w0")

# (ncx, ncy, dz_array kind, dx, buildings, inflow)
CASES = {
    "cube_power": (60, 50, ("u", 30, 1.0), 1.0, [(20.0, 15.0, 12.0, 10.0, 14.0, 0.0, 0.0)],
                   (2, 5.0, 10.0, 0.15, 270.0)),
    "rotated_log": (70, 64, ("g", 24, 1.0), 2.0,
                    [(40.0, 30.0, 30.0, 14.0, 20.0, 0.0, 30.0), (90.0, 70.0, 22.0, 22.0, 35.0, 0.0, 45.0),
                     (20.0, 90.0, 18.0, 9.0, 12.0, 4.0, 117.0)],
                    (1, 4.5, 10.0, 0.5, 200.0)),
    "outside_and_tall_uniform": (40, 36, ("u", 12, 2.0), 2.0,
                                 [(-10.0, 5.0, 30.0, 20.0, 500.0, 0.0, 0.0), (60.0, 60.0, 40.0, 40.0, 9.0, 0.0, 0.0),
                                  (30.0, -4.0, 10.0, 10.0, 3.0, 2.5, 10.0)],
                                 (0, 3.0, 10.0, 0.1, 91.0)),
    "tiny": (2, 2, ("u", 1, 1.0), 1.0, [], (1, 5.0, 10.0, 0.1, 270.0)),
    "no_buildings_stretched": (33, 17, ("g", 9, 1.5), 3.0, [], (1, 5.0, 10.0, 0.1, 333.0)),
}


def dz_of(qesmc, spec):
    kind, ncz, dz0 = spec
    return qesmc.dz_uniform(ncz, dz0) if kind == "u" else qesmc.dz_geometric(ncz, dz0, 1.02)


def assert_inputs_equal(dev, ref, what):
    for name in FIELDS:
        a, b = getattr(dev, name), getattr(ref, name)
        assert a.shape == b.shape, (what, name)
        if not np.array_equal(a.view(np.uint32 if a.dtype == np.float32 else np.int32),
                              b.view(np.uint32 if b.dtype == np.float32 else np.int32)):
            raise AssertionError(f"{what}: {name} differs in {int((a != b).sum())} entries")


@pytest.mark.parametrize("case", sorted(CASES))
def test_device_city_equals_oracle(qesmc, oracle, case):
    ncx, ncy, dzs, dx, blds, inflow = CASES[case]
    dz = dz_of(qesmc, dzs)
    ref = oracle.build_city(ncx, ncy, dx, dz, blds, inflow)
    with qesmc.Context(0) as ctx:
        ctx.set_city(ncx, ncy, dx, dz, blds, inflow)
        assert_inputs_equal(ctx.get_inputs(), ref, case)
        # a new direction and profile on the resident geometry
        for inf in ((1, 6.0, 10.0, 0.3, 10.0), (2, 4.0, 10.0, 0.2, 181.0), (0, 2.0, 10.0, 0.1, 300.0)):
            ctx.set_city_inflow(inf)
            assert_inputs_equal(ctx.get_inputs(), oracle.build_city(ncx, ncy, dx, dz, blds, inf), f"{case} {inf}")


def test_device_city_solve_equals_oracle(qesmc, oracle):
    ncx, ncy, dzs, dx, blds, inflow = CASES["rotated_log"]
    dz = dz_of(qesmc, dzs)
    ref = oracle.build_city(ncx, ncy, dx, dz, blds, inflow)
    with qesmc.Context(0) as ctx:
        ctx.set_city(ncx, ncy, dx, dz, blds, inflow)
        ctx.set_lambda(None)
        r = ctx.run(itermax=120)
        u, v, w = ctx.get_velocity()
        lam = ctx.get_lambda()
    ou, ov, ow, olam, oit, _ = oracle.solve_rb_f32(ref, itermax=120)
    assert r["iterations"] == oit
    for a, b in ((u, ou), (v, ov), (w, ow), (lam, olam)):
        assert np.array_equal(a, b)


def test_device_city_large_random(qesmc, oracle):
    """400 x 400 x 120 cells, 200 seeded buildings, stretched dz: the
    full-size synthetic city, device-built versus the host builder (which the
    CPU tests pin to the oracle) and the oracle on a corner."""
    ncx, ncy, ncz, dx = 400, 400, 120, 2.0
    dz = qesmc.dz_geometric(ncz, 1.0, 1.02)
    blds = qesmc.random_buildings(2024, ncx, ncy, dx, 200, 10.0, 40.0, 10.0, 60.0)
    inflow = (1, 0.5 / 0.4 * np.log(10.5 / 0.5), 10.0, 0.5, 217.0)
    host = qesmc.city_build(ncx, ncy, dx, dz, blds, inflow)
    with qesmc.Context(0) as ctx:
        ctx.set_city(ncx, ncy, dx, dz, blds, inflow)
        assert_inputs_equal(ctx.get_inputs(), host, "400x400x120")


def test_device_city_lognormal_heights(qesmc, oracle):
    ncx, ncy, dx = 150, 120, 1.0
    dz = qesmc.dz_geometric(40, 1.0, 1.02)
    blds = qesmc.random_buildings(7, ncx, ncy, dx, 40, 10.0, 30.0, 20.0, 0.6, lognormal=True, hcap=120.0)
    inflow = (1, 5.0, 10.0, 0.5, 45.0)
    with qesmc.Context(0) as ctx:
        ctx.set_city(ncx, ncy, dx, dz, blds, inflow)
        assert_inputs_equal(ctx.get_inputs(), oracle.build_city(ncx, ncy, dx, dz, blds, inflow), "lognormal")


def test_device_city_errors(qesmc):
    dz = qesmc.dz_uniform(8, 1.0)
    with qesmc.Context(0) as ctx:
        with pytest.raises(qesmc.QesmcError) as e:
            ctx.set_city_inflow((1, 5.0, 10.0, 0.1, 270.0))
        assert e.value.code == qesmc.ERR_STATE
        with pytest.raises(qesmc.QesmcError) as e:
            ctx.set_city(20, 20, 1.0, dz, [], (3, 5.0, 10.0, 0.1, 270.0))
        assert e.value.code == qesmc.ERR_ARG
        with pytest.raises(qesmc.QesmcError) as e:
            ctx.set_city(20, 20, 1.0, dz, [(float("nan"), 1.0, 5.0, 5.0, 5.0, 0.0, 0.0)], (1, 5.0, 10.0, 0.1, 0.0))
        assert e.value.code == qesmc.ERR_ARG
        with pytest.raises(qesmc.QesmcError) as e:
            ctx.set_city(20, 20, 1.0, np.array([1.0, 1.0, -1.0] + [1.0] * 6, np.float32), [], (1, 5.0, 10.0, 0.1, 0.0))
        assert e.value.code == qesmc.ERR_ARG
        # still usable afterwards
        ctx.set_city(20, 20, 1.0, dz, [(5.0, 5.0, 6.0, 6.0, 4.0, 0.0, 0.0)], (1, 5.0, 10.0, 0.1, 0.0))
        assert ctx.run(itermax=5)["iterations"] == 5

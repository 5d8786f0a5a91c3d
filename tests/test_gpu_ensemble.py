"""Ensembles on one geometry: lambda combinations, the device L_inf velocity
difference, and the direction-superposition warm start (the problem is
linear in the inflow, so converged x and y basis solutions combine into a
near-solution for any wind direction)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NCX, NCY, NCZ, DX = 64, 48, 20, 2.0
U_REF = 0.5 / 0.4 * math.log(10.5 / 0.5)


def city_inputs(qesmc):
    dz = qesmc.dz_geometric(NCZ, 1.0, 1.02)
    blds = qesmc.random_buildings(77, NCX, NCY, DX, 10, 6.0, 20.0, 5.0, 25.0, margin=2)
    return dz, blds


def test_combine_and_velocity_diff_bitwise(qesmc):
    dz, blds = city_inputs(qesmc)
    rng = np.random.default_rng(1)
    with qesmc.Context(0) as a, qesmc.Context(0) as b, qesmc.Context(0) as c:
        for x, wind in ((a, 10.0), (b, 250.0), (c, 90.0)):
            x.set_city(NCX, NCY, DX, dz, blds, (1, U_REF, 10.0, 0.5, wind))
        la = rng.standard_normal(a.city.ncell).astype(np.float32)
        lb = rng.standard_normal(a.city.ncell).astype(np.float32)
        a.set_lambda(la)
        b.set_lambda(lb)
        co = np.array([0.3, -1.7], np.float32)
        qesmc.lambda_combine(c, [a, b], co)
        assert np.array_equal(c.get_lambda(), co[0] * la + co[1] * lb)
        qesmc.lambda_combine(c, [c], np.array([2.0], np.float32))  # in place
        assert np.array_equal(c.get_lambda(), np.float32(2.0) * (co[0] * la + co[1] * lb))
        a.run(itermax=30)
        b.run(itermax=45)
        d = qesmc.velocity_linf_diff(a, b)
        ref = max(float(np.abs(x.astype(np.float64) - y).max()) for x, y in zip(a.get_velocity(), b.get_velocity()))
        assert d == ref and d > 0
        assert qesmc.velocity_linf_diff(a, a) == 0.0
        with pytest.raises(qesmc.QesmcError):
            qesmc.lambda_combine(c, [a] * 9, np.ones(9, np.float32))


def test_direction_superposition_warm_start(qesmc, oracle):
    """Basis solves for winds along +x (from 270 deg) and +y (from 180 deg),
    combined for 37 deg: 30 iterations from the combination land within
    1e-4 m/s of the FP64 oracle (30 cold iterations are far from it), and
    under the stop rule max|dlambda| <= 1e-4 the warm solve needs a fraction
    of the cold iterations."""
    dz, blds = city_inputs(qesmc)
    wind = 37.0
    th = math.radians(270.0 - wind)
    coefs = np.array([math.cos(th), math.sin(th)], np.float32)
    with qesmc.Context(0) as bx, qesmc.Context(0) as by, qesmc.Context(0) as d:
        bx.set_city(NCX, NCY, DX, dz, blds, (1, U_REF, 10.0, 0.5, 270.0))
        by.set_city(NCX, NCY, DX, dz, blds, (1, U_REF, 10.0, 0.5, 180.0))
        d.set_city(NCX, NCY, DX, dz, blds, (1, U_REF, 10.0, 0.5, wind))
        qesmc.run_batch([bx, by], itermax=6000, tol=-1.0)  # far past the fp32 floor
        d.set_lambda(None)
        d.run(itermax=30, tol=-1.0)
        cold30 = d.get_velocity()
        qesmc.lambda_combine(d, [bx, by], coefs)
        d.run(itermax=30, tol=-1.0)
        warm30 = d.get_velocity()
        d.set_lambda(None)
        cold_it = d.run(itermax=20000, tol=1e-4)["iterations"]
        qesmc.lambda_combine(d, [bx, by], coefs)
        warm_it = d.run(itermax=20000, tol=1e-4)["iterations"]
        inputs = d.get_inputs()
    p = oracle.Problem(**inputs.arrays())
    ref = oracle.solve_rb_f64(p, tol=1e-9)[:3]

    def err(uvw):
        return max(float(np.abs(a.astype(np.float64) - b).max()) for a, b in zip(uvw, ref))

    assert err(warm30) <= 1e-4, err(warm30)
    assert err(cold30) > 10 * err(warm30), (err(cold30), err(warm30))
    assert warm_it < 0.5 * cold_it < 5000, (warm_it, cold_it)


def test_probes_gather_the_resident_field(qesmc):
    """qesmc_get_probes returns u, v, w of the corrected field at the listed
    faces, in list order (repeats and the first and last face included), and
    refuses indices outside the face arrays."""
    dz, blds = city_inputs(qesmc)
    with qesmc.Context(0) as a:
        a.set_city(NCX, NCY, DX, dz, blds, (1, U_REF, 10.0, 0.5, 123.0))
        a.run(itermax=40)
        u, v, w = a.get_velocity()
        n = u.size
        idx = np.random.default_rng(4).integers(0, n, 5000)
        idx = np.concatenate([[0, n - 1, 7, 7], idx]).astype(np.int64)
        qesmc.set_probes(a, idx)
        pu, pv, pw = qesmc.get_probes(a)
        assert np.array_equal(pu, u[idx]) and np.array_equal(pv, v[idx]) and np.array_equal(pw, w[idx])
        # a later run: the same list reads the new field
        a.run(itermax=10)
        u2, _, w2 = a.get_velocity()
        pu2, _, pw2 = qesmc.get_probes(a)
        assert np.array_equal(pu2, u2[idx]) and np.array_equal(pw2, w2[idx])
        for bad in ([n], [-1]):
            with pytest.raises(qesmc.QesmcError):
                qesmc.set_probes(a, np.array(bad, np.int64))

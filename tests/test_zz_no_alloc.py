"""qesmc_run_static (the entry point that allocates nothing) against
qesmc_run through the C ABI: bitwise equal velocity, lambda, iteration count
and last residual, on two calls in a row (the second on a smaller problem
with fewer iterations, so stale convergence rows of the first call would
show), on the split and the classic-layout paths and with a tolerance that
stops the solve early; its limits are refused with clear codes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _city(Q, ncx, ncy, ncz, seed, wind):
    dz = Q.dz_geometric(ncz, 1.0, 1.02)
    blds = Q.random_buildings(seed, ncx, ncy, 2.0, 12, 6.0, 20.0, 4.0, 24.0, margin=2)
    return dz, blds, (1, 4.0, 10.0, 0.5, wind)


def _pair(Q, ncx, ncy, ncz, seed, wind):
    dz, blds, inflow = _city(Q, ncx, ncy, ncz, seed, wind)
    a, b = Q.Context(0), Q.Context(0)
    for c in (a, b):
        c.set_city(ncx, ncy, 2.0, dz, blds, inflow)
    return a, b


def _same(ra, rb, a, b):
    for key in ("iterations", "final_max_dlambda", "path", "compressed", "status"):
        assert ra[key] == rb[key], (key, ra[key], rb[key])
    for x, y in zip(a.get_velocity(), b.get_velocity()):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    assert np.array_equal(a.get_lambda().view(np.uint32), b.get_lambda().view(np.uint32))


@pytest.mark.parametrize("path", ["split", "plain"])
def test_run_static_matches_run_twice(qesmc, path):
    Q = qesmc
    pv = {"split": Q.PATH_SPLIT, "plain": Q.PATH_PLAIN}[path]
    # first call: 96 x 80 x 30 cells, 700 iterations, stopped by tol on the way
    a, b = _pair(Q, 96, 80, 30, 11, 250.0)
    try:
        ra = a.run(itermax=700, tol=2e-4, path=pv)
        rb = b.run_static(itermax=700, tol=2e-4, path=pv)
        assert 0 < rb["iterations"] < 700
        _same(ra, rb, a, b)
    finally:
        a.close()
        b.close()
    # second call: a smaller city, fewer iterations, the reference's tol
    a, b = _pair(Q, 40, 34, 12, 12, 20.0)
    try:
        ra = a.run(itermax=90, tol=1e-9, path=pv)
        rb = b.run_static(itermax=90, tol=1e-9, path=pv)
        assert rb["iterations"] == 90
        _same(ra, rb, a, b)
        # a warm second solve on the same contexts
        ra = a.run(itermax=40, path=pv)
        rb = b.run_static(itermax=40, path=pv)
        _same(ra, rb, a, b)
    finally:
        a.close()
        b.close()


def test_run_static_limits(qesmc):
    Q = qesmc
    a, b = _pair(Q, 24, 20, 8, 13, 100.0)
    try:
        with pytest.raises(Q.QesmcError) as e:
            b.run_static(itermax=65537)
        assert e.value.code == Q.ERR_LIMIT
        with pytest.raises(Q.QesmcError) as e:
            b.run_static(itermax=10, path=Q.PATH_FUSED)
        assert e.value.code == Q.ERR_STATE
        # after a fused qesmc_run the fused path's buffers exist
        ra = a.run(itermax=10, path=Q.PATH_FUSED)
        b.run(itermax=10, path=Q.PATH_FUSED)
        a.set_lambda(None)
        b.set_lambda(None)
        ra = a.run(itermax=25, path=Q.PATH_FUSED)
        rb = b.run_static(itermax=25, path=Q.PATH_FUSED)
        _same(ra, rb, a, b)
        # the largest allowed count runs (stopped early by tol)
        b.set_lambda(None)
        rb = b.run_static(itermax=65536, tol=1e-3)
        assert 0 < rb["iterations"] < 65536
    finally:
        a.close()
        b.close()

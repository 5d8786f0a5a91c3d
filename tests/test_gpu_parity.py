"""Parity of the HIP solver (through the C ABI) with the CPU oracle.

The library keeps the reference's operand order and promotions and is built
without FMA contraction, so the bar here is bit-exact equality of lambda, u,
v, w, the iteration count and the last max|dlambda| with the oracle's
restatement of Solver_CPU_RB::solve.  The 1e-4 m/s tolerance
(max |du,dv,dw| <= 1e-4 m/s) is asserted as well, and again against the FP64
converged oracle where a time-to-tolerance is claimed.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

AUTO, PLAIN, SPLIT, FUSED = 0, 1, 2, 3


def single_cube_city(O):
    """100 x 100 x 50 cells, dx = dz = 1 m, 20 m cube centred at (50, 50),
    power law 5 (z/10)^0.15 from 270 deg."""
    return O.build_city(100, 100, 1.0, O.dz_uniform(50, 1.0), [(40.0, 40.0, 20.0, 20.0, 20.0, 0.0, 0.0)],
                        (2, 5.0, 10.0, 0.15, 270.0))


def random_city(O, Q, seed, ncx, ncy, ncz, dx=2.0, nb=6, kind=1, wind=None):
    rng = np.random.default_rng(seed)
    dz = Q.dz_geometric(ncz, 1.0, 1.02) if seed % 2 else Q.dz_uniform(ncz, 1.0)
    blds = Q.random_buildings(seed, ncx, ncy, dx, nb, 2.0 * dx, max(3.0 * dx, ncx * dx / 4), 2.0,
                              max(3.0, 0.6 * float(dz[1:].sum())), margin=1) if ncx > 8 and ncy > 8 else []
    wind = float(rng.uniform(0, 360)) if wind is None else wind
    prof = (kind, 4.0, 10.0, {0: 0.1, 1: 0.5, 2: 0.15}[kind], wind)
    return O.build_city(ncx, ncy, dx, dz, blds, prof)


def assert_same(gpu, ref, what):
    a, b = gpu, ref
    assert a.shape == b.shape, what
    if not np.array_equal(a, b, equal_nan=True):
        d = np.abs(a.astype(np.float64) - b.astype(np.float64))
        raise AssertionError(f"{what}: {int((a != b).sum())} entries differ, max |diff| {np.nanmax(d):.3e}")


def check_solve(ctx, O, p, path, lam0=None, **kw):
    u, v, w, lam, res = ctx.solve(p, lam0, path=path, **kw)
    kw.setdefault("itermax", 500)
    kw.setdefault("tol", 1e-9)
    ou, ov, ow, olam, oit, ome = O.solve_rb_f32(p, lam=lam0, itermax=kw["itermax"], tol=kw["tol"])
    assert res["iterations"] == oit
    assert np.float32(res["final_max_dlambda"]) == np.float32(ome)
    assert_same(lam, olam, "lambda")
    for g, o, name in ((u, ou, "u"), (v, ov, "v"), (w, ow, "w")):
        assert_same(g, o, name)
        assert np.abs(g.astype(np.float64) - o).max() <= 1e-4
    return res


@pytest.mark.parametrize("path", [PLAIN, SPLIT, FUSED])
def test_single_cube_100x100x50_reference_schedule_bit_exact(gpu_ctx, oracle, path):
    p = single_cube_city(oracle)
    res = check_solve(gpu_ctx, oracle, p, path)
    assert res["iterations"] == 500 and res["path"] == path
    if path != PLAIN:
        assert res["compressed"] == 1


@pytest.mark.parametrize("shape", [(2, 2, 1), (3, 3, 1), (5, 4, 2), (7, 5, 3), (13, 11, 6), (65, 33, 9),
                                   (130, 67, 17), (201, 3, 4)])
@pytest.mark.parametrize("path", [PLAIN, SPLIT, FUSED])
def test_ragged_shapes_bit_exact(gpu_ctx, oracle, qesmc, shape, path):
    ncx, ncy, ncz = shape
    seed = ncx * 31 + ncy * 7 + ncz
    p = random_city(oracle, qesmc, seed, ncx, ncy, ncz)
    lam0 = np.random.default_rng(seed).standard_normal(p.ncell).astype(np.float32)
    for itermax in (0, 1, 2, 37):
        check_solve(gpu_ctx, oracle, p, path, lam0=None, itermax=itermax)
    check_solve(gpu_ctx, oracle, p, path, lam0=lam0, itermax=23)


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_inflow_kinds_and_directions(gpu_ctx, oracle, qesmc, kind):
    for wind in (0.0, 90.0, 187.5, 333.0):
        p = random_city(oracle, qesmc, 5 + kind, 48, 40, 12, kind=kind, wind=wind)
        check_solve(gpu_ctx, oracle, p, SPLIT, itermax=60)


@pytest.mark.parametrize("check_every", [1, 3, 64])
def test_tolerance_stop_matches_oracle(gpu_ctx, oracle, qesmc, check_every):
    """A tolerance reached mid-run: same stopping iteration, whatever the
    host polling interval."""
    p = random_city(oracle, qesmc, 21, 40, 36, 10)
    R = oracle.divergence_f32(p)
    errs, lam = [], None
    for _ in range(40):
        lam, it, me = oracle.sor_rb_f32(p, R, lam, itermax=1, tol=0.0)
        errs.append(me)
    tol = float(np.float32(errs[17]))
    for path in (PLAIN, SPLIT, FUSED):
        res = check_solve(gpu_ctx, oracle, p, path, itermax=400, tol=tol, check_every=check_every)
        assert res["iterations"] == 18


@pytest.mark.parametrize("tol", [1.0, 5.0, float("nan")])
def test_no_iteration_when_initial_error_not_above_tol(gpu_ctx, oracle, qesmc, tol):
    p = random_city(oracle, qesmc, 8, 20, 18, 6)
    lam0 = np.random.default_rng(2).standard_normal(p.ncell).astype(np.float32)
    for path in (PLAIN, SPLIT, FUSED):
        res = check_solve(gpu_ctx, oracle, p, path, lam0=lam0, itermax=50, tol=tol)
        assert res["iterations"] == 0


def test_fractional_coefficients_use_full_stencil(gpu_ctx, oracle, qesmc):
    """Coefficients that are not stair-step walls (cut-cell style fractions)
    cannot be mask-encoded; the split path falls back to full coefficients
    and stays bit-exact."""
    p = random_city(oracle, qesmc, 33, 60, 44, 12)
    rng = np.random.default_rng(4)
    for a in p.coeffs:
        a *= rng.uniform(0.3, 1.0, a.size).astype(np.float32)
    res = check_solve(gpu_ctx, oracle, p, SPLIT, itermax=80)
    assert res["compressed"] == 0
    check_solve(gpu_ctx, oracle, p, PLAIN, itermax=80)
    # the fused path has no full-coefficient variant: it refuses
    with pytest.raises(qesmc.QesmcError) as ei:
        gpu_ctx.solve(p, path=FUSED, itermax=5)
    assert ei.value.code == qesmc.ERR_UNSUPPORTED
    # back to a stair-step geometry on the same context: mask path again
    q = random_city(oracle, qesmc, 34, 60, 44, 12)
    assert check_solve(gpu_ctx, oracle, q, SPLIT, itermax=20)["compressed"] == 1


def test_zero_diagonal_is_an_error(gpu_ctx, oracle, qesmc):
    p = random_city(oracle, qesmc, 9, 20, 16, 6)
    NX, NY = p.nx - 1, p.ny - 1
    c = 5 + 6 * NX + 3 * NX * NY
    for a in p.coeffs:
        a[c] = 0.0
    with pytest.raises(qesmc.QesmcError) as ei:
        gpu_ctx.solve(p)
    assert ei.value.code == qesmc.ERR_ZERO_DIAG


def test_non_finite_lambda_is_reported(gpu_ctx, oracle, qesmc):
    p = random_city(oracle, qesmc, 10, 20, 16, 6)
    nx, ny = p.nx, p.ny
    p.u0[7 + 5 * nx + 3 * nx * ny] = np.float32(3e38)
    p.u0[8 + 5 * nx + 3 * nx * ny] = np.float32(-3e38)
    with pytest.raises(qesmc.QesmcError) as ei:
        gpu_ctx.solve(p, itermax=20)
    assert ei.value.code == qesmc.ERR_NONFINITE


def test_resident_interface_many_directions(qesmc, oracle):
    """Geometry uploaded once; one inflow per wind direction; lambda reset
    between runs (cold start), then a warm-started run."""
    base = random_city(oracle, qesmc, 41, 70, 52, 14)
    with qesmc.Context(0) as ctx:
        ctx.set_geometry(base)
        for wind in (0.0, 45.0, 101.25, 270.0):
            p = oracle.build_city(70, 52, 2.0, base.dz_array, [], (1, 4.0, 10.0, 0.5, wind))
            # same geometry: reuse the base arrays, new inflow only
            q = oracle.Problem(base.nx, base.ny, base.nz, base.dx, base.dy, base.dz_array, *base.coeffs,
                               base.icellflag, p.u0, p.v0, p.w0)
            oracle.lib().oracle_set_velocity_zero(q.nx, q.ny, q.nz, q.icellflag, q.u0, q.v0, q.w0)
            ctx.set_inflow(q.u0, q.v0, q.w0)
            ctx.set_lambda(None)
            res = ctx.run(itermax=90)
            u, v, w = ctx.get_velocity()
            lam = ctx.get_lambda()
            ou, ov, ow, olam, oit, ome = oracle.solve_rb_f32(q, itermax=90)
            assert res["iterations"] == oit
            assert_same(lam, olam, "lambda")
            assert_same(u, ou, "u")
            assert_same(v, ov, "v")
            assert_same(w, ow, "w")
        # warm start: continue from the last lambda for 30 more iterations
        res = ctx.run(itermax=30)
        ou2, ov2, ow2, olam2, oit2, _ = oracle.solve_rb_f32(q, lam=olam, itermax=30)
        assert_same(ctx.get_lambda(), olam2, "lambda warm")
        assert_same(ctx.get_velocity()[0], ou2, "u warm")


def test_device_divergence_norms_match_oracle(gpu_ctx, oracle):
    p = single_cube_city(oracle)
    u, v, w, lam, res = gpu_ctx.solve(p, itermax=200)
    for physical in (False, True):
        gd = gpu_ctx.divergence_linf(physical)
        od = oracle.div_linf(p, u, v, w, physical)
        assert math.isclose(gd, od, rel_tol=1e-12, abs_tol=1e-15), (physical, gd, od)
    d0 = oracle.div_linf(p, p.u0, p.v0, p.w0)
    assert gpu_ctx.divergence_linf(False) < 0.1 * d0


def test_time_to_tolerance_against_fp64_oracle(gpu_ctx, oracle):
    """The 100 x 100 x 50 single-cube case run until max|dlambda| reaches
    fp32's floor: within 1e-4 m/s of the FP64 converged oracle."""
    p = single_cube_city(oracle)
    u64, v64, w64, lam64, it64, me64 = oracle.solve_rb_f64(p, tol=1e-9)
    assert me64 <= 1e-9
    u, v, w, lam, res = gpu_ctx.solve(p, itermax=1000, tol=1e-9)
    err = max(np.abs(u - u64).max(), np.abs(v - v64).max(), np.abs(w - w64).max())
    assert err <= 1e-4, err
    # and the post-correction divergence is no worse than the oracle's fp32 one
    ou, ov, ow, *_ = oracle.solve_rb_f32(p, itermax=1000)
    assert gpu_ctx.divergence_linf(False) <= oracle.div_linf(p, ou, ov, ow) * (1 + 1e-12)


def test_city_400x400x120_reference_schedule(gpu_ctx, oracle, qesmc):
    """400 x 400 x 120, dx = 2 m, geometric dz from 1 m at ratio 1.02, 200
    seeded buildings, log law (u* = 0.5 m/s, z0 = 0.5 m): 500 iterations on
    both paths, bit-exact against the oracle (run on the host cores)."""
    ncx, ncy, ncz, dx = 400, 400, 120, 2.0
    dz = qesmc.dz_geometric(ncz, 1.0, 1.02)
    blds = qesmc.random_buildings(2024, ncx, ncy, dx, 200, 10.0, 40.0, 10.0, 60.0, margin=2)
    assert len(blds) == 200
    u_ref = 0.5 / 0.4 * math.log((10.0 + 0.5) / 0.5)
    p = oracle.build_city(ncx, ncy, dx, dz, blds, (1, u_ref, 10.0, 0.5, 217.0))
    ou, ov, ow, olam, oit, ome = oracle.solve_rb_f32(p, itermax=500)
    for path in (FUSED, SPLIT, PLAIN, AUTO):
        u, v, w, lam, res = gpu_ctx.solve(p, path=path)
        # the automatic choice: these arrays fit the Infinity Cache
        assert res["path"] == (SPLIT if path == AUTO else path)
        assert res["iterations"] == oit == 500
        assert_same(lam, olam, f"lambda path {path}")
        assert_same(u, ou, "u")
        assert_same(v, ov, "v")
        assert_same(w, ow, "w")


def test_city_1000x1000x150_bit_exact(oracle, qesmc):
    """The largest workload: 1000 x 1000 x 150 cells (151 M), dx = 1 m,
    geometric dz, 3000 lognormal buildings, log law from 270 deg; 12
    iterations on every path, bit-exact against the oracle (whose OpenMP
    loop runs them on the host cores in seconds)."""
    dz = qesmc.dz_geometric(150, 1.0, 1.02)
    blds = qesmc.random_buildings(555, 1000, 1000, 1.0, 3000, 6.0, 16.0, 20.0, 0.6, lognormal=True, hcap=120.0,
                                  margin=2)
    u_ref = 0.5 / 0.4 * math.log((10.0 + 0.5) / 0.5)
    p = oracle.build_city(1000, 1000, 1.0, dz, blds, (1, u_ref, 10.0, 0.5, 270.0))
    assert p.ncell == 1000 * 1000 * 151
    ou, ov, ow, olam, oit, ome = oracle.solve_rb_f32(p, itermax=12)
    with qesmc.Context(0) as ctx:
        for path, slab in ((SPLIT, 0), (SPLIT, -1), (SPLIT, 20), (FUSED, 0), (PLAIN, 0), (AUTO, 0)):
            u, v, w, lam, res = ctx.solve(p, path=path, itermax=12, slab_planes=slab)
            # the automatic choice: the two-colour march beyond the Infinity Cache
            assert res["iterations"] == oit == 12 and res["path"] == (FUSED if path == AUTO else path)
            assert np.float32(res["final_max_dlambda"]) == np.float32(ome)
            assert_same(lam, olam, f"lambda path {path} slab {slab}")
            assert_same(u, ou, "u")
            assert_same(v, ov, "v")
            assert_same(w, ow, "w")
    # the Chebyshev schedule on the march: a factor per colour within its launch
    R = oracle.divergence_f32(p)
    sched = oracle.sor_rb_sched_f32(p, R, oracle.chebyshev_omegas(0.99999, 2 * 12))
    with qesmc.Context(0) as ctx:
        u, v, w, lam, res = ctx.solve(p, itermax=12, tol=-1.0, chebyshev_rho=0.99999)
        assert res["path"] == FUSED and res["iterations"] == 12
        assert_same(lam, sched, "lambda, Chebyshev schedule on the march")
    # a no-allocation run chooses the march only once its buffers exist
    with qesmc.Context(0) as ctx:
        ctx.set_geometry(p)
        ctx.set_inflow(p.u0, p.v0, p.w0)
        assert ctx.run_static(itermax=2, tol=-1.0)["path"] == SPLIT
        assert ctx.run(itermax=2, tol=-1.0)["path"] == FUSED
        assert ctx.run_static(itermax=2, tol=-1.0)["path"] == FUSED
        # refinement runs on the march too, with the split kernel's values
        ctx.set_lambda(None)
        assert ctx.run(itermax=3, tol=-1.0, refine_every=2)["path"] == FUSED
        lam = ctx.get_lambda()
        ctx.set_lambda(None)
        assert ctx.run(itermax=3, tol=-1.0, refine_every=2, path=SPLIT)["path"] == SPLIT
        assert_same(ctx.get_lambda(), lam, "refined lambda, march against split")


def test_ragged_city_beyond_the_infinity_cache_auto_is_the_split_result(qesmc):
    """703 x 517 x 97 cells (35 M; colour arrays 331 MB, beyond the 256 MB
    Infinity Cache): QESMC_PATH_AUTO runs the two-colour march (x-tiles of
    uneven width, a partial last k-chunk, 64-row segments) and returns the
    split kernel's lambda and field bit for bit (the split kernel is checked
    against the oracle above), plain and with refine_every, to a tolerance
    that stops mid-run."""
    dz = qesmc.dz_geometric(97, 1.0, 1.02)
    blds = qesmc.random_buildings(909, 703, 517, 1.0, 900, 6.0, 16.0, 20.0, 0.6, lognormal=True, hcap=80.0,
                                  margin=2)
    with qesmc.Context(0) as ctx:
        ctx.set_city(703, 517, 1.0, dz, blds, (1, 4.0, 10.0, 0.5, 233.0))
        # a tolerance the reference's rule reaches by iteration 60
        ctx.set_lambda(None)
        tol = float(np.float32(ctx.run(itermax=60, tol=-1.0)["final_max_dlambda"])) * 1.0001
        for kw in (dict(itermax=23, tol=-1.0), dict(itermax=41, tol=-1.0, omega=1.95, refine_every=10),
                   dict(itermax=400, tol=tol)):
            out = {}
            for path in (AUTO, SPLIT):
                ctx.set_lambda(None)
                r = ctx.run(path=path, **kw)
                out[path] = (r, ctx.get_lambda(), ctx.get_velocity())
            assert out[AUTO][0]["path"] == FUSED and out[SPLIT][0]["path"] == SPLIT
            assert out[AUTO][0]["iterations"] == out[SPLIT][0]["iterations"]
            assert_same(out[AUTO][1], out[SPLIT][1], f"lambda {kw}")
            for a, b, n in zip(out[AUTO][2], out[SPLIT][2], "uvw"):
                assert_same(a, b, f"{n} {kw}")
        assert out[AUTO][0]["iterations"] <= 60


@pytest.mark.parametrize("tile", [(1, 1, 1), (3, 1, 2), (8, 2, 8), (16, 4, 4), (5, 4, 16), (64, 2, 1)])
def test_tile_parameters_never_change_results(gpu_ctx, oracle, qesmc, tile):
    """(tile_k, tile_v, wg_per_cu) reshape the work decomposition only."""
    tk, tv, wg = tile
    p = random_city(oracle, qesmc, 77, 150, 90, 21)
    res = check_solve(gpu_ctx, oracle, p, SPLIT, itermax=40, tile_k=tk, tile_v=tv, wg_per_cu=wg)
    assert res["compressed"] == 1
    check_solve(gpu_ctx, oracle, p, PLAIN, itermax=12, tile_k=tk)
    assert check_solve(gpu_ctx, oracle, p, FUSED, itermax=40, tile_k=tk)["tile_k"] == tk


@pytest.mark.parametrize("slab", [1, 2, 3, 7, 18, 19, 40])
@pytest.mark.parametrize("check_every", [1, 64])
def test_slab_schedule_never_changes_results(gpu_ctx, oracle, qesmc, slab, check_every):
    """The interleaved red/black schedule over k-slabs (slab_planes) reorders
    the launches only: bit-exact against the oracle for slabs of one plane up
    to more than the domain, with a tol that stops the solve on the way."""
    p = random_city(oracle, qesmc, 91, 90, 70, 21)
    check_solve(gpu_ctx, oracle, p, SPLIT, itermax=300, tol=2e-3, slab_planes=slab, check_every=check_every)
    check_solve(gpu_ctx, oracle, p, SPLIT, itermax=25, slab_planes=slab, tile_k=3, tile_v=1)

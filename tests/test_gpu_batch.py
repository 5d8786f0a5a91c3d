"""Batched solves (one launch per colour pass for many problems, optional
HIP graph replay) and the tile autotuner, against the CPU oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def city(O, Q, seed, ncx, ncy, ncz, fractional=False):
    rng = np.random.default_rng(seed)
    blds = Q.random_buildings(seed, ncx, ncy, 2.0, 5, 4.0, max(6.0, ncx / 3), 2.0, 14.0, margin=1) \
        if ncx > 10 and ncy > 10 else []
    p = O.build_city(ncx, ncy, 2.0, Q.dz_geometric(ncz, 1.0, 1.02), blds,
                     (1, 3.8, 10.0, 0.5, float(rng.uniform(0, 360))))
    if fractional:
        for a in p.coeffs:
            a *= rng.uniform(0.4, 1.0, a.size).astype(np.float32)
    return p


SHAPES = [(40, 32, 12), (7, 5, 3), (64, 48, 20), (3, 3, 1), (33, 65, 9), (90, 20, 14), (16, 16, 16), (50, 50, 6)]


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("tol", [1e-9, 3e-3])
@pytest.mark.parametrize("tile_v", [0, 1, 2])
def test_batch_equals_oracle_per_problem(qesmc, oracle, use_graph, tol, tile_v):
    """Uneven sizes, one problem with fractional (unmasked) coefficients, a
    tolerance that stops problems at different iterations; the batch's own
    tiling choice, or one or two half-cells per lane."""
    probs = [city(oracle, qesmc, 200 + q, *s, fractional=(q == 2)) for q, s in enumerate(SHAPES)]
    ctxs = [qesmc.Context(0) for _ in probs]
    try:
        for c, p in zip(ctxs, probs):
            c.set_geometry(p)
            c.set_inflow(p.u0, p.v0, p.w0)
            c.set_lambda(None)
        res, ms = qesmc.run_batch(ctxs, itermax=150, tol=tol, check_every=16, use_graph=use_graph, tile_v=tile_v)
        assert len({(r["tile_k"], r["tile_v"], r["wg_per_cu"]) for r in res}) == 1  # one tiling for the batch
        assert res[0]["tile_v"] == tile_v if tile_v else res[0]["tile_v"] in (1, 2)
        iters = set()
        for q, (c, p, r) in enumerate(zip(ctxs, probs, res)):
            ou, ov, ow, olam, oit, ome = oracle.solve_rb_f32(p, itermax=150, tol=tol)
            assert r["iterations"] == oit, (q, r["iterations"], oit)
            iters.add(oit)
            assert np.float32(r["final_max_dlambda"]) == np.float32(ome)
            u, v, w = c.get_velocity()
            assert np.array_equal(c.get_lambda(), olam), q
            assert np.array_equal(u, ou) and np.array_equal(v, ov) and np.array_equal(w, ow), q
        if tol > 1e-6:
            assert len(iters) > 2  # problems really stopped at different iterations
        assert ms > 0
    finally:
        for c in ctxs:
            c.close()


def test_batch_graph_replays_match_eager_and_warm_start(qesmc, oracle):
    """Two consecutive batched runs (the second warm-starts from the first):
    graph replay and eager launches give identical fields."""
    probs = [city(oracle, qesmc, 300 + q, *s) for q, s in enumerate(SHAPES[:5])]
    out = {}
    for g in (False, True):
        ctxs = [qesmc.Context(0) for _ in probs]
        try:
            for c, p in zip(ctxs, probs):
                c.set_geometry(p)
                c.set_inflow(p.u0, p.v0, p.w0)
            qesmc.run_batch(ctxs, itermax=70, tol=1e-9, check_every=32, use_graph=g)
            qesmc.run_batch(ctxs, itermax=45, tol=1e-9, check_every=32, use_graph=g)
            out[g] = [c.get_lambda() for c in ctxs]
        finally:
            for c in ctxs:
                c.close()
    for a, b, p in zip(out[False], out[True], probs):
        assert np.array_equal(a, b)
        lam, _, _ = oracle.sor_rb_f32(p, oracle.divergence_f32(p), None, itermax=70, tol=1e-9)
        lam, _, _ = oracle.sor_rb_f32(p, oracle.divergence_f32(p), lam, itermax=45, tol=1e-9)
        # the oracle mirrors k = 0 after every iteration; so does the library at the end of a run
        assert np.array_equal(a, lam)


def test_autotune_keeps_results(qesmc, oracle):
    p = city(oracle, qesmc, 400, 120, 100, 30)
    with qesmc.Context(0) as c:
        c.set_geometry(p)
        c.set_inflow(p.u0, p.v0, p.w0)
        best = qesmc.autotune(c, iterations=20)
        assert best.tile_k > 0 and best.tile_v in (1, 2, 4) and best.wg_per_cu > 0
        c.set_lambda(None)
        r = c.run(itermax=60)
        assert (r["tile_k"], r["tile_v"], r["wg_per_cu"]) == (best.tile_k, best.tile_v, best.wg_per_cu)
        lam = c.get_lambda()
    olam = oracle.solve_rb_f32(p, itermax=60)[3]
    assert np.array_equal(lam, olam)


def test_batch_reports_its_tiling(qesmc, oracle):
    """One tiling serves the whole batch: unless the caller sets it, the most
    work per wave tile (4 planes x 2 half-cells down to 1 x 1) that leaves at
    least 6 wave tiles per CU over all members, and 5 work-groups per CU from
    48 wave tiles per CU up (4 below, at two half-cells per lane).  A
    128 x 128 x 64 city has 127 rows of 68 half-cells and 64 updated planes:
    1088 wave tiles of 4 x 2, 2176 of 2 x 2; MI355X has 256 CUs (1536 tiles
    needed).  A member's autotuned tiling is left to its single solves."""
    probs = [city(oracle, qesmc, 410 + q, 128, 128, 64) for q in range(2)]
    ctxs = [qesmc.Context(0) for _ in probs]
    try:
        for c, p in zip(ctxs, probs):
            c.set_geometry(p)
            c.set_inflow(p.u0, p.v0, p.w0)
        best = qesmc.autotune(ctxs[0], iterations=10)
        res, _ = qesmc.run_batch(ctxs, itermax=20, tol=1e-9, use_graph=False)
        assert [(r["tile_k"], r["tile_v"], r["wg_per_cu"]) for r in res] == [(4, 2, 4)] * 2
        res, _ = qesmc.run_batch(ctxs[:1], itermax=20, tol=1e-9, use_graph=False)
        assert (res[0]["tile_k"], res[0]["tile_v"], res[0]["wg_per_cu"]) == (2, 2, 4)
        for c in ctxs:
            c.set_lambda(None)
        res, _ = qesmc.run_batch(ctxs, itermax=20, tol=1e-9, use_graph=False, tile_k=2, tile_v=1, wg_per_cu=3)
        assert [(r["tile_k"], r["tile_v"], r["wg_per_cu"]) for r in res] == [(2, 1, 3)] * 2
        lam_b = ctxs[0].get_lambda()
        ctxs[0].set_lambda(None)
        r = ctxs[0].run(itermax=20, tol=1e-9)
        assert (r["tile_k"], r["tile_v"], r["wg_per_cu"]) == (best.tile_k, best.tile_v, best.wg_per_cu)
        assert np.array_equal(ctxs[0].get_lambda(), lam_b)
    finally:
        for c in ctxs:
            c.close()


def test_batch_larger_than_the_infinity_cache(qesmc, oracle):
    """Two 400 x 400 x 120 cities in one batch (≈ 350 MB of split arrays,
    more than the 256 MB Infinity Cache): per-city results bitwise equal to
    the oracle's."""
    import math
    dz = qesmc.dz_geometric(120, 1.0, 1.02)
    u_ref = 0.5 / 0.4 * math.log((10.0 + 0.5) / 0.5)
    probs = [oracle.build_city(400, 400, 2.0, dz,
                               qesmc.random_buildings(2024 + q, 400, 400, 2.0, 200, 10.0, 40.0, 10.0, 60.0, margin=2),
                               (1, u_ref, 10.0, 0.5, 217.0 + 40.0 * q)) for q in range(2)]
    ctxs = [qesmc.Context(0) for _ in probs]
    try:
        for c, p in zip(ctxs, probs):
            c.set_geometry(p)
            c.set_inflow(p.u0, p.v0, p.w0)
            c.set_lambda(None)
        res, _ = qesmc.run_batch(ctxs, itermax=7, tol=1e-9, check_every=4, use_graph=True)
        assert [(r["tile_k"], r["tile_v"], r["wg_per_cu"]) for r in res] == [(4, 2, 5)] * 2  # 149 tiles per CU
        for c, p, r in zip(ctxs, probs, res):
            ou, ov, ow, olam, oit, ome = oracle.solve_rb_f32(p, itermax=7)
            assert r["iterations"] == oit == 7
            assert np.array_equal(c.get_lambda(), olam)
            u, v, w = c.get_velocity()
            assert np.array_equal(u, ou) and np.array_equal(v, ov) and np.array_equal(w, ow)
    finally:
        for c in ctxs:
            c.close()


@pytest.mark.timeout(600)
def test_batch_of_the_benchmark_cities_equals_oracle(qesmc, oracle):
    """The benchmark's own 64-city batch workload (bench.batch_cities: 64-256 x
    64-256 x 64-96 cells, up to 120 buildings, directions 5.625 deg apart), the
    first 16 members at full size: inputs built by the oracle's restatement of
    the reference's builders, 10 batched iterations replayed as HIP graphs of 5,
    each member bitwise equal to the oracle's fp32 red-black solve."""
    import os
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    if root not in sys.path:
        sys.path.insert(0, root)
    import bench

    specs = bench.batch_cities(64)[:16]
    probs = [oracle.build_city(ncx, ncy, 2.0, dz, [tuple(b) for b in blds], inflow)
             for ncx, ncy, dz, blds, inflow in specs]
    assert max(p.ncell for p in probs) > 2_000_000
    ctxs = [qesmc.Context(0) for _ in probs]
    try:
        for c, p in zip(ctxs, probs):
            c.set_geometry(p)
            c.set_inflow(p.u0, p.v0, p.w0)
            c.set_lambda(None)
        res, _ = qesmc.run_batch(ctxs, itermax=10, tol=1e-9, check_every=5, use_graph=True)
        for q, (c, p, r) in enumerate(zip(ctxs, probs, res)):
            u, v, w, lam, its, me = oracle.solve_rb_f32(p, itermax=10)
            assert r["iterations"] == its == 10, q
            assert r["final_max_dlambda"] == np.float32(me), q
            gu, gv, gw = c.get_velocity()
            for a, b in ((gu, u), (gv, v), (gw, w), (c.get_lambda(), lam)):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), q
    finally:
        for c in ctxs:
            c.close()

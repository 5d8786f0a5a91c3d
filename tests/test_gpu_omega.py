"""The over-relaxation estimate (qesmc_estimate_omega): the Jacobi spectral
radius rho_J of the resident problem, by Lanczos in FP64.

Pinned by (1) the closed form on a box without buildings one relaxed layer
deep (dx = dz = 1 m, ncz = 2: the layer k = 1 has the ground wall below, n =
0, and the fixed top above, so every relaxed cell has D = 5 and couples only
horizontally): the modes are Dirichlet sines in x and y, rho_J =
(2 cos(pi/(nx+1)) + 2 cos(pi/(ny+1)))/5 over the nx, ny relaxed cells per
axis; (2) scipy's sparse eigensolver
(ARPACK) on a city with buildings and stretched dz, applied to the same
operator assembled here from the resident coefficients: the fluid cells the
sweep relaxes, couplings e..n, the k = 0 mirror as a diagonal term, weighted
by dz_k (symmetric) and scaled by (D dz_k)^-1/2; (3) what it is for: SOR at
the returned omega reaches a field within 1e-4 m/s of the FP64 oracle in far
fewer iterations than at the reference's 1.78, on held-out cities."""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

pytestmark = pytest.mark.gpu


def test_box_closed_form(qesmc, oracle):
    ncx, ncy, ncz = 40, 30, 2
    p = oracle.build_city(ncx, ncy, 1.0, oracle.dz_uniform(ncz, 1.0), [], (1, 4.0, 10.0, 0.5, 270.0))
    # the relaxed cells: i, j in [1, N-2] of the N = ncx, ncy cells, the one layer k = 1
    nx, ny = ncx - 2, ncy - 2
    rho = (2 * math.cos(math.pi / (nx + 1)) + 2 * math.cos(math.pi / (ny + 1))) / 5
    with qesmc.Context(0) as ctx:
        ctx.set_geometry(p)
        ctx.set_inflow(p.u0, p.v0, p.w0)
        r, om = ctx.estimate_omega(steps=400)
    assert abs(r - rho) <= 1e-9, (r, rho)
    opt = 2 / (1 + math.sqrt(1 - rho * rho))
    assert abs(om - (2 - 1.2 * (2 - opt))) <= 1e-6


def _operator(p):
    """The symmetrised Jacobi operator of the problem's relaxed fluid cells."""
    NX, NY, NZ = p.nx - 1, p.ny - 1, p.nz - 1
    fl = p.icellflag.reshape(NZ, NY, NX)
    co = [a.reshape(NZ, NY, NX).astype(np.float64) for a in p.coeffs]
    act = np.zeros((NZ, NY, NX), bool)
    act[1:NZ - 1, 1:NY - 1, 1:NX - 1] = True
    act &= (fl != 0) & (fl != 2)
    idx = -np.ones((NZ, NY, NX), np.int64)
    idx[act] = np.arange(act.sum())
    D = sum(co)
    dzk = p.dz_array.astype(np.float64)[:, None, None] * np.ones((NZ, NY, NX))
    s = np.sqrt(D * dzk)
    rows, cols, vals = [], [], []
    K, J, I = np.nonzero(act)
    for a, (dk, dj, di) in enumerate(((0, 0, 1), (0, 0, -1), (0, 1, 0), (0, -1, 0), (1, 0, 0), (-1, 0, 0))):
        k2, j2, i2 = K + dk, J + dj, I + di
        if a == 5:  # the k = 0 mirror: cell k = 1 couples to itself
            k2 = np.where(K == 1, K, k2)
        ok = act[k2, j2, i2]
        c = co[a][K, J, I][ok] * dzk[K, J, I][ok] / (s[K, J, I][ok] * s[k2[ok], j2[ok], i2[ok]])
        rows.append(idx[K, J, I][ok])
        cols.append(idx[k2[ok], j2[ok], i2[ok]])
        vals.append(c)
    n = int(act.sum())
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


def test_city_against_arpack(qesmc, oracle):
    dz = oracle.dz_stretched(1.0 * 1.04 ** np.arange(16))
    blds = qesmc.random_buildings(5, 60, 44, 2.0, 10, 6.0, 18.0, 3.0, 20.0, margin=2)
    p = oracle.build_city(60, 44, 2.0, dz, blds, (1, 4.0, 10.0, 0.5, 250.0))
    M = _operator(p)
    # the dz weighting makes it symmetric up to the fp32 rounding of the
    # stored coefficients (dz_k m_k against dz_k+1 n_k+1)
    assert abs(M - M.T).max() <= 1e-6
    lam = float(spla.eigsh(0.5 * (M + M.T), k=1, which="LA", tol=1e-12)[0][0])
    with qesmc.Context(0) as ctx:
        ctx.set_geometry(p)
        ctx.set_inflow(p.u0, p.v0, p.w0)
        r, _ = ctx.estimate_omega(steps=400)
    assert abs(r - lam) <= 1e-7, (r, lam)


@pytest.mark.timeout(600)
def test_estimated_omega_reaches_the_field_criterion_sooner(qesmc, oracle):
    from qesmc import warmstart as W

    with qesmc.Context(0) as ctx:
        for seed in (1_000_000, 1_000_001):
            city = W.seeded_city(seed)
            u64, v64, w64, _, _, _ = oracle.solve_rb_f64(oracle.Problem(**city.arrays()), tol=1e-9)
            W.load_problem(ctx, city)
            _, om = ctx.estimate_omega(steps=200)
            assert 1.78 < om < 2.0
            its = {}
            for omega in (1.78, om):
                ctx.set_lambda(None)
                n = 0
                while n < 6000:
                    ctx.run(itermax=25, tol=-1.0, omega=omega)
                    n += 25
                    u, v, w = ctx.get_velocity()
                    err = max(float(np.abs(a.astype(np.float64) - b).max()) for a, b in ((u, u64), (v, v64), (w, w64)))
                    if err <= 1e-4:
                        break
                its[omega] = n if err <= 1e-4 else None
            print(seed, om, its)
            assert its[om] is not None and its[1.78] is not None
            assert its[om] * 2 <= its[1.78]


def test_estimate_refusals_and_small_domains(qesmc, oracle):
    """Arguments and state: steps in [1, 100000], geometry resident; on a
    domain with a single relaxed cell the estimate is that cell's own
    coupling (its k = 0 mirror term over its diagonal), and the one-step
    estimate is a Ritz value, never above the spectral radius."""
    with qesmc.Context(0) as ctx:
        with pytest.raises(qesmc.QesmcError) as e:
            ctx.estimate_omega(steps=10)
        assert e.value.code == qesmc.ERR_STATE
        p = oracle.build_city(40, 30, 1.0, oracle.dz_uniform(2, 1.0), [], (1, 4.0, 10.0, 0.5, 270.0))
        ctx.set_geometry(p)
        for bad in (0, 100001):
            with pytest.raises(qesmc.QesmcError) as e:
                ctx.estimate_omega(steps=bad)
            assert e.value.code == qesmc.ERR_ARG
        r1, _ = ctx.estimate_omega(steps=1)
        r, om = ctx.estimate_omega(steps=400)
        assert 0.0 < r1 <= r + 1e-12 < 1.0
        assert 1.0 < om < 2.0
        # one relaxed cell (i = j = k = 1 of a 3 x 3 x 3-cell box) with a
        # coupling n = 0.5 to the ground ghost below it, which the mirror makes
        # its own value: rho_J = n / D
        q = oracle.build_city(3, 3, 1.0, oracle.dz_uniform(2, 1.0), [], (1, 4.0, 10.0, 0.5, 270.0))
        c = 1 + 1 * 3 + 1 * 9
        q.coeffs[5][c] = 0.5
        ctx.set_geometry(q)
        r, _ = ctx.estimate_omega(steps=5)
        co = [float(a[c]) for a in q.coeffs]
        assert co[5] == 0.5 and abs(r - 0.5 / sum(co)) <= 1e-12, (r, co)

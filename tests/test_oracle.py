"""Pins of the CPU oracle (oracle/qesmc_oracle.c) against what the reference and
the mathematics fix.  CPU only.

Each check is chosen so that a plausible mistake in the oracle (a dropped
stencil term, a wrong sign or neighbour index, a misplaced mirror or
correction) fails at least one of them:
  * the reference's own stored solver outputs (legacy 40^3 single building),
  * a sparse direct solve of the same discrete operator (scipy),
  * a divergence-free invariant of the converged, corrected field,
  * fp32 and FP64 trajectories iteration by iteration,
  * thread-count independence, zero-source and mirror-symmetry invariants,
  * closed forms of the stair-step flags, wall coefficients and profiles.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def small_city(O, ncx=30, ncy=24, ncz=14, dx=2.0, stretched=True, wind=233.0, kind=1):
    dz = O.dz_stretched(1.0 * 1.05 ** np.arange(ncz)) if stretched else O.dz_uniform(ncz, 1.0)
    blds = [(14.0, 10.0, 12.0, 8.0, 6.0, 0.0, 0.0), (36.0, 24.0, 10.0, 14.0, 9.5, 0.0, 0.0)]
    prof = {0: (0, 4.0, 10.0, 0.1, wind), 1: (1, 4.0, 10.0, 0.5, wind), 2: (2, 5.0, 10.0, 0.15, wind)}[kind]
    return O.build_city(ncx, ncy, dx, dz, blds, prof)


# ---------------------------------------------------------------------------
# the reference's stored outputs
# ---------------------------------------------------------------------------
def legacy_problem(O):
    """The 2018 single-building case of
    scratch/solverTestCases/Single Building/CPU/Single_building.cpp:33-264:
    40^3 cells of 5 m, a 20 m cube at x, y in [90, 110] m, log profile
    U_ref = 5 m/s at 10 m, z0 = 0.1 m; walls applied on the interior range
    only and every layer k in [1, nz-2] normalised (:198-264)."""
    NX = NY = 40
    NZ = 41
    nx, ny, nz = 41, 41, 42
    dx = 5.0
    flag = np.ones((NZ, NY, NX), np.int32)
    flag[0:5, 18:22, 18:22] = 0
    flag[0] = 0
    co = np.ones((6, NZ, NY, NX), np.float32)
    e, f, g, h, m, n = co
    fl = flag
    inner = (slice(1, nz - 2), slice(1, ny - 2), slice(1, nx - 2))
    fluid = fl[inner] != 0
    K, J, I = np.meshgrid(np.arange(1, nz - 2), np.arange(1, ny - 2), np.arange(1, nx - 2), indexing="ij")
    for arr, (dk, dj, di) in ((n, (-1, 0, 0)), (m, (1, 0, 0)), (f, (0, 0, -1)), (e, (0, 0, 1)),
                              (h, (0, -1, 0)), (g, (0, 1, 0))):
        wall = fluid & (fl[K + dk, J + dj, I + di] == 0)
        arr[inner][wall] = 0.0
    co[:, 1:nz - 1] /= np.float32(dx * dx)
    z = ((np.arange(nz) - 0.5) * dx).astype(np.float32)
    z0, zr, U = np.float32(0.1), np.float32(10.0), np.float32(5.0)
    with np.errstate(invalid="ignore"):
        prof = (U * (np.log((z + z0) / z0) / np.log((zr + z0) / z0))).astype(np.float32)
    u0 = np.zeros((nz, ny, nx), np.float32)
    v0 = np.zeros_like(u0)
    w0 = np.zeros_like(u0)
    u0[1:] = prof[1:, None, None]
    for k, j, i in zip(*np.nonzero(flag == 0)):
        u0[k, j, i] = u0[k, j, i + 1] = 0.0
        v0[k, j, i] = v0[k, j + 1, i] = 0.0
        w0[k, j, i] = w0[k + 1, j, i] = 0.0
    return O.Problem(nx, ny, nz, dx, dx, np.full(nz - 1, dx, np.float32), *[c.ravel() for c in co],
                     flag.ravel(), u0.ravel(), v0.ravel(), w0.ravel())


def test_fp64_oracle_matches_reference_stored_outputs(oracle):
    """FP64 red-black oracle vs the reference's stored legacy outputs
    (tests/golden/make_legacy_fixture.py extracts them).  Both legacy files
    print 6 significant digits; the GPU file carries the reference GPU's
    Neumann-kernel overshoot (quirk A3), hence its looser bound.  Only
    k <= 38 is compared: the legacy correction also touches the top ghost
    layer's w face, which the current reference (and the oracle) leave at w0."""
    O = oracle
    p = legacy_problem(O)
    u, v, w, lam, it, me = O.solve_rb_f64(p, tol=1e-11)
    assert me <= 1e-11
    nx, ny, nz = p.nx, p.ny, p.nz
    U, V, W = (a.reshape(nz, ny, nx) for a in (u, v, w))
    fx = np.load(os.path.join(GOLDEN, "legacy_single_building.npz"))
    i, j, k = fx["ijk"].astype(int).T
    mine = np.stack([0.5 * (U[k, j, i + 1] + U[k, j, i]), 0.5 * (V[k, j + 1, i] + V[k, j, i]),
                     0.5 * (W[k + 1, j, i] + W[k, j, i])], 1)
    sel = (k >= 1) & (k <= 38)
    d_cpu = np.abs(mine - fx["cpu_uvw"])[sel].max()
    d_gpu = np.abs(mine - fx["gpu_uvw"])[sel].max()
    assert d_cpu <= 2e-5, d_cpu
    assert d_gpu <= 1e-4, d_gpu
    # the comparison is not vacuous: the building perturbs the flow strongly
    assert np.abs(fx["cpu_uvw"][sel][:, 1]).max() > 0.5


# ---------------------------------------------------------------------------
# the SOR operator against a sparse direct solve
# ---------------------------------------------------------------------------
def assemble(p, R):
    """The discrete problem the red-black sweep relaxes, assembled directly:
    for every updated cell, S*lam_c - sum_nb c_nb*lam_nb = -R_c, where
    never-updated cells keep lambda = 0 (Dirichlet ring and top layer) and
    the k = 0 ghost mirrors k = 1 (so the n-neighbour of a k = 1 cell is the
    cell itself)."""
    NX, NY, NZ = p.nx - 1, p.ny - 1, p.nz - 1
    cells = [(i, j, k) for k in range(1, NZ - 1) for j in range(1, NY - 1) for i in range(1, NX - 1)]
    index = {c: q for q, c in enumerate(cells)}
    e, f, g, h, m, n = (a.reshape(NZ, NY, NX).astype(np.float64) for a in p.coeffs)
    rows, cols, vals = [], [], []
    b = np.zeros(len(cells))
    Rg = R.reshape(NZ, NY, NX)
    for q, (i, j, k) in enumerate(cells):
        S = e[k, j, i] + f[k, j, i] + g[k, j, i] + h[k, j, i] + m[k, j, i] + n[k, j, i]
        diag = S
        for coef, nb in ((e[k, j, i], (i + 1, j, k)), (f[k, j, i], (i - 1, j, k)), (g[k, j, i], (i, j + 1, k)),
                         (h[k, j, i], (i, j - 1, k)), (m[k, j, i], (i, j, k + 1)), (n[k, j, i], (i, j, k - 1))):
            if nb[2] == 0:
                nb = (nb[0], nb[1], 1)
            if nb == (i, j, k):
                diag -= coef
            elif nb in index:
                rows.append(q)
                cols.append(index[nb])
                vals.append(-coef)
        rows.append(q)
        cols.append(q)
        vals.append(diag)
        b[q] = -Rg[k, j, i]
    A = sp.csr_matrix((vals, (rows, cols)), shape=(len(cells), len(cells)))
    return A, b, cells


def test_fp64_sor_converges_to_sparse_direct_solution(oracle):
    O = oracle
    p = O.build_city(14, 12, 2.0, O.dz_stretched([1.0, 1.1, 1.2, 1.3, 1.5, 1.7, 2.0, 2.3]),
                     [(8.0, 6.0, 8.0, 6.0, 4.0, 0.0, 0.0)], (1, 4.0, 10.0, 0.5, 200.0))
    R = O.divergence_f64(p)
    lam, it, me = O.sor_rb_f64(p, R, tol=1e-13, itermax=100000)
    assert me <= 1e-13
    A, b, cells = assemble(p, R)
    x = spla.spsolve(A.tocsc(), b)
    NX, NY = p.nx - 1, p.ny - 1
    mine = np.array([lam[i + j * NX + k * NX * NY] for (i, j, k) in cells])
    scale = np.abs(x).max()
    assert scale > 1e-2
    assert np.abs(mine - x).max() <= 1e-9 * max(1.0, scale)


def test_manufactured_lambda_is_recovered(oracle):
    """Pick lambda*, set R := A lambda* with the assembled operator, and the
    FP64 sweep must return lambda* (pins neighbour indices and signs)."""
    O = oracle
    p = small_city(O, 16, 12, 8)
    A, _, cells = assemble(p, np.zeros(p.ncell))
    rng = np.random.default_rng(3)
    lam_star = rng.standard_normal(len(cells))
    rhs = A @ lam_star  # = -R
    NX, NY = p.nx - 1, p.ny - 1
    R = np.zeros(p.ncell)
    for q, (i, j, k) in enumerate(cells):
        R[i + j * NX + k * NX * NY] = -rhs[q]
    lam, it, me = O.sor_rb_f64(p, R, tol=1e-13, itermax=200000)
    mine = np.array([lam[i + j * NX + k * NX * NY] for (i, j, k) in cells])
    assert np.abs(mine - lam_star).max() <= 1e-8


# ---------------------------------------------------------------------------
# invariants of the whole step
# ---------------------------------------------------------------------------
def ref_operator_div(p, u, v, w):
    """R(u)/(-2) of Solver_CPU_RB.cpp:56-80 in float64, as a (NZ,NY,NX) array."""
    nx, ny, nz = p.nx, p.ny, p.nz
    NX, NY, NZ = nx - 1, ny - 1, nz - 1
    e, f, g, h, m, n = (a.reshape(NZ, NY, NX).astype(np.float64) for a in p.coeffs)
    U, V, W = (np.asarray(a, np.float64).reshape(nz, ny, nx) for a in (u, v, w))
    dz = p.dz_array.astype(np.float64)
    D = np.zeros((NZ, NY, NX))
    for k in range(1, nz - 2):
        D[k] = ((e[k] * U[k, :NY, 1:nx] - f[k] * U[k, :NY, :NX]) * p.dx
                + (g[k] * V[k, 1:ny, :NX] - h[k] * V[k, :NY, :NX]) * p.dy
                + m[k] * dz[k] * 0.5 * (dz[k] + dz[k + 1]) * W[k + 1, :NY, :NX]
                - n[k] * dz[k] * 0.5 * (dz[k] + dz[k - 1]) * W[k, :NY, :NX])
    return D


def test_converged_field_is_divergence_free(oracle):
    """With dz = 1 m (where quirk A1 is exact) the corrected field of the
    converged FP64 solve has zero reference-operator divergence in every
    updated fluid cell below the top layer, for an arbitrary u0, v0, w0.
    A sign or index error in R, the sweep or the correction breaks this."""
    O = oracle
    p = O.build_city(20, 16, 2.0, O.dz_uniform(10, 1.0),
                     [(10.0, 8.0, 12.0, 10.0, 5.0, 0.0, 0.0), (26.0, 16.0, 6.0, 8.0, 3.0, 0.0, 0.0)],
                     (1, 4.0, 10.0, 0.5, 240.0))
    rng = np.random.default_rng(11)
    for a in (p.u0, p.v0, p.w0):
        a += rng.uniform(-1, 1, a.size).astype(np.float32)
    D0 = ref_operator_div(p, p.u0, p.v0, p.w0)
    u, v, w, lam, it, me = O.solve_rb_f64(p, tol=1e-13, itermax=200000)
    D = ref_operator_div(p, u, v, w)
    NX, NY, NZ = p.nx - 1, p.ny - 1, p.nz - 1
    fl = p.icellflag.reshape(NZ, NY, NX)
    sel = np.zeros_like(fl, bool)
    sel[1:NZ - 2, 1:NY - 1, 1:NX - 1] = True
    sel &= fl == 1
    assert np.abs(D0[sel]).max() > 0.1
    assert np.abs(D[sel]).max() <= 1e-9, np.abs(D[sel]).max()


def test_fp32_and_fp64_trajectories_agree(oracle):
    """Same iteration in fp32 and FP64 for 1, 7 and 60 iterations: pins the
    fp32 loop (colour order, mirror timing, stop test) to the FP64 one."""
    O = oracle
    p = small_city(O)
    R32 = O.divergence_f32(p)
    R64 = O.divergence_f64(p)
    assert np.abs(R32 - R64).max() <= 1e-6 * np.abs(R64).max()
    for n in (1, 7, 60):
        l32, it32, e32 = O.sor_rb_f32(p, R32, itermax=n, tol=0.0)
        l64, it64, e64 = O.sor_rb_f64(p, R64, itermax=n, tol=0.0)
        assert it32 == it64 == n
        scale = np.abs(l64).max()
        assert np.abs(l32 - l64).max() <= 2e-5 * scale, (n, np.abs(l32 - l64).max(), scale)
        # max|dlambda| differs by at most the lambda discrepancy of two iterates
        assert abs(e32 - e64) <= 4e-5 * scale


def test_red_black_is_thread_count_independent(oracle):
    O = oracle
    p = small_city(O)
    t0 = O.num_threads()
    try:
        O.set_num_threads(1)
        a = O.solve_rb_f32(p, itermax=40)
        O.set_num_threads(max(2, os.cpu_count() or 2))
        b = O.solve_rb_f32(p, itermax=40)
    finally:
        O.set_num_threads(t0)
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(x, y)
    assert a[4] == b[4] and a[5] == b[5]


def test_zero_source_gives_zero_lambda(oracle):
    """Uniform inflow over flat ground: R = 0, lambda stays 0, u = u0."""
    O = oracle
    p = O.build_city(18, 14, 2.0, O.dz_uniform(9, 1.0), [], (0, 3.0, 10.0, 0.1, 250.0))
    R = O.divergence_f32(p)
    assert np.all(R == 0)
    u, v, w, lam, it, me = O.solve_rb_f32(p, itermax=25)
    assert np.all(lam == 0) and it == 1 and me == 0.0
    assert np.array_equal(u, p.u0) and np.array_equal(v, p.v0) and np.array_equal(w, p.w0)


def test_mirror_symmetry_of_centred_cube(oracle):
    """Wind from 270 deg on a cube centred in y: the converged v is odd and u
    even under y -> Y - y (face j <-> NY - j)."""
    O = oracle
    p = O.build_city(40, 40, 1.0, O.dz_uniform(20, 1.0), [(15.0, 15.0, 10.0, 10.0, 8.0, 0.0, 0.0)],
                     (2, 5.0, 10.0, 0.15, 270.0))
    u, v, w, lam, it, me = O.solve_rb_f64(p, tol=1e-12)
    nx, ny, nz = p.nx, p.ny, p.nz
    U, V = u.reshape(nz, ny, nx), v.reshape(nz, ny, nx)
    NY = ny - 1
    Vr = V[:, NY - np.arange(ny), :]
    Ur = U[:, (NY - 1 - np.arange(NY)), :]
    assert np.abs(V).max() > 0.1
    assert np.abs(V + Vr).max() <= 1e-9
    assert np.abs(U[:, :NY, :] - Ur).max() <= 1e-9


def test_serial_and_red_black_reach_the_same_field(oracle):
    """Solver_CPU (lexicographic) and Solver_CPU_RB converge to one fixed point."""
    O = oracle
    p = small_city(O, 24, 20, 10)
    u64, v64, w64, *_ = O.solve_rb_f64(p, tol=1e-12)
    for solve in (O.solve_serial_f32, O.solve_rb_f32):
        u, v, w, lam, it, me = solve(p, itermax=3000)
        err = max(np.abs(u - u64).max(), np.abs(v - v64).max(), np.abs(w - w64).max())
        assert err <= 1e-4, (solve.__name__, err)


# ---------------------------------------------------------------------------
# loop control (Solver_CPU_RB.cpp:95)
# ---------------------------------------------------------------------------
def test_stop_rule_matches_per_iteration_errors(oracle):
    O = oracle
    p = small_city(O)
    R = O.divergence_f32(p)
    errs, lam = [], None
    for _ in range(30):
        lam, it, me = O.sor_rb_f32(p, R, lam, itermax=1, tol=0.0)
        errs.append(me)
    errs = np.array(errs)
    tol = float(errs[12]) * 1.0001
    expect = int(np.argmax(errs <= tol)) + 1
    lam2, it, me = O.sor_rb_f32(p, R, None, itermax=1000, tol=tol)
    assert it == expect and me == errs[expect - 1]
    assert np.array_equal(lam2, O.sor_rb_f32(p, R, None, itermax=expect, tol=0.0)[0])


@pytest.mark.parametrize("itermax,tol,expect_iters", [(0, 1e-9, 0), (5, 1.0, 0), (5, 2.0, 0),
                                                      (5, float("nan"), 0), (3, 1e-9, 3)])
def test_loop_entry_conditions(oracle, itermax, tol, expect_iters):
    O = oracle
    p = small_city(O, 12, 10, 6)
    rng = np.random.default_rng(1)
    lam0 = rng.standard_normal(p.ncell).astype(np.float32)
    u, v, w, lam, it, me = O.solve_rb_f32(p, lam=lam0, itermax=itermax, tol=tol)
    assert it == expect_iters
    if expect_iters == 0:
        assert np.array_equal(lam, lam0)


def test_chebyshev_factors_closed_form(oracle):
    """The cyclic Chebyshev factors from their recurrence against the
    Chebyshev-polynomial closed form w_h = 2 C_h(1/rho) / (rho C_(h+1)(1/rho))
    (h >= 1; C_n(x) = cosh(n acosh x)), w_0 = 1, and the limit, Young's
    optimum 2 / (1 + sqrt(1 - rho^2)); a running start continues the same
    sequence."""
    O = oracle
    for rho in (0.5, 0.9, 0.99):
        w = O.chebyshev_omegas(rho, 60).astype(np.float64)
        a = np.arccosh(1.0 / rho)
        h = np.arange(1, 60)
        closed = 2.0 * np.cosh(h * a) / (rho * np.cosh((h + 1) * a))
        assert w[0] == 1.0
        np.testing.assert_allclose(w[1:], closed, rtol=2e-7)
        assert np.all(np.diff(w[1:]) <= 0)  # falls to the optimum from above
        lim = O.chebyshev_omegas(rho, 1, start=4000)[0]
        assert lim == np.float32(2.0 / (1.0 + np.sqrt(1.0 - rho * rho)))
        assert np.array_equal(O.chebyshev_omegas(rho, 10, start=7), O.chebyshev_omegas(rho, 17)[7:])
        # past the recurrence's fixed point every factor is the limit (and a
        # far start returns at once)
        assert np.array_equal(O.chebyshev_omegas(rho, 5, start=4000), O.chebyshev_omegas(rho, 4005)[4000:])
        assert np.all(O.chebyshev_omegas(rho, 3, start=2 ** 40) == lim)


def test_scheduled_loop_pass_factors(oracle):
    """A constant schedule is the fixed-omega loop bit for bit; and pass 2t is
    the red pass: a factor of 0 leaves that colour's cells as they were."""
    O = oracle
    p = small_city(O, 20, 16, 9)
    R = O.divergence_f32(p)
    ref = O.sor_rb_f32(p, R, None, omega=1.6, itermax=7, tol=-1.0)[0]
    assert np.array_equal(O.sor_rb_sched_f32(p, R, np.full(14, 1.6, np.float32)), ref)
    NX, NY, NZ = p.nx - 1, p.ny - 1, p.nz - 1
    k, j, i = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
    inner = (i >= 1) & (i <= NX - 2) & (j >= 1) & (j <= NY - 2) & (k >= 1) & (k <= NZ - 2)
    red = (inner & ((i + j + k) % 2 == 0)).ravel()
    black = (inner & ((i + j + k) % 2 == 1)).ravel()
    a = O.sor_rb_sched_f32(p, R, np.array([0.0, 1.0], np.float32))
    assert np.all(a[red] == 0) and np.count_nonzero(a[black]) > 0
    b = O.sor_rb_sched_f32(p, R, np.array([1.0, 0.0], np.float32))
    assert np.all(b[black] == 0) and np.count_nonzero(b[red]) > 0


# ---------------------------------------------------------------------------
# inputs: grid, flags, walls, coefficients, profiles (closed forms)
# ---------------------------------------------------------------------------
def test_vertical_grid_closed_form(oracle):
    O = oracle
    dz = O.dz_stretched(1.0 * 1.02 ** np.arange(30))
    z_face, z = O.vertical_grid(dz)
    d = dz.astype(np.float64)
    zf = np.concatenate([[-d[0], 0.0], np.cumsum(d[1:])])
    assert np.allclose(z_face, zf, rtol=1e-6, atol=1e-6)
    assert np.allclose(z[1:], 0.5 * (zf[1:-1] + zf[2:]), rtol=1e-6)
    assert z[0] == np.float32(-0.5) * dz[0]


def test_cube_flags_and_walls_closed_form(oracle):
    """The single 20 m cube of the 100 x 100 x 50 case: x, y in [40, 60] m, 20 m high, dx = dz = 1:
    solid cells exactly i, j in [40, 59], k in [1, 20]; one zeroed
    coefficient per fluid face of the cube; ground walls n = 0 at k = 1."""
    O = oracle
    p = O.build_city(100, 100, 1.0, O.dz_uniform(50, 1.0), [(40.0, 40.0, 20.0, 20.0, 20.0, 0.0, 0.0)],
                     (2, 5.0, 10.0, 0.15, 270.0))
    NX, NY, NZ = 100, 100, 51
    fl = p.icellflag.reshape(NZ, NY, NX)
    expect = np.ones_like(fl)
    expect[0] = 2
    expect[1:21, 40:60, 40:60] = 0
    assert np.array_equal(fl, expect)
    e, f, g, h, m, n = (a.reshape(NZ, NY, NX) for a in p.coeffs)
    one = np.float32(1.0)
    # fluid neighbours of the cube faces
    assert np.all(e[1:21, 40:60, 39] == 0) and np.all(f[1:21, 40:60, 60] == 0)
    assert np.all(g[1:21, 39, 40:60] == 0) and np.all(h[1:21, 60, 40:60] == 0)
    assert np.all(n[21, 40:60, 40:60] == 0)
    assert np.all(n[1][fl[1] == 1] == 0)
    zero_counts = [(a[1:NZ - 1] == 0).sum() for a in (e, f, g, h, m, n)]
    assert zero_counts[:4] == [400, 400, 400, 400]
    assert zero_counts[4] == 0 and zero_counts[5] == NX * NY - 400 + 400
    # normalised values
    assert np.all(e[1:NZ - 1][e[1:NZ - 1] != 0] == one) and np.all(m[1:NZ - 1][m[1:NZ - 1] != 0] == one)


def test_stretched_coefficients_closed_form(oracle):
    O = oracle
    dzc = 1.0 * 1.02 ** np.arange(20)
    dz = O.dz_stretched(dzc)
    p = O.build_city(10, 8, 2.0, dz, [], (1, 4.0, 10.0, 0.5, 270.0))
    NX, NY, NZ = 10, 8, 21
    e, f, g, h, m, n = (a.reshape(NZ, NY, NX).astype(np.float64) for a in p.coeffs)
    d = dz.astype(np.float64)
    for k in range(1, NZ - 1):
        assert np.allclose(m[k], 1.0 / (d[k] * 0.5 * (d[k] + d[k + 1])), rtol=1e-6)
        if k > 1:
            assert np.allclose(n[k], 1.0 / (d[k] * 0.5 * (d[k] + d[k - 1])), rtol=1e-6)
        assert np.allclose(e[k], 0.25, rtol=0) and np.allclose(h[k, 1:], 0.25, rtol=0)
    assert np.all(m[NZ - 1] == 1.0) and np.all(e[0] == 1.0)  # un-normalised layers


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_profiles_closed_form(oracle, kind):
    O = oracle
    dz = O.dz_stretched(1.0 * 1.03 ** np.arange(25))
    z_face, z = O.vertical_grid(dz)
    wind = 200.0
    p = O.build_city(6, 5, 2.0, dz, [], (kind, 4.0, 10.0, {0: 0.1, 1: 0.5, 2: 0.15}[kind], wind))
    nx, ny, nz = p.nx, p.ny, p.nz
    U = p.u0.reshape(nz, ny, nx)
    V = p.v0.reshape(nz, ny, nx)
    th = np.deg2rad(270.0 - wind)
    zz = z.astype(np.float64)
    if kind == 0:
        mag = np.full(nz - 1, 4.0)
    elif kind == 1:
        ustar = 4.0 * 0.4 / np.log((10.0 + 0.5) / 0.5)
        mag = ustar / 0.4 * np.log((zz + 0.5) / 0.5)
    else:
        mag = 4.0 * (zz / 10.0) ** 0.15
    for k in range(1, nz - 1):
        assert np.allclose(U[k, 2:4, 2:4], np.cos(th) * mag[k], rtol=2e-6, atol=1e-6)
        assert np.allclose(V[k, 2:4, 2:4], np.sin(th) * mag[k], rtol=2e-6, atol=1e-6)
    assert np.all(U[0] == 0) and np.all(U[nz - 1] == 0) and np.all(p.w0 == 0)


def test_set_velocity_zero_only_inside_solids(oracle):
    O = oracle
    p = O.build_city(20, 16, 1.0, O.dz_uniform(8, 1.0), [(5.0, 4.0, 6.0, 5.0, 3.0, 0.0, 0.0)],
                     (0, 3.0, 10.0, 0.1, 270.0))
    nx, ny, nz = p.nx, p.ny, p.nz
    U = p.u0.reshape(nz, ny, nx)
    W = p.w0.reshape(nz, ny, nx)
    # building cells i in [5,10], j in [4,8], k in [1,3]: faces i = 5..11 zero
    assert np.all(U[1:4, 4:9, 5:12] == 0)
    assert np.all(U[1:4, 4:9, 4] == np.float32(-np.cos(0.0) * 3.0) * -1)
    # ground: w0 at k = 1 zeroed for interior i, j (the k = 0 ghost is solid)
    assert np.all(W[1, 1:ny - 1, 1:nx - 1] == 0)


def _linear_faces(O, nx, ny, nz, dx, dy, dz_array, A):
    """Face values of the linear field (u, v, w) = A @ (x, y, z) at the
    staggered positions: u at west faces, v at south faces, w at bottom faces
    (z of u, v: cell centres; z of w: cell faces)."""
    z_face, z = O.vertical_grid(dz_array)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    zc = np.append(z, z[-1] + dz_array[-1]).astype(np.float64)[k]
    zf = z_face.astype(np.float64)[k]
    xs = {"u": (i * dx, (j + 0.5) * dy, zc), "v": ((i + 0.5) * dx, j * dy, zc), "w": ((i + 0.5) * dx, (j + 0.5) * dy, zf)}
    out = []
    for row, name in zip(A, "uvw"):
        X, Y, Z = xs[name]
        out.append((row[0] * X + row[1] * Y + row[2] * Z).astype(np.float32).ravel())
    return out


@pytest.mark.parametrize("stretched,dy", [(False, 2.0), (True, 2.0), (True, 3.25)])
def test_velocity_gradient_exact_for_linear_fields(oracle, stretched, dy):
    """TURBGeneralData.cpp:479-552 restated: every component of a linear
    field's gradient is recovered (distinct coefficients, so a swapped
    component or operand fails), only fluid cells in the icellfluid range are
    written."""
    O = oracle
    ncx, ncy, ncz, dx = 14, 11, 9, 2.0
    dz = O.dz_stretched(np.geomspace(1.0, 3.0, ncz).astype(np.float32)) if stretched else O.dz_uniform(ncz, 1.5)
    p = O.build_city(ncx, ncy, dx, dz, [(8.0, 6.0, 6.0, 4.0, 5.0, 0.0, 0.0)], (1, 5.0, 10.0, 0.1, 270.0))
    p = O.Problem(p.nx, p.ny, p.nz, dx, dy, p.dz_array, *p.coeffs, p.icellflag, p.u0, p.v0, p.w0)
    A = np.array([[0.7, -1.3, 0.45], [2.1, 0.35, -0.8], [-0.55, 1.6, 0.25]])
    u, v, w = _linear_faces(O, p.nx, p.ny, p.nz, p.dx, p.dy, p.dz_array, A)
    G = O.velocity_gradient(p, u, v, w, uniform_z=not stretched, dz=1.5)
    NX, NY, NZ = p.nx - 1, p.ny - 1, p.nz - 1
    flag = p.icellflag.reshape(NZ, NY, NX)
    sel = np.zeros((NZ, NY, NX), bool)
    sel[1:p.nz - 2, 1:p.ny - 2, 1:p.nx - 2] = True
    sel &= (flag != 0) & (flag != 2)
    assert sel.sum() > 100 and (~sel & (flag == 0)).any()
    expect = A.ravel()  # Gxx Gxy Gxz Gyx Gyy Gyz Gzx Gzy Gzz
    for n in range(9):
        g = G[n].reshape(NZ, NY, NX)
        assert np.allclose(g[sel], expect[n], rtol=0, atol=2e-4), (n, np.abs(g[sel] - expect[n]).max())
        assert np.all(g[~sel] == 0)


# ---------------------------------------------------------------------------
# the post-correction divergence norm (oracle_div_linf)
# ---------------------------------------------------------------------------
def _with_dy(O, p, dy):
    """The same problem with a different dy (build_city sets dy = dx)."""
    return O.Problem(p.nx, p.ny, p.nz, p.dx, dy, p.dz_array, *p.coeffs, p.icellflag, p.u0, p.v0, p.w0)


def _linear_div_faces(p, a, b, c, offsets=False, seed=0):
    """u = a x_face, v = b y_face, w = c z_face (face (i,j,k) sits at x = i dx,
    y = j dy, z = z_face[k]); with offsets, plus an arbitrary function of the
    two other indices in each component (which every difference cancels)."""
    O_zf = np.concatenate([[-p.dz_array[0]], [0.0], np.cumsum(p.dz_array[1:].astype(np.float64))])
    K, J, I = np.meshgrid(np.arange(p.nz), np.arange(p.ny), np.arange(p.nx), indexing="ij")
    u = a * I * p.dx
    v = b * J * p.dy
    w = c * O_zf[K]
    if offsets:
        rng = np.random.default_rng(seed)
        u = u + rng.uniform(-3, 3, (p.nz, p.ny, 1))
        v = v + rng.uniform(-3, 3, (p.nz, 1, p.nx))
        w = w + rng.uniform(-3, 3, (1, p.ny, p.nx))
    return [np.ascontiguousarray(x.ravel(), np.float32) for x in (u, v, w)], O_zf


def _probe_cells(p, rng, nrandom=150):
    """Interior fluid cells to probe: every one next to a zeroed coefficient
    plus a random sample of the others."""
    NX, NY, NZ = p.nx - 1, p.ny - 1, p.nz - 1
    fl = p.icellflag.reshape(NZ, NY, NX)
    co = [a.reshape(NZ, NY, NX) for a in p.coeffs]
    sel = np.zeros_like(fl, bool)
    sel[1:NZ - 2, 1:NY - 2, 1:NX - 2] = True
    sel &= fl == 1
    walled = sel & np.any([x == 0 for x in co], axis=0)
    plain = np.argwhere(sel & ~walled)
    picks = [tuple(x) for x in np.argwhere(walled)]
    picks += [tuple(x) for x in plain[rng.choice(len(plain), min(nrandom, len(plain)), replace=False)]]
    return picks


def _single_cell_div(O, p, u, v, w, cell, physical):
    """oracle_div_linf with only `cell` (k, j, i) marked fluid: |div| of that cell alone."""
    k, j, i = cell
    NX, NY = p.nx - 1, p.ny - 1
    flag = np.zeros(p.ncell, np.int32)
    flag[i + j * NX + k * NX * NY] = 1
    q = O.Problem(p.nx, p.ny, p.nz, p.dx, p.dy, p.dz_array, *p.coeffs, flag, p.u0, p.v0, p.w0)
    return O.div_linf(q, u, v, w, physical=physical)


def test_div_linf_physical_operator_closed_form(oracle):
    """Physical divergence of u = a x, v = b y, w = c z on a stretched grid
    with dx != dy and buildings: exactly a + b + c in every interior fluid
    cell (wall-adjacent ones included, the physical operator ignores the
    coefficients), with arbitrary offsets that the differences cancel.  A
    dropped /dz_k, a swapped dx/dy, a sign error or a shifted face index in
    any term changes the value.  Single-cell flag arrays isolate each cell."""
    O = oracle
    p = _with_dy(O, small_city(O, 26, 20, 12, dx=2.0), 3.0)
    a, b, c = 0.75, -0.5, 1.25
    (u, v, w), _ = _linear_div_faces(p, a, b, c, offsets=True, seed=5)
    rng = np.random.default_rng(1)
    cells = _probe_cells(p, rng)
    assert len(cells) > 200
    for cell in cells:
        d = _single_cell_div(O, p, u, v, w, cell, True)
        assert abs(d - abs(a + b + c)) <= 2e-5, (cell, d)
    # over the whole domain the norm is the same value
    assert abs(O.div_linf(p, u, v, w, physical=True) - 1.5) <= 2e-5


def test_div_linf_reference_operator_closed_form(oracle):
    """Reference operator R(u)/(-2) (Solver_CPU_RB.cpp:56-80) of the same
    linear field: with e..n either 0 or their normalised values 1/dx^2 (g, h
    too: quirk A2) and 1/(dz_k (dz_k + dz_k+-1)/2) (Wall.cpp:827-846), each
    axis contributes
        x: a ((i+1) [e != 0] - i [f != 0])
        y: b (dy/dx)^2 ((j+1) [g != 0] - j [h != 0])
        z: c (z_face[k+1] [m != 0] - z_face[k] [n != 0])
    so an interior cell gives a + b (dy/dx)^2 + c dz_k (the vertical term is
    the plain difference of w, quirk A1).  A swapped e/f (or g/h, m/n), a
    dropped dx or dz factor or a wrong face index changes the value at the
    wall-adjacent cells, which are all probed."""
    O = oracle
    p = _with_dy(O, small_city(O, 26, 20, 12, dx=2.0), 3.0)
    a, b, c = 0.75, -0.5, 1.25
    (u, v, w), zf = _linear_div_faces(p, a, b, c)
    NX, NY, NZ = p.nx - 1, p.ny - 1, p.nz - 1
    e, f, g, h, m, n = (x.reshape(NZ, NY, NX) != 0 for x in p.coeffs)
    r2 = (p.dy / p.dx) ** 2
    rng = np.random.default_rng(2)
    cells = _probe_cells(p, rng)
    checked_walls = 0
    for (k, j, i) in cells:
        expect = (a * ((i + 1) * e[k, j, i] - i * f[k, j, i]) + b * r2 * ((j + 1) * g[k, j, i] - j * h[k, j, i])
                  + c * (zf[k + 1] * m[k, j, i] - zf[k] * n[k, j, i]))
        d = _single_cell_div(O, p, u, v, w, (k, j, i), False)
        assert abs(d - abs(expect)) <= 1e-5 * max(1.0, abs(expect)), ((k, j, i), d, expect)
        checked_walls += not (e[k, j, i] and f[k, j, i] and g[k, j, i] and h[k, j, i] and m[k, j, i] and n[k, j, i])
    assert checked_walls > 100
    # interior cells: a + b (dy/dx)^2 + c dz_k
    dzk = p.dz_array[5]
    d = _single_cell_div(O, p, u, v, w, (5, 3, 3), False)
    assert abs(d - abs(a + b * r2 + c * dzk)) <= 1e-5


def test_div_linf_matches_the_fp64_divergence_source(oracle):
    """On an arbitrary field the reference-operator norm equals max |R| / 2 of
    the FP64 divergence source (oracle_divergence_f64, pinned by the
    reference's stored outputs and the divergence-free invariant) over the
    interior fluid cells, bitwise, and the numpy restatement ref_operator_div
    to rounding."""
    O = oracle
    p = small_city(O, 30, 24, 14)
    rng = np.random.default_rng(9)
    u, v, w = (rng.uniform(-5, 5, p.nface).astype(np.float32) for _ in range(3))
    q = O.Problem(p.nx, p.ny, p.nz, p.dx, p.dy, p.dz_array, *p.coeffs, p.icellflag, u, v, w)
    R = O.divergence_f64(q)
    NX, NY, NZ = p.nx - 1, p.ny - 1, p.nz - 1
    fl = p.icellflag.reshape(NZ, NY, NX)
    sel = np.zeros_like(fl, bool)
    sel[1:NZ - 2, 1:NY - 2, 1:NX - 2] = True
    sel &= fl == 1
    expect = np.abs(R.reshape(NZ, NY, NX)[sel]).max() / 2.0
    got = O.div_linf(p, u, v, w, physical=False)
    assert got == expect
    D = ref_operator_div(p, u, v, w)
    assert abs(np.abs(D[sel]).max() - got) <= 1e-12 * got


# ---------------------------------------------------------------------------
# rotated buildings (RectangularBuilding vertex list + PNPOLY)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("rot", [15.0, 30.0, 45.0, 120.0, 200.0, 333.0])
def test_rotated_building_matches_half_plane_test(oracle, rot):
    """A rectangle with corner (x0, y0), side L along (cos r, sin r) and W
    along (-sin r, cos r) (RectangularBuilding.h:86-103): a cell centre is
    inside iff both of its coordinates along those axes lie in [0, L] and
    [0, W].  Cells whose centre is within 1e-3 m of an edge are skipped (the
    crossing-number test decides those by float rounding); the layers are
    k in [1, k_end) with z[k_end - 1] <= H < z[k_end] (PolyBuilding.cpp:
    160-175)."""
    O = oracle
    ncx, ncy, dx = 60, 50, 1.5
    dz = O.dz_stretched(1.0 * 1.04 ** np.arange(16))
    x0, y0, L, W, H = 40.0, 25.0, 23.0, 11.0, 9.3
    p = O.build_city(ncx, ncy, dx, dz, [(x0, y0, L, W, H, 0.0, rot)], (0, 3.0, 10.0, 0.1, 270.0))
    NX, NY, NZ = ncx, ncy, dz.size
    fl = p.icellflag.reshape(NZ, NY, NX)
    _, z = O.vertical_grid(dz)
    r = np.deg2rad(rot)
    J, I = np.meshgrid(np.arange(NY), np.arange(NX), indexing="ij")
    px, py = (I + 0.5) * dx - x0, (J + 0.5) * dx - y0
    s = px * np.cos(r) + py * np.sin(r)
    t = -px * np.sin(r) + py * np.cos(r)
    inside = (s >= 0) & (s <= L) & (t >= 0) & (t <= W)
    near = (np.minimum.reduce([np.abs(s), np.abs(s - L), np.abs(t), np.abs(t - W)]) < 1e-3)
    k_end = int(np.argmax(z[1:] > H)) + 1
    assert 2 < k_end < NZ - 1
    for k in range(1, NZ - 1):
        want = inside & (k < k_end)
        got = fl[k] == 0
        assert np.array_equal(got[~near], want[~near]), (k, np.argwhere(got != want)[:5])
    assert inside.sum() > 100  # about L W / dx^2 = 112 cells per layer at dx 1.5: not vacuous


# ---------------------------------------------------------------------------
# Wall::wallLogBC on stair-step walls (oracle_wall_log_bc)
# ---------------------------------------------------------------------------
def _log_law_closed_form(V, d1, d2, z0, kappa=0.4, steps=20):
    """Final wall-parallel speed of the 20-step fixed point of Wall.cpp:677-797
    in closed form: u*_{n+1} = a - b u*_n with a = kappa V / ln(d1/z0),
    b = ln(d2/d1) / ln(d1/z0), u*_0 = 0.1, so
    u*_n = a (1 - (-b)^n) / (1 + b) + (-b)^n 0.1, and the last step writes
    V - u*_{steps-1} ln(d2/d1) / kappa."""
    L1, L2 = np.log(d1 / z0), np.log(d2 / d1)
    a, b = kappa * V / L1, L2 / L1
    n = steps - 1
    us = a * (1 - (-b) ** n) / (1 + b) + (-b) ** n * 0.1
    return V - us * L2 / kappa


def _box_city(O, dz_kind, z0):
    """One box building on flat ground, random wall-parallel velocities on
    every face (so each pass's result depends on which neighbour it reads)."""
    ncx, ncy, ncz, dx = 24, 20, 12, 2.0
    dz = O.dz_uniform(ncz, 1.0) if dz_kind == "uniform" else O.dz_stretched(1.0 * 1.1 ** np.arange(ncz))
    p = O.build_city(ncx, ncy, dx, dz, [(16.0, 12.0, 14.0, 10.0, 6.5, 0.0, 0.0)], (1, 4.0, 10.0, z0, 250.0))
    rng = np.random.default_rng(3)
    u, v, w = (rng.uniform(0.5, 4.0, p.nface).astype(np.float32) * s for s in (1.0, -1.0, 0.5))
    NX, NY, NZ = p.nx - 1, p.ny - 1, p.nz - 1
    solid = (p.icellflag.reshape(NZ, NY, NX)[1:] == 0)
    ks, js, is_ = np.nonzero(solid)
    box = (is_.min(), is_.max() + 1, js.min(), js.max() + 1, ks.max() + 1)  # i0, i1, j0, j1, top layer
    assert solid.sum() == (box[1] - box[0]) * (box[3] - box[2]) * box[4]  # a box
    return p, u, v, w, box


@pytest.mark.parametrize("dz_kind,z0", [("uniform", 0.1), ("stretched", 0.1), ("uniform", 0.3)])
def test_wall_log_bc_closed_form(oracle, dz_kind, z0):
    """Each of the six passes against the closed form of its fixed point, on a
    face that only that pass touches, with the neighbour the reference reads
    (Wall.cpp:681-797: +nx*ny, -nx*ny, +1, -1, +nx, -nx) and the components it
    writes (normal 0, parallel = speed x cos/sin of the neighbour's
    direction); the face counts of a box building (every k = 1 fluid cell has
    the ground below it); a face in two lists takes the later pass's result
    computed from the neighbour the earlier pass rewrote (pass order); faces in
    no list are untouched.  z0 = 0.3 m makes the recurrence diverge (b > 1),
    which pins the step count."""
    O = oracle
    p, u, v, w, (i0, i1, j0, j1, H) = _box_city(O, dz_kind, z0)
    nx, ny = p.nx, p.ny
    NX, NY = nx - 1, ny - 1
    fz = nx * ny
    F = lambda i, j, k: i + j * nx + k * fz  # noqa: E731
    dz = float(np.min(p.dz_array))
    dx = p.dx
    U, Vv, Wv, nfaces = O.wall_log_bc(p, u, v, w, z0)
    L, Wd = i1 - i0, j1 - j0
    assert nfaces == NX * NY + 2 * H * (L + Wd)

    def check(f, nb, comps, normal, d, src=(u, v, w), tol=2e-5):
        A, B = (src[c][nb].astype(np.float64) for c in comps)
        V = np.hypot(A, B)
        m = _log_law_closed_form(V, 0.5 * d, 1.5 * d, z0)
        th = np.arctan2(B, A)
        out = (U, Vv, Wv)
        assert out[normal][f] == 0.0
        for c, want in zip(comps, (m * np.cos(th), m * np.sin(th))):
            assert abs(out[c][f] - want) <= tol * max(1.0, abs(want)), (f, c, out[c][f], want)

    jm, km = (j0 + j1) // 2, max(2, H // 2)
    im = (i0 + i1) // 2
    check(F(3, 4, 1), F(3, 4, 2), (0, 1), 2, dz)                 # below (ground)
    check(F(im, jm, H + 1), F(im, jm, H + 2), (0, 1), 2, dz)     # below (roof)
    check(F(i1, jm, km), F(i1 + 1, jm, km), (1, 2), 0, dx)       # back: solid at i-1
    check(F(i0 - 1, jm, km), F(i0 - 2, jm, km), (1, 2), 0, dx)   # front: solid at i+1
    check(F(im, j1, km), F(im, j1 + 1, km), (0, 2), 1, dx)       # right: solid at j-1
    check(F(im, j0 - 1, km), F(im, j0 - 2, km), (0, 2), 1, dx)   # left: solid at j+1
    # a k = 1 cell beside the building is in the below and the back list: the
    # back pass reads its neighbour after the below pass rewrote it
    nb = F(i1 + 1, jm, 1)
    mid = [a.copy() for a in (u, v, w)]
    A, B = (mid[c][F(i1 + 1, jm, 2)].astype(np.float64) for c in (0, 1))
    m = _log_law_closed_form(np.hypot(A, B), 0.5 * dz, 1.5 * dz, z0)
    th = np.arctan2(B, A)
    mid[0][nb], mid[1][nb], mid[2][nb] = m * np.cos(th), m * np.sin(th), 0.0
    check(F(i1, jm, 1), nb, (1, 2), 0, dx, src=mid, tol=1e-4)
    # faces of cells in no list keep their values
    far = F(3, 4, 5)
    assert U[far] == u[far] and Vv[far] == v[far] and Wv[far] == w[far]


def test_wall_log_bc_refuses_cut_cells(oracle):
    O = oracle
    p, u, v, w, _ = _box_city(O, "uniform", 0.1)
    flag = p.icellflag.copy()
    flag[flag.size // 2] = 7
    q = O.Problem(p.nx, p.ny, p.nz, p.dx, p.dy, p.dz_array, *p.coeffs, flag, p.u0, p.v0, p.w0)
    with pytest.raises(ValueError):
        O.wall_log_bc(q, u, v, w, 0.1)


@pytest.mark.parametrize("wind", [45.0, 135.0, 333.0])
def test_fp64_solution_is_linear_in_the_wind_direction(oracle, wind):
    """The premise of the large-domain FP64 fixture: the discrete problem is
    linear in (u0, v0, w0) and a single-sensor inflow from direction d is
    cos(t) x the inflow from 270 deg + sin(t) x the inflow from 180 deg,
    t = 270 - d (WindProfilerSensorType.cpp:571-621), so the converged FP64
    field of any direction is that combination of the two basis fields, up to
    the fp32 rounding of the inflow values (~1e-7 relative)."""
    O = oracle
    fields = {}
    for d in (270.0, 180.0, wind):
        p = small_city(O, 30, 24, 14, wind=d)
        u, v, w, lam, its, me = O.solve_rb_f64(p, tol=1e-11)
        fields[d] = np.stack([u, v, w])
    th = np.radians(270.0 - wind)
    comb = np.cos(th) * fields[270.0] + np.sin(th) * fields[180.0]
    scale = np.abs(fields[wind]).max()
    assert np.abs(comb - fields[wind]).max() <= 2e-6 * scale
    # and a wrong combination (sign of the second basis flipped) is far off
    bad = np.cos(th) * fields[270.0] - np.sin(th) * fields[180.0]
    assert np.abs(bad - fields[wind]).max() > 1e-2 * scale


# ---------------------------------------------------------------------------
# f4: wall corrections of the turbulence velocity gradient (stair-step)
# ---------------------------------------------------------------------------
def _turb_wall_city(O):
    """A box building on flat ground (flags from the builder), then flags set
    by hand so that every pass has cells: a block in the air (walls above), a
    one-cell gap in x and in y between solid columns (walls on both sides),
    a fluid cell between a roof and the block (walls below and above), and
    terrain cells (flag 2) next to building cells."""
    ncx, ncy, ncz, dx = 24, 20, 12, 2.0
    dz = O.dz_stretched(1.0 * 1.1 ** np.arange(ncz))
    p = O.build_city(ncx, ncy, dx, dz, [(16.0, 12.0, 10.0, 8.0, 4.5, 0.0, 0.0)], (1, 4.0, 10.0, 0.1, 250.0))
    NX, NY, NZ = p.nx - 1, p.ny - 1, p.nz - 1
    f = p.icellflag.reshape(NZ, NY, NX).copy()
    f[6:8, 3:6, 3:7] = 0          # floating block at k = 6, 7: walls above the cells under it
    f[4, 3:6, 3:7] = 0            # a slab at k = 4: the cells at k = 5 have walls below and above
    f[1:4, 10, 14] = 0            # solid columns at i = 14 and i = 16 around a gap at i = 15
    f[1:4, 10, 16] = 0
    f[1:4, 14, 4] = 0             # solid columns at j = 14 and j = 16 around a gap at j = 15
    f[1:4, 16, 4] = 0
    f[1:3, 2, 18:21] = 2          # terrain cells, and a building cell beside them at i = 17
    f[1:3, 2, 17] = 0
    f[1:3, 2, 22] = 0             # a fluid cell at i = 21 between terrain (i = 20) and building (i = 22)
    q = O.Problem(p.nx, p.ny, p.nz, p.dx, p.dy, p.dz_array, *p.coeffs, f.ravel().astype(np.int32), p.u0, p.v0,
                  p.w0)
    return q, f


def _turb_lists(f, cf):
    """TURBWall::get_stairstep_wall_id's six lists (TURBWall.cpp:76-118) as
    sets of flat cell ids: i in [1,NX-1], j in [1,NY-1], k in [1,NZ-2], flat
    neighbours -sz, +sz, -1, +1, -NX, +NX."""
    NZ, NY, NX = f.shape
    flat = f.ravel()
    sz = NX * NY
    out = [set() for _ in range(6)]
    for k in range(1, NZ - 1):
        for j in range(1, NY):
            for i in range(1, NX):
                c = i + j * NX + k * sz
                if flat[c] in (0, 2):
                    continue
                for p_, off in enumerate((-sz, sz, -1, 1, -NX, NX)):
                    if flat[c + off] == cf:
                        out[p_].add(c)
    return out


def test_turb_wall_finite_differences_of_linear_fields(oracle):
    """Mode 2 (TURBWall.cpp:206-300): one-sided differences of fields linear
    in x, y and z give the slopes exactly (to float rounding) in every pass's
    cells, and cells in no list keep the centred gradient."""
    O = oracle
    p, f = _turb_wall_city(O)
    NZ, NY, NX = f.shape
    nx, ny, nz = p.nx, p.ny, p.nz
    _, z = O.vertical_grid(p.dz_array)
    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    zc = z[np.minimum(K, nz - 2)].astype(np.float64)
    a, al, b, be, c, ga = 0.7, -0.3, 0.45, 0.25, -0.6, 0.35
    u = (a * zc + al * J * p.dx).astype(np.float32).ravel()
    v = (b * I * p.dx + be * zc).astype(np.float32).ravel()
    w = (c * I * p.dx + ga * J * p.dx).astype(np.float32).ravel()
    G0 = O.velocity_gradient(p, u, v, w)
    G = O.turb_wall_stairstep(p, G0, u, v, w, 2, 2)
    want = {1: al, 2: a, 3: b, 5: be, 6: c, 7: ga}   # Gxy, Gxz, Gyx, Gyz, Gzx, Gzy
    lists = [a_ | b_ for a_, b_ in zip(_turb_lists(f, 0), _turb_lists(f, 2))]
    touched = {1: lists[4] | lists[5], 2: lists[0] | lists[1], 3: lists[2] | lists[3], 5: lists[0] | lists[1],
               6: lists[2] | lists[3], 7: lists[4] | lists[5]}
    for comp, slope in want.items():
        cells = np.array(sorted(touched[comp]))
        assert cells.size > 0, comp
        np.testing.assert_allclose(G[comp][cells], slope, rtol=2e-4, atol=2e-4, err_msg=str(comp))
        rest = np.setdiff1d(np.arange(p.ncell), cells)
        assert np.array_equal(G[comp][rest], G0[comp][rest]), comp
    for comp in (0, 4, 8):
        assert np.array_equal(G[comp], G0[comp])


def test_turb_wall_later_passes_win(oracle):
    """Mode 2 with quadratic fields, where the forward and backward one-sided
    differences differ: a cell with walls below and above takes the backward
    (above) difference, one between back and front walls the front one, one
    between right and left walls the left one, and a cell with a building
    behind and terrain in front the terrain object's front pass (terrain
    after buildings)."""
    O = oracle
    p, f = _turb_wall_city(O)
    NZ, NY, NX = f.shape
    nx, ny, nz = p.nx, p.ny, p.nz
    _, z = O.vertical_grid(p.dz_array)
    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    zc = z[np.minimum(K, nz - 2)].astype(np.float64)
    dx = float(p.dx)
    u = (0.3 * zc ** 2).astype(np.float32).ravel()
    v = (0.02 * (I * dx) ** 2).astype(np.float32).ravel()
    w = (0.02 * (J * dx) ** 2).astype(np.float32).ravel()
    G = O.turb_wall_stairstep(p, O.velocity_gradient(p, u, v, w), u, v, w, 2, 2)
    cell = lambda i, j, k: i + j * NX + k * NX * NY  # noqa: E731
    Lb, Lt = _turb_lists(f, 0), _turb_lists(f, 2)
    # between the slab (k = 4) and the block (k = 6): the backward (above) difference
    sand = sorted(Lb[0] & Lb[1])
    assert len(sand) == 3 * 4
    for cs in sand:
        kk = cs // (NX * NY)
        backward = 0.3 * (z[kk] + z[kk - 1])
        forward = 0.3 * (z[kk + 1] + z[kk])
        assert abs(G[2][cs] - backward) <= 2e-4 * abs(backward) and abs(forward - backward) > 1e-2
    # between the columns at i = 14 and 16: the front difference
    cx = cell(15, 10, 2)
    assert cx in Lb[2] and cx in Lb[3]
    back = 0.02 * dx * (2 * 15 + 1)     # (v(i+1) - v(i)) / dx
    front = 0.02 * dx * (2 * 15 - 1)    # (v(i) - v(i-1)) / dx
    assert abs(G[3][cx] - front) <= 1e-4 * abs(front) and abs(G[3][cx] - back) > 1e-3
    # between the columns at j = 14 and 16: the left difference
    cy = cell(4, 15, 2)
    assert cy in Lb[4] and cy in Lb[5]
    right = 0.02 * dx * (2 * 15 + 1)
    left = 0.02 * dx * (2 * 15 - 1)
    assert abs(G[7][cy] - left) <= 1e-4 * abs(left) and abs(G[7][cy] - right) > 1e-3
    # terrain behind (i = 20) and a building in front (i = 22) of cell (21, 2, k): the building
    # object's front pass runs first, the terrain object's back pass after it and wins
    for k in (1, 2):
        ct = cell(21, 2, k)
        assert ct in Lb[3] and ct in Lt[2]
        tb = 0.02 * dx * (2 * 21 + 1)   # the back (forward) difference
        assert abs(G[3][ct] - tb) <= 1e-4 * tb


def test_turb_wall_log_law_closed_form(oracle):
    """Mode 1 (TURBWallBuilding.cpp:87-169 and the terrain variant) on uniform
    fields: each wall-normal derivative is q / (0.5 d ln(0.5 d / z0)) with
    d = dz_k, dx, dy = dx; the sixth pass ("left of the cell") walks the
    cells with a wall below, so a cell with a wall only on its left keeps its
    centred Gxy, and every cell with a wall below gets the y-formula; the
    terrain object takes z0 from the mean of the four surrounding
    z0_domain_u / z0_domain_v entries, summed in float and divided by 4.0."""
    O = oracle
    p, f = _turb_wall_city(O)
    NZ, NY, NX = f.shape
    U, V, W, z0 = 3.0, -2.0, 0.5, 0.1
    u = np.full(p.nface, U, np.float32)
    v = np.full(p.nface, V, np.float32)
    w = np.full(p.nface, W, np.float32)
    rng = np.random.default_rng(7)
    z0u = rng.uniform(0.05, 0.2, p.nx * p.ny).astype(np.float32)
    z0v = rng.uniform(0.05, 0.2, p.nx * p.ny).astype(np.float32)
    G0 = O.velocity_gradient(p, u, v, w)
    G = O.turb_wall_stairstep(p, G0, u, v, w, 1, 1, z0, z0u, z0v)
    dx = np.float64(np.float32(p.dx))
    dz = p.dz_array.astype(np.float64)

    def ll(q, d, zz):
        return np.float32(0.5 * np.float64(np.float32(q) + np.float32(q)) / (0.5 * d * np.log(0.5 * d / np.float64(np.float32(zz)))))

    def z0_terrain(c):
        k, r = divmod(c, NX * NY)
        j, i = divmod(r, NX)
        s = np.float32(np.float32(np.float32(z0u[i + j * p.nx] + z0u[i + 1 + j * p.nx]) + z0v[i + j * p.nx])
                       + z0v[i + (j + 1) * p.nx])
        return np.float32(np.float64(s) / 4.0)

    Lb, Lt = _turb_lists(f, 0), _turb_lists(f, 2)
    # the final value of each component comes from the last object whose pass list holds the cell
    for comp, passes, qv, dfun in ((2, (0, 1), U, lambda c: dz[c // (NX * NY)]),
                                   (5, (0, 1), V, lambda c: dz[c // (NX * NY)]),
                                   (3, (2, 3), V, lambda c: dx), (6, (2, 3), W, lambda c: dx),
                                   (1, (4, 0), U, lambda c: dx), (7, (4, 0), W, lambda c: dx)):
        cells_b = set().union(*(Lb[q] for q in passes))
        cells_t = set().union(*(Lt[q] for q in passes))
        assert cells_b, comp
        for c in cells_b | cells_t:
            want = ll(qv, dfun(c), z0_terrain(c) if c in cells_t else z0)
            assert G[comp][c] == want, (comp, c, G[comp][c], want)
        rest = np.setdiff1d(np.arange(p.ncell), np.array(sorted(cells_b | cells_t)))
        assert np.array_equal(G[comp][rest], G0[comp][rest]), comp
    # the quirk, explicitly: a wall on the left only leaves Gxy alone
    left_only = (Lb[5] | Lt[5]) - (Lb[4] | Lt[4] | Lb[0] | Lt[0])
    assert left_only
    for c in left_only:
        assert G[1][c] == G0[1][c]

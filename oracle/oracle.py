"""ctypes wrapper over the CPU oracle (oracle/qesmc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the reference they compare against.  The product library
(libqesmc) never imports this module.

Arrays follow the reference's layouts (src/qes/Domain.h:143-175): cell arrays
of (nx-1)*(ny-1)*(nz-1) entries and face arrays of nx*ny*nz entries, x fastest.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle_qesmc.so")
_lib = None

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    I, Fl, D = C.c_int, C.c_float, C.c_double
    L.oracle_divergence_f32.argtypes = [I, I, I, Fl, Fl, f32p] + [f32p] * 6 + [f32p] * 3 + [I, f32p]
    L.oracle_divergence_f64.argtypes = [I, I, I, D, D, f32p] + [f32p] * 6 + [f32p] * 3 + [I, f64p]
    L.oracle_sor_rb_f32.argtypes = [I, I, I] + [f32p] * 6 + [f32p, f32p, f32p, Fl, Fl, I, C.POINTER(Fl)]
    L.oracle_sor_rb_f32.restype = I
    L.oracle_sor_serial_f32.argtypes = L.oracle_sor_rb_f32.argtypes
    L.oracle_sor_serial_f32.restype = I
    L.oracle_sor_rb_sched_f32.argtypes = [I, I, I] + [f32p] * 6 + [f32p, f32p, f32p, I]
    L.oracle_sor_rb_f64.argtypes = [I, I, I] + [f32p] * 6 + [f64p, f64p, f64p, D, D, I, C.POINTER(D)]
    L.oracle_sor_rb_f64.restype = I
    L.oracle_velocity_f32.argtypes = [I, I, I, Fl, Fl, f32p, f32p, f32p, f32p, i32p,
                                      f32p, f32p, f32p, f32p, I, I, f32p, f32p, f32p]
    L.oracle_velocity_f64.argtypes = [I, I, I, D, D, f32p, f32p, f32p, f32p, i32p,
                                      f32p, f32p, f32p, f64p, I, I, f64p, f64p, f64p]
    L.oracle_div_linf.argtypes = [I, I, I, D, D, f32p] + [f32p] * 6 + [i32p, f32p, f32p, f32p, I]
    L.oracle_div_linf.restype = D
    L.oracle_vertical_grid.argtypes = [I, f32p, f32p, f32p]
    L.oracle_init_cells.argtypes = [I, I, I] + [f32p] * 6 + [i32p]
    L.oracle_rect_building.argtypes = [I, I, I, Fl, Fl, f32p, f32p] + [Fl] * 7 + [i32p]
    L.oracle_define_walls.argtypes = [I, I, I, i32p] + [f32p] * 6
    L.oracle_solver_coefficients.argtypes = [I, I, I, Fl, f32p] + [f32p] * 6
    L.oracle_profile.argtypes = [I, f32p, I, Fl, Fl, Fl, Fl, f32p, f32p]
    L.oracle_extrude_profile.argtypes = [I, I, I, f32p, f32p, f32p, f32p, f32p]
    L.oracle_set_velocity_zero.argtypes = [I, I, I, i32p, f32p, f32p, f32p]
    L.oracle_velocity_gradient.argtypes = [I, I, I, Fl, Fl, Fl, f32p, I, i32p, f32p, f32p, f32p, f32p]
    L.oracle_wall_log_bc.argtypes = [I, I, I, Fl, Fl, Fl, i32p, f32p, f32p, f32p]
    L.oracle_wall_log_bc.restype = C.c_long
    L.oracle_turb_wall_stairstep.argtypes = [I, I, I, Fl, f32p, Fl, f32p, f32p, I, I, i32p, f32p, f32p, f32p, f32p]
    L.oracle_turb_wall_stairstep.restype = I
    L.oracle_num_threads.restype = I
    L.oracle_set_num_threads.argtypes = [I]
    _lib = L
    return L


# ----------------------------------------------------------------------------
# Problem container
# ----------------------------------------------------------------------------
class Problem:
    """Inputs of Solver::solve: the Domain sizes and the WINDSGeneralData
    arrays it reads (src/winds/Solver_CPU_RB.cpp:41-291)."""

    def __init__(self, nx, ny, nz, dx, dy, dz_array, e, f, g, h, m, n, icellflag, u0, v0, w0):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.dx, self.dy = float(dx), float(dy)
        self.dz_array = np.ascontiguousarray(dz_array, dtype=np.float32)
        self.e, self.f, self.g, self.h, self.m, self.n = (np.ascontiguousarray(a, dtype=np.float32)
                                                          for a in (e, f, g, h, m, n))
        self.icellflag = np.ascontiguousarray(icellflag, dtype=np.int32)
        self.u0, self.v0, self.w0 = (np.ascontiguousarray(a, dtype=np.float32) for a in (u0, v0, w0))
        assert self.dz_array.size == self.nz - 1
        assert self.e.size == self.ncell and self.u0.size == self.nface

    @property
    def ncell(self):
        return (self.nx - 1) * (self.ny - 1) * (self.nz - 1)

    @property
    def nface(self):
        return self.nx * self.ny * self.nz

    @property
    def coeffs(self):
        return (self.e, self.f, self.g, self.h, self.m, self.n)

    def interior_lups(self):
        """Cells updated per SOR iteration: (nx-3)(ny-3)(nz-3)."""
        return (self.nx - 3) * (self.ny - 3) * (self.nz - 3)


# ----------------------------------------------------------------------------
# Solver pieces
# ----------------------------------------------------------------------------
def divergence_f32(p: Problem, alpha1=1):
    R = np.zeros(p.ncell, np.float32)
    lib().oracle_divergence_f32(p.nx, p.ny, p.nz, p.dx, p.dy, p.dz_array, *p.coeffs,
                                p.u0, p.v0, p.w0, alpha1, R)
    return R


def divergence_f64(p: Problem, alpha1=1):
    R = np.zeros(p.ncell, np.float64)
    lib().oracle_divergence_f64(p.nx, p.ny, p.nz, p.dx, p.dy, p.dz_array, *p.coeffs,
                                p.u0, p.v0, p.w0, alpha1, R)
    return R


def sor_rb_f32(p: Problem, R, lam=None, omega=1.78, tol=1e-9, itermax=500):
    lam = np.zeros(p.ncell, np.float32) if lam is None else np.array(lam, np.float32, copy=True)
    old = np.zeros_like(lam)
    me = C.c_float(0)
    it = lib().oracle_sor_rb_f32(p.nx, p.ny, p.nz, *p.coeffs, np.ascontiguousarray(R, np.float32),
                                 lam, old, omega, tol, itermax, C.byref(me))
    return lam, it, me.value


def chebyshev_omegas(rho, passes, start=0):
    """The factors of colour passes start .. start+passes-1 of cyclic
    Chebyshev semi-iteration for a 2-cyclic system whose Jacobi matrix has
    spectral radius rho (Golub & Varga, Numer. Math. 3 (1961) 147-168; Varga,
    Matrix Iterative Analysis, sec. 5.2): w_0 = 1, w_1 = 1 / (1 - rho^2 / 2),
    w_h = 1 / (1 - rho^2 w_(h-1) / 4), computed in double and rounded to
    float32 (the SOR pass's type).  Pass 2t is iteration t's red pass."""
    r2 = float(rho) * float(rho)
    out, w = [], 1.0
    for h in range(start + passes):
        prev = w
        w = 1.0 if h == 0 else 1.0 / (1.0 - 0.5 * r2) if h == 1 else 1.0 / (1.0 - 0.25 * r2 * w)
        if h >= 2 and w == prev:  # the fixed point: every later factor is w
            out.extend([w] * (start + passes - max(h, start)))
            break
        if h >= start:
            out.append(w)
    return np.array(out, np.float32)


def sor_rb_sched_f32(p: Problem, R, omegas, lam=None):
    """len(omegas) // 2 red-black iterations, pass h at omegas[h]."""
    lam = np.zeros(p.ncell, np.float32) if lam is None else np.array(lam, np.float32, copy=True)
    om = np.ascontiguousarray(omegas, np.float32)
    lib().oracle_sor_rb_sched_f32(p.nx, p.ny, p.nz, *p.coeffs, np.ascontiguousarray(R, np.float32), lam, om,
                                  om.size // 2)
    return lam


def sor_serial_f32(p: Problem, R, lam=None, omega=1.78, tol=1e-9, itermax=500):
    lam = np.zeros(p.ncell, np.float32) if lam is None else np.array(lam, np.float32, copy=True)
    old = np.zeros_like(lam)
    me = C.c_float(0)
    it = lib().oracle_sor_serial_f32(p.nx, p.ny, p.nz, *p.coeffs, np.ascontiguousarray(R, np.float32),
                                     lam, old, omega, tol, itermax, C.byref(me))
    return lam, it, me.value


def sor_rb_f64(p: Problem, R, lam=None, omega=float(np.float32(1.78)), tol=1e-9, itermax=100000):
    lam = np.zeros(p.ncell, np.float64) if lam is None else np.array(lam, np.float64, copy=True)
    old = np.zeros_like(lam)
    me = C.c_double(0)
    it = lib().oracle_sor_rb_f64(p.nx, p.ny, p.nz, *p.coeffs, np.ascontiguousarray(R, np.float64),
                                 lam, old, omega, tol, itermax, C.byref(me))
    return lam, it, me.value


def velocity_f32(p: Problem, lam, alpha1=1, alpha2=1):
    u, v, w = (np.zeros(p.nface, np.float32) for _ in range(3))
    lib().oracle_velocity_f32(p.nx, p.ny, p.nz, p.dx, p.dy, p.dz_array, p.f, p.h, p.n, p.icellflag,
                              p.u0, p.v0, p.w0, np.ascontiguousarray(lam, np.float32),
                              alpha1, alpha2, u, v, w)
    return u, v, w


def velocity_f64(p: Problem, lam, alpha1=1, alpha2=1):
    u, v, w = (np.zeros(p.nface, np.float64) for _ in range(3))
    lib().oracle_velocity_f64(p.nx, p.ny, p.nz, p.dx, p.dy, p.dz_array, p.f, p.h, p.n, p.icellflag,
                              p.u0, p.v0, p.w0, np.ascontiguousarray(lam, np.float64),
                              alpha1, alpha2, u, v, w)
    return u, v, w


def solve_rb_f32(p: Problem, lam=None, omega=1.78, tol=1e-9, itermax=500, alpha1=1, alpha2=1):
    """Whole Solver_CPU_RB::solve: returns (u, v, w, lambda, iterations, max_error)."""
    R = divergence_f32(p, alpha1)
    lam, it, me = sor_rb_f32(p, R, lam, omega, tol, itermax)
    u, v, w = velocity_f32(p, lam, alpha1, alpha2)
    return u, v, w, lam, it, me


def solve_serial_f32(p: Problem, lam=None, omega=1.78, tol=1e-9, itermax=500, alpha1=1, alpha2=1):
    """Whole Solver_CPU::solve (lexicographic)."""
    R = divergence_f32(p, alpha1)
    lam, it, me = sor_serial_f32(p, R, lam, omega, tol, itermax)
    u, v, w = velocity_f32(p, lam, alpha1, alpha2)
    return u, v, w, lam, it, me


def solve_rb_f64(p: Problem, tol=1e-9, itermax=200000, omega=float(np.float32(1.78))):
    """FP64 converged oracle: returns (u, v, w, lambda, iterations, max_error) in double."""
    R = divergence_f64(p)
    lam, it, me = sor_rb_f64(p, R, None, omega, tol, itermax)
    u, v, w = velocity_f64(p, lam)
    return u, v, w, lam, it, me


def div_linf(p: Problem, u, v, w, physical=False):
    """max |div| over interior fluid cells, reference operator (R/(-2)) or physical."""
    return lib().oracle_div_linf(p.nx, p.ny, p.nz, p.dx, p.dy, p.dz_array, *p.coeffs, p.icellflag,
                                 np.ascontiguousarray(u, np.float32), np.ascontiguousarray(v, np.float32),
                                 np.ascontiguousarray(w, np.float32), 1 if physical else 0)


# ----------------------------------------------------------------------------
# Inputs: grid, stair-step buildings, walls, coefficients, inflow
# ----------------------------------------------------------------------------
def vertical_grid(dz_array):
    dz = np.ascontiguousarray(dz_array, np.float32)
    nz = dz.size + 1
    z_face = np.zeros(nz, np.float32)
    z = np.zeros(nz - 1, np.float32)
    lib().oracle_vertical_grid(nz, dz, z_face, z)
    return z_face, z


def dz_uniform(ncz, dz):
    """Domain::defineVerticalStretching(float), Domain.cpp:123-131."""
    return np.full(ncz + 1, dz, np.float32)


def dz_stretched(dz_cells):
    """Domain::defineVerticalStretching(vector), Domain.cpp:133-144: the ncz
    layer thicknesses with dz_array[0] = dz_array[1] for the ghost cell."""
    dz_cells = np.asarray(dz_cells, np.float32)
    out = np.zeros(dz_cells.size + 1, np.float32)
    out[1:] = dz_cells
    out[0] = out[1]
    return out


def build_city(ncx, ncy, dx, dz_array, buildings=(), profile=(1, 5.0, 10.0, 0.1, 270.0)):
    """The WINDSGeneralData set-up for a flat, stair-step city:
    allocateMemory + ground flags, RectangularBuilding/PolyBuilding cell flags,
    Wall::defineWalls + solverCoefficients, single-sensor profile extrusion and
    Wall::setVelocityZero (the call order of WINDSGeneralData.cpp:63-912 and
    qesWindsMain.cpp:140-158).

    buildings: iterable of (x_start, y_start, L, W, H, base_height, rotation_deg)
    profile:   (kind, U_ref, z_ref, z0_or_p, wind_dir), kind 0 uniform, 1 log, 2 power.
    """
    L = lib()
    dz_array = np.ascontiguousarray(dz_array, np.float32)
    nx, ny, nz = ncx + 1, ncy + 1, dz_array.size + 1
    ncell = (nx - 1) * (ny - 1) * (nz - 1)
    nface = nx * ny * nz
    z_face, z = vertical_grid(dz_array)
    coefs = [np.empty(ncell, np.float32) for _ in range(6)]
    flag = np.empty(ncell, np.int32)
    L.oracle_init_cells(nx, ny, nz, *coefs, flag)
    for b in buildings:
        xs, ys, bl, bw, bh, base, rot = (list(b) + [0.0, 0.0])[:7]
        L.oracle_rect_building(nx, ny, nz, dx, dx, z_face, z, xs, ys, bl, bw, bh, base, rot, flag)
    L.oracle_define_walls(nx, ny, nz, flag, *coefs)
    L.oracle_solver_coefficients(nx, ny, nz, dx, dz_array, *coefs)
    up = np.zeros(nz, np.float32)
    vp = np.zeros(nz, np.float32)
    kind, U_ref, z_ref, z0p, wdir = profile
    L.oracle_profile(nz, z, int(kind), U_ref, z_ref, z0p, wdir, up, vp)
    u0, v0, w0 = (np.zeros(nface, np.float32) for _ in range(3))
    L.oracle_extrude_profile(nx, ny, nz, up, vp, u0, v0, w0)
    L.oracle_set_velocity_zero(nx, ny, nz, flag, u0, v0, w0)
    return Problem(nx, ny, nz, dx, dx, dz_array, *coefs, flag, u0, v0, w0)


def velocity_gradient(p: Problem, u, v, w, uniform_z=False, dz=None):
    """TURBGeneralData::derivativeVelocity (TURBGeneralData.cpp:479-552):
    returns a (9, ncell) array Gxx, Gxy, Gxz, Gyx, Gyy, Gyz, Gzx, Gzy, Gzz."""
    G = np.empty((9, p.ncell), np.float32)
    dzv = float(p.dz_array[1]) if dz is None else float(dz)
    lib().oracle_velocity_gradient(p.nx, p.ny, p.nz, p.dx, p.dy, dzv, p.dz_array, 1 if uniform_z else 0,
                                   p.icellflag, np.ascontiguousarray(u, np.float32),
                                   np.ascontiguousarray(v, np.float32), np.ascontiguousarray(w, np.float32), G)
    return G


def wall_log_bc(p: Problem, u, v, w, z0, dz=None):
    """Wall::wallLogBC on a solved stair-step field (Wall.cpp:659-797, applied
    as qesWindsMain.cpp:157-165 does: u0 = u, wallLogBC, u = u0).  Returns
    (u, v, w, number of wall faces); dz defaults to the Domain's dz, the
    smallest dz_array entry (Domain.cpp:89,131)."""
    u, v, w = (np.array(a, np.float32, copy=True) for a in (u, v, w))
    dzv = float(np.min(p.dz_array)) if dz is None else float(dz)
    n = lib().oracle_wall_log_bc(p.nx, p.ny, p.nz, p.dx, dzv, float(z0), p.icellflag, u, v, w)
    if n < 0:
        raise ValueError("cut-cell flags (7/8) present: only the stair-step passes are restated")
    return u, v, w, int(n)


def turb_wall_stairstep(p: Problem, G, u, v, w, building_mode=2, terrain_mode=2, z0=0.1, z0_domain_u=None,
                        z0_domain_v=None):
    """The wall step of TURBGeneralData::derivativeVelocity on stair-step
    walls (TURBGeneralData.cpp:553-556): buildings (flag 0) then terrain
    (flag 2), each in its mode (1 log law, 2 one-sided differences, 0 none).
    Returns a corrected copy of G (9, ncell).  z0_domain_u / z0_domain_v
    (nx*ny each) are needed only for the terrain log law."""
    G = np.array(G, np.float32, copy=True)
    if terrain_mode == 1 and (z0_domain_u is None or z0_domain_v is None):
        raise ValueError("the terrain log law needs z0_domain_u and z0_domain_v")
    z0u = np.ascontiguousarray(np.zeros(1) if z0_domain_u is None else z0_domain_u, np.float32)
    z0v = np.ascontiguousarray(np.zeros(1) if z0_domain_v is None else z0_domain_v, np.float32)
    rc = lib().oracle_turb_wall_stairstep(p.nx, p.ny, p.nz, p.dx, p.dz_array, float(z0), z0u, z0v,
                                          int(building_mode), int(terrain_mode), p.icellflag,
                                          np.ascontiguousarray(u, np.float32), np.ascontiguousarray(v, np.float32),
                                          np.ascontiguousarray(w, np.float32), G)
    if rc < 0:
        raise ValueError("cut-cell flags (7/8) present or a mode outside 0..2")
    return G


def set_num_threads(t: int):
    lib().oracle_set_num_threads(int(t))


def num_threads() -> int:
    return lib().oracle_num_threads()

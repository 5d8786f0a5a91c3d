"""Python binding of libqesmc (include/qesmc.h) over ctypes.

The product is the C ABI and its HIP kernels; this module only marshals numpy
arrays into it.  It never falls back to a CPU implementation: if libqesmc.so
is missing, or no HIP device is present, the calls raise.

Layouts follow QES-Winds (src/qes/Domain.h:143-175): nx = ncx+1, ny = ncy+1,
nz = ncz+2; cell arrays (nx-1)(ny-1)(nz-1), face arrays nx*ny*nz, x fastest.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# QESMC_LIB may point at another build of the same library (A/B timing runs)
LIB_PATH = os.environ.get("QESMC_LIB") or os.path.join(_HERE, "libqesmc.so")

OK = 0
ERR_ARG = -1
ERR_NO_DEVICE = -2
ERR_OOM = -3
ERR_HIP = -4
ERR_ZERO_DIAG = -5
ERR_NONFINITE = -6
ERR_STATE = -7
ERR_UNSUPPORTED = -8
ERR_IO = -9
ERR_LIMIT = -10

PATH_AUTO = 0
PATH_PLAIN = 1
PATH_SPLIT = 2
PATH_FUSED = 3
RUN_REENCODE = 1

# every symbol include/qesmc.h declares
EXPORTS = (
    "qesmc_create", "qesmc_destroy", "qesmc_default_params", "qesmc_status_string",
    "qesmc_solve", "qesmc_set_geometry", "qesmc_set_inflow", "qesmc_set_lambda", "qesmc_run",
    "qesmc_get_velocity", "qesmc_get_lambda", "qesmc_divergence_linf", "qesmc_city_build",
    "qesmc_city_random_buildings", "qesmc_run_static", "qesmc_refine",
)


class QesmcError(RuntimeError):
    def __init__(self, code: int, where: str):
        self.code = code
        super().__init__(f"{where}: {status_string(code)} ({code})")


_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dx", C.c_float),
                ("dy", C.c_float), ("dz_array", _fp)]


class Coeffs(C.Structure):
    _fields_ = [("e", _fp), ("f", _fp), ("g", _fp), ("h", _fp), ("m", _fp), ("n", _fp),
                ("icellflag", _ip)]


class Fields(C.Structure):
    _fields_ = [("u0", _fp), ("v0", _fp), ("w0", _fp), ("u", _fp), ("v", _fp), ("w", _fp)]


class Params(C.Structure):
    _fields_ = [("omega", C.c_float), ("tol", C.c_float), ("itermax", C.c_int),
                ("alpha1", C.c_int), ("alpha2", C.c_int), ("path", C.c_int),
                ("check_every", C.c_int), ("flags", C.c_int), ("tile_k", C.c_int),
                ("tile_v", C.c_int), ("wg_per_cu", C.c_int), ("slab_planes", C.c_int),
                ("stop_extend", C.c_float), ("refine_every", C.c_int),
                ("chebyshev_rho", C.c_double), ("chebyshev_start", C.c_int)]


class Result(C.Structure):
    _fields_ = [("iterations", C.c_int), ("path", C.c_int), ("compressed", C.c_int),
                ("status", C.c_int), ("final_max_dlambda", C.c_double), ("ms_total", C.c_double),
                ("ms_sor", C.c_double), ("glups", C.c_double), ("tile_k", C.c_int),
                ("tile_v", C.c_int), ("wg_per_cu", C.c_int), ("reserved", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Building(C.Structure):
    _fields_ = [("x_start", C.c_float), ("y_start", C.c_float), ("length", C.c_float),
                ("width", C.c_float), ("height", C.c_float), ("base_height", C.c_float),
                ("rotation_deg", C.c_float)]


class Inflow(C.Structure):
    _fields_ = [("kind", C.c_int), ("U_ref", C.c_float), ("z_ref", C.c_float),
                ("z0_or_p", C.c_float), ("wind_dir_deg", C.c_float)]


_lib = None


def lib():
    """Load libqesmc.so (built by `make -C qes-public_amd/csrc`)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libqesmc.so not built: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    L.qesmc_create.argtypes = [C.POINTER(P), C.c_int]
    L.qesmc_destroy.argtypes = [P]
    L.qesmc_default_params.argtypes = [C.POINTER(Params)]
    L.qesmc_default_params.restype = None
    L.qesmc_status_string.argtypes = [C.c_int]
    L.qesmc_status_string.restype = C.c_char_p
    L.qesmc_solve.argtypes = [P, C.POINTER(Grid), C.POINTER(Coeffs), C.POINTER(Fields), _fp,
                              C.POINTER(Params), C.POINTER(Result)]
    L.qesmc_set_geometry.argtypes = [P, C.POINTER(Grid), C.POINTER(Coeffs)]
    L.qesmc_set_inflow.argtypes = [P, _fp, _fp, _fp]
    L.qesmc_set_lambda.argtypes = [P, _fp]
    L.qesmc_run.argtypes = [P, C.POINTER(Params), C.POINTER(Result)]
    L.qesmc_run_static.argtypes = [P, C.POINTER(Params), C.POINTER(Result)]
    L.qesmc_refine.argtypes = [P, C.POINTER(Params), C.POINTER(Result)]
    L.qesmc_get_velocity.argtypes = [P, _fp, _fp, _fp]
    L.qesmc_get_lambda.argtypes = [P, _fp]
    L.qesmc_divergence_linf.argtypes = [P, C.c_int, C.POINTER(C.c_double)]
    L.qesmc_city_build.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, _fp, C.c_int,
                                   C.POINTER(Building), C.POINTER(Inflow)] + [_fp] * 6 + [_ip] + [_fp] * 3
    L.qesmc_city_random_buildings.argtypes = [C.c_ulonglong, C.c_int, C.c_int, C.c_float, C.c_int,
                                              C.c_float, C.c_float, C.c_float, C.c_float, C.c_int,
                                              C.c_float, C.c_int, C.POINTER(Building)]
    for name in EXPORTS:
        fn = getattr(L, name)
        if name not in ("qesmc_default_params", "qesmc_status_string"):
            fn.restype = C.c_int
    _lib = L
    return L


def status_string(code: int) -> str:
    return lib().qesmc_status_string(int(code)).decode()


def _f(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_fp)


def _i(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def _check(rc, where):
    if rc != OK:
        raise QesmcError(rc, where)


def default_params(**kw) -> Params:
    p = Params()
    lib().qesmc_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class City:
    """The solver's inputs (WINDSGeneralData arrays + Domain sizes)."""

    def __init__(self, nx, ny, nz, dx, dy, dz_array, e, f, g, h, m, n, icellflag, u0, v0, w0):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.dx, self.dy = float(dx), float(dy)
        c = np.ascontiguousarray
        self.dz_array = c(dz_array, np.float32)
        self.e, self.f, self.g, self.h, self.m, self.n = (c(a, np.float32) for a in (e, f, g, h, m, n))
        self.icellflag = c(icellflag, np.int32)
        self.u0, self.v0, self.w0 = (c(a, np.float32) for a in (u0, v0, w0))

    @property
    def ncell(self):
        return (self.nx - 1) * (self.ny - 1) * (self.nz - 1)

    @property
    def nface(self):
        return self.nx * self.ny * self.nz

    def interior_lups(self):
        return (self.nx - 3) * (self.ny - 3) * (self.nz - 3)

    def arrays(self):
        return dict(nx=self.nx, ny=self.ny, nz=self.nz, dx=self.dx, dy=self.dy,
                    dz_array=self.dz_array, e=self.e, f=self.f, g=self.g, h=self.h, m=self.m,
                    n=self.n, icellflag=self.icellflag, u0=self.u0, v0=self.v0, w0=self.w0)


def _buildings(buildings):
    bl = list(buildings)
    barr = (Building * max(1, len(bl)))()
    for q, b in enumerate(bl):
        vals = (list(b) + [0.0, 0.0])[:7] if not isinstance(b, Building) else None
        barr[q] = b if vals is None else Building(*[float(v) for v in vals])
    return barr, len(bl)


def _inflow(inflow):
    return Inflow(int(inflow[0]), *[float(v) for v in inflow[1:5]])


def city_build(ncx, ncy, dx, dz_array, buildings=(), inflow=(1, 5.0, 10.0, 0.1, 270.0)) -> City:
    """Flat stair-step city through qesmc_city_build."""
    dz_array = np.ascontiguousarray(dz_array, np.float32)
    nx, ny, nz = ncx + 1, ncy + 1, dz_array.size + 1
    ncell, nface = (nx - 1) * (ny - 1) * (nz - 1), nx * ny * nz
    barr, nb = _buildings(buildings)
    inf = _inflow(inflow)
    co = [np.empty(ncell, np.float32) for _ in range(6)]
    flag = np.empty(ncell, np.int32)
    u0, v0, w0 = (np.empty(nface, np.float32) for _ in range(3))
    rc = lib().qesmc_city_build(nx, ny, nz, float(dx), _f(dz_array), nb, barr, C.byref(inf),
                                *[_f(a) for a in co], _i(flag), _f(u0), _f(v0), _f(w0))
    _check(rc, "qesmc_city_build")
    return City(nx, ny, nz, dx, dx, dz_array, *co, flag, u0, v0, w0)


def random_buildings(seed, ncx, ncy, dx, count, lmin=10.0, lmax=40.0, hmin=10.0, hmax=60.0,
                     lognormal=False, hcap=120.0, margin=2):
    out = (Building * max(1, count))()
    n = lib().qesmc_city_random_buildings(int(seed), ncx + 1, ncy + 1, float(dx), int(count),
                                          lmin, lmax, hmin, hmax, 1 if lognormal else 0, hcap,
                                          int(margin), out)
    if n < 0:
        raise QesmcError(n, "qesmc_city_random_buildings")
    return [(b.x_start, b.y_start, b.length, b.width, b.height, b.base_height, b.rotation_deg)
            for b in out[:n]]


def dz_uniform(ncz, dz):
    return np.full(ncz + 1, dz, np.float32)


def dz_geometric(ncz, dz0, ratio):
    """ncz layers growing geometrically from dz0, ghost dz_array[0] = dz_array[1]."""
    cells = (dz0 * np.float64(ratio) ** np.arange(ncz)).astype(np.float32)
    out = np.empty(ncz + 1, np.float32)
    out[1:] = cells
    out[0] = out[1]
    return out


class Context:
    """One libqesmc context (device buffers + a HIP stream)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        _check(lib().qesmc_create(C.byref(self._h), int(device)), "qesmc_create")
        self.city = None

    def close(self):
        if self._h:
            lib().qesmc_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_geometry(self, city: City):
        g = Grid(city.nx, city.ny, city.nz, city.dx, city.dy, _f(city.dz_array))
        co = Coeffs(*[_f(a) for a in (city.e, city.f, city.g, city.h, city.m, city.n)], _i(city.icellflag))
        _check(lib().qesmc_set_geometry(self._h, C.byref(g), C.byref(co)), "qesmc_set_geometry")
        self.city = city

    def set_inflow(self, u0, v0, w0):
        u0, v0, w0 = (np.ascontiguousarray(a, np.float32) for a in (u0, v0, w0))
        _check(lib().qesmc_set_inflow(self._h, _f(u0), _f(v0), _f(w0)), "qesmc_set_inflow")

    def set_lambda(self, lam=None):
        if lam is None:
            _check(lib().qesmc_set_lambda(self._h, None), "qesmc_set_lambda")
        else:
            lam = np.ascontiguousarray(lam, np.float32)
            _check(lib().qesmc_set_lambda(self._h, _f(lam)), "qesmc_set_lambda")

    def run(self, params: Params | None = None, allow_nonfinite=False, **kw) -> dict:
        p = params if params is not None else default_params(**kw)
        r = Result()
        rc = lib().qesmc_run(self._h, C.byref(p), C.byref(r))
        if rc != OK and not (allow_nonfinite and rc == ERR_NONFINITE):
            raise QesmcError(rc, "qesmc_run")
        return r.as_dict()

    def run_static(self, params: Params | None = None, allow_nonfinite=False, **kw) -> dict:
        """qesmc_run without device allocation (qesmc_run_static): itermax <= 65536."""
        p = params if params is not None else default_params(**kw)
        r = Result()
        rc = lib().qesmc_run_static(self._h, C.byref(p), C.byref(r))
        if rc != OK and not (allow_nonfinite and rc == ERR_NONFINITE):
            raise QesmcError(rc, "qesmc_run_static")
        return r.as_dict()

    def refine(self, params: Params | None = None, allow_nonfinite=False, **kw) -> dict:
        """One mixed-precision refinement of the resident lambda (qesmc_refine):
        FP64 residual, SOR on the correction from zero, lambda += correction,
        velocity correction."""
        p = params if params is not None else default_params(**kw)
        r = Result()
        rc = lib().qesmc_refine(self._h, C.byref(p), C.byref(r))
        if rc != OK and not (allow_nonfinite and rc == ERR_NONFINITE):
            raise QesmcError(rc, "qesmc_refine")
        return r.as_dict()

    def get_velocity(self):
        n = self.city.nface
        u, v, w = (np.empty(n, np.float32) for _ in range(3))
        _check(lib().qesmc_get_velocity(self._h, _f(u), _f(v), _f(w)), "qesmc_get_velocity")
        return u, v, w

    def get_lambda(self):
        lam = np.empty(self.city.ncell, np.float32)
        _check(lib().qesmc_get_lambda(self._h, _f(lam)), "qesmc_get_lambda")
        return lam

    def divergence_linf(self, physical=False) -> float:
        out = C.c_double()
        _check(lib().qesmc_divergence_linf(self._h, 1 if physical else 0, C.byref(out)),
               "qesmc_divergence_linf")
        return out.value

    def set_city(self, ncx, ncy, dx, dz_array, buildings=(), inflow=(1, 5.0, 10.0, 0.1, 270.0)):
        """Build a flat stair-step city on the device (qesmc_set_city)."""
        dz_array = np.ascontiguousarray(dz_array, np.float32)
        barr, nb = _buildings(buildings)
        inf = _inflow(inflow)
        nx, ny, nz = ncx + 1, ncy + 1, dz_array.size + 1
        _check(_city_lib().qesmc_set_city(self._h, nx, ny, nz, float(dx), _f(dz_array), nb, barr, C.byref(inf)),
               "qesmc_set_city")
        self.city = CityShape(nx, ny, nz, dx, dx, dz_array)
        return self.city

    def set_city_inflow(self, inflow):
        """New sensor profile on the resident geometry (qesmc_set_city_inflow)."""
        inf = _inflow(inflow)
        _check(_city_lib().qesmc_set_city_inflow(self._h, C.byref(inf)), "qesmc_set_city_inflow")

    def get_inputs(self) -> City:
        """The resident inputs as a City (qesmc_get_inputs)."""
        s = self.city
        co = [np.empty(s.ncell, np.float32) for _ in range(6)]
        flag = np.empty(s.ncell, np.int32)
        u0, v0, w0 = (np.empty(s.nface, np.float32) for _ in range(3))
        _check(_city_lib().qesmc_get_inputs(self._h, *[_f(a) for a in co], _i(flag), _f(u0), _f(v0), _f(w0)),
               "qesmc_get_inputs")
        return City(s.nx, s.ny, s.nz, s.dx, s.dy, s.dz_array, *co, flag, u0, v0, w0)

    def velocity_gradient(self, uniform_z=False, dz=None, copy=True):
        """Velocity gradient tensor of the resident corrected field
        (qesmc_velocity_gradient): ((9, ncell) array or None, kernel ms)."""
        dzv = float(self.city.dz_array[1]) if dz is None else float(dz)
        G = np.empty((9, self.city.ncell), np.float32) if copy else None
        ms = C.c_double()
        _check(_city_lib().qesmc_velocity_gradient(self._h, 1 if uniform_z else 0, dzv,
                                                   G.ctypes.data_as(_fp) if copy else None, C.byref(ms)),
               "qesmc_velocity_gradient")
        return G, ms.value

    def velocity_gradient_walls(self, building_mode=2, terrain_mode=2, z0=0.1, z0_domain_u=None, z0_domain_v=None,
                                copy=True):
        """The wall step of the gradient on stair-step walls
        (qesmc_velocity_gradient_walls), on the gradient of the last
        velocity_gradient call: ((9, ncell) array or None, kernel ms)."""
        G = np.empty((9, self.city.ncell), np.float32) if copy else None
        zu = None if z0_domain_u is None else np.ascontiguousarray(z0_domain_u, np.float32)
        zv = None if z0_domain_v is None else np.ascontiguousarray(z0_domain_v, np.float32)
        ms = C.c_double()
        _check(_city_lib().qesmc_velocity_gradient_walls(
            self._h, int(building_mode), int(terrain_mode), float(z0), _f(zu) if zu is not None else None,
            _f(zv) if zv is not None else None, G.ctypes.data_as(_fp) if copy else None, C.byref(ms)),
            "qesmc_velocity_gradient_walls")
        return G, ms.value

    def estimate_omega(self, steps=200):
        """(rho_J, omega) of the resident problem (qesmc_estimate_omega)."""
        rho, om = C.c_double(), C.c_float()
        _check(_city_lib().qesmc_estimate_omega(self._h, int(steps), C.byref(rho), C.byref(om)),
               "qesmc_estimate_omega")
        return rho.value, om.value

    def wall_log_bc(self, z0) -> float:
        """Wall::wallLogBC on the resident corrected field (qesmc_wall_log_bc);
        returns the device ms of the six passes."""
        ms = C.c_double()
        _check(_city_lib().qesmc_wall_log_bc(self._h, float(z0), C.byref(ms)), "qesmc_wall_log_bc")
        return ms.value

    def solve(self, city: City, lam=None, params: Params | None = None, **kw):
        """One-call drop-in (qesmc_solve): returns (u, v, w, lambda, result)."""
        p = params if params is not None else default_params(**kw)
        g = Grid(city.nx, city.ny, city.nz, city.dx, city.dy, _f(city.dz_array))
        co = Coeffs(*[_f(a) for a in (city.e, city.f, city.g, city.h, city.m, city.n)], _i(city.icellflag))
        u, v, w = (np.empty(city.nface, np.float32) for _ in range(3))
        fl = Fields(_f(city.u0), _f(city.v0), _f(city.w0), _f(u), _f(v), _f(w))
        lam_io = np.zeros(city.ncell, np.float32) if lam is None else np.array(lam, np.float32, copy=True)
        r = Result()
        rc = lib().qesmc_solve(self._h, C.byref(g), C.byref(co), C.byref(fl), _f(lam_io), C.byref(p),
                               C.byref(r))
        _check(rc, "qesmc_solve")
        self.city = city
        return u, v, w, lam_io, r.as_dict()


EXPORTS = EXPORTS + ("qesmc_run_batch", "qesmc_autotune")


def _batch_lib():
    L = lib()
    if not getattr(L, "_batch_sigs_set", False):
        L.qesmc_run_batch.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(Params), C.POINTER(Result), C.c_int,
                                      C.POINTER(C.c_double)]
        L.qesmc_run_batch.restype = C.c_int
        L.qesmc_autotune.argtypes = [C.c_void_p, C.POINTER(Params), C.c_int, C.POINTER(Params)]
        L.qesmc_autotune.restype = C.c_int
        L._batch_sigs_set = True
    return L


def run_batch(ctxs, params: Params | None = None, use_graph=True, **kw):
    """qesmc_run_batch over resident contexts: ([result dict], loop ms)."""
    p = params if params is not None else default_params(**kw)
    n = len(ctxs)
    hs = (C.c_void_p * n)(*[c._h.value for c in ctxs])
    res = (Result * n)()
    ms = C.c_double()
    rc = _batch_lib().qesmc_run_batch(hs, n, C.byref(p), res, 1 if use_graph else 0, C.byref(ms))
    _check(rc, "qesmc_run_batch")
    return [r.as_dict() for r in res], ms.value


def autotune(ctx: "Context", iterations=40, params: Params | None = None, **kw) -> Params:
    p = params if params is not None else default_params(**kw)
    best = Params()
    _check(_batch_lib().qesmc_autotune(ctx._h, C.byref(p), int(iterations), C.byref(best)), "qesmc_autotune")
    return best


EXPORTS = EXPORTS + ("qesmc_set_city", "qesmc_set_city_inflow", "qesmc_get_inputs", "qesmc_lambda_combine",
                     "qesmc_velocity_linf_diff", "qesmc_velocity_gradient", "qesmc_set_probes",
                     "qesmc_get_probes", "qesmc_wall_log_bc", "qesmc_estimate_omega",
                     "qesmc_velocity_gradient_walls")


class CityShape:
    """Sizes of a city that lives only on the device."""

    def __init__(self, nx, ny, nz, dx, dy, dz_array):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.dx, self.dy = float(dx), float(dy)
        self.dz_array = np.ascontiguousarray(dz_array, np.float32)

    ncell = City.ncell
    nface = City.nface
    interior_lups = City.interior_lups


def _city_lib():
    L = lib()
    if not getattr(L, "_city_sigs_set", False):
        L.qesmc_set_city.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_float, _fp, C.c_int,
                                     C.POINTER(Building), C.POINTER(Inflow)]
        L.qesmc_set_city.restype = C.c_int
        L.qesmc_set_city_inflow.argtypes = [C.c_void_p, C.POINTER(Inflow)]
        L.qesmc_set_city_inflow.restype = C.c_int
        L.qesmc_get_inputs.argtypes = [C.c_void_p] + [_fp] * 6 + [_ip] + [_fp] * 3
        L.qesmc_get_inputs.restype = C.c_int
        L.qesmc_lambda_combine.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), _fp]
        L.qesmc_lambda_combine.restype = C.c_int
        L.qesmc_velocity_linf_diff.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]
        L.qesmc_velocity_linf_diff.restype = C.c_int
        L.qesmc_velocity_gradient.argtypes = [C.c_void_p, C.c_int, C.c_float, _fp, C.POINTER(C.c_double)]
        L.qesmc_velocity_gradient.restype = C.c_int
        L.qesmc_velocity_gradient_walls.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_float, _fp, _fp, _fp,
                                                    C.POINTER(C.c_double)]
        L.qesmc_velocity_gradient_walls.restype = C.c_int
        L.qesmc_set_probes.argtypes = [C.c_void_p, C.c_longlong, C.POINTER(C.c_longlong)]
        L.qesmc_set_probes.restype = C.c_int
        L.qesmc_get_probes.argtypes = [C.c_void_p, _fp, _fp, _fp]
        L.qesmc_get_probes.restype = C.c_int
        L.qesmc_estimate_omega.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_float)]
        L.qesmc_estimate_omega.restype = C.c_int
        L.qesmc_wall_log_bc.argtypes = [C.c_void_p, C.c_float, C.POINTER(C.c_double)]
        L.qesmc_wall_log_bc.restype = C.c_int
        L._city_sigs_set = True
    return L


EXPORTS = EXPORTS + ("qesmc_workspace_write", "qesmc_workspace_dims", "qesmc_workspace_read")


def _ws_lib():
    L = lib()
    if not getattr(L, "_ws_sigs_set", False):
        L.qesmc_workspace_write.argtypes = [C.c_char_p, C.POINTER(Grid), C.POINTER(Coeffs), C.POINTER(Fields)]
        L.qesmc_workspace_write.restype = C.c_int
        L.qesmc_workspace_dims.argtypes = [C.c_char_p] + [C.POINTER(C.c_int)] * 3
        L.qesmc_workspace_dims.restype = C.c_int
        L.qesmc_workspace_read.argtypes = ([C.c_char_p, C.c_int, C.c_int, C.c_int, _fp, _fp, _fp] + [_fp] * 6 + [_ip]
                                           + [_fp] * 6)
        L.qesmc_workspace_read.restype = C.c_int
        L._ws_sigs_set = True
    return L


def workspace_write(path, nx, ny, nz, dx, dy, dz_array, coeffs, icellflag, u0, v0, w0, u, v, w):
    """Write a QES-Winds workspace file, NetCDF classic (qesmc_workspace_write).
    coeffs = (e, f, g, h, m, n) cell arrays; u0..w face arrays."""
    dz_array = np.ascontiguousarray(dz_array, np.float32)
    co = [np.ascontiguousarray(a, np.float32) for a in coeffs]
    flag = np.ascontiguousarray(icellflag, np.int32)
    fl = [np.ascontiguousarray(a, np.float32) for a in (u0, v0, w0, u, v, w)]
    g = Grid(int(nx), int(ny), int(nz), float(dx), float(dy), _f(dz_array))
    cs = Coeffs(*[_f(a) for a in co], _i(flag))
    fs = Fields(*[_f(a) for a in fl])
    _check(_ws_lib().qesmc_workspace_write(os.fsencode(path), C.byref(g), C.byref(cs), C.byref(fs)),
           "qesmc_workspace_write")


def workspace_dims(path):
    """Staggered (nx, ny, nz) of a workspace file (qesmc_workspace_dims)."""
    d = [C.c_int() for _ in range(3)]
    _check(_ws_lib().qesmc_workspace_dims(os.fsencode(path), *[C.byref(v) for v in d]), "qesmc_workspace_dims")
    return tuple(v.value for v in d)


def workspace_read(path, velocity=True):
    """Arrays of a workspace file (qesmc_workspace_read): a dict with nx, ny,
    nz, dx, dy, dz_array, e..n, icellflag, u0, v0, w0 and, with velocity,
    u, v, w."""
    nx, ny, nz = workspace_dims(path)
    ncell, nface = (nx - 1) * (ny - 1) * (nz - 1), nx * ny * nz
    dx, dy = C.c_float(), C.c_float()
    out = {"nx": nx, "ny": ny, "nz": nz, "dz_array": np.empty(nz - 1, np.float32)}
    for k in ("e", "f", "g", "h", "m", "n"):
        out[k] = np.empty(ncell, np.float32)
    out["icellflag"] = np.empty(ncell, np.int32)
    for k in ("u0", "v0", "w0") + (("u", "v", "w") if velocity else ()):
        out[k] = np.empty(nface, np.float32)
    vel = [_f(out[k]) if k in out else None for k in ("u", "v", "w")]
    _check(_ws_lib().qesmc_workspace_read(os.fsencode(path), nx, ny, nz, C.byref(dx), C.byref(dy),
                                          _f(out["dz_array"]), *[_f(out[k]) for k in ("e", "f", "g", "h", "m", "n")],
                                          _i(out["icellflag"]), _f(out["u0"]), _f(out["v0"]), _f(out["w0"]), *vel),
           "qesmc_workspace_read")
    out["dx"], out["dy"] = dx.value, dy.value
    return out


def lambda_combine(ctx: "Context", srcs, coefs):
    """lambda(ctx) = sum coefs[t] * lambda(srcs[t]) on the device (qesmc_lambda_combine)."""
    n = len(srcs)
    hs = (C.c_void_p * n)(*[c._h.value for c in srcs])
    cf = np.ascontiguousarray(coefs, np.float32)
    _check(_city_lib().qesmc_lambda_combine(ctx._h, n, hs, _f(cf)), "qesmc_lambda_combine")


def velocity_linf_diff(a: "Context", b: "Context") -> float:
    """max |(u,v,w)_a - (u,v,w)_b| over all faces, FP64 on the device."""
    out = C.c_double()
    _check(_city_lib().qesmc_velocity_linf_diff(a._h, b._h, C.byref(out)), "qesmc_velocity_linf_diff")
    return out.value


def set_probes(ctx: "Context", faces):
    """Resident list of face indices (i + j nx + k nx ny) whose u, v, w
    get_probes gathers on the device (qesmc_set_probes)."""
    idx = np.ascontiguousarray(faces, np.int64)
    _check(_city_lib().qesmc_set_probes(ctx._h, idx.size, idx.ctypes.data_as(C.POINTER(C.c_longlong))),
           "qesmc_set_probes")
    ctx._n_probe = int(idx.size)


def get_probes(ctx: "Context"):
    """u, v, w of the resident corrected field at the probe faces."""
    n = getattr(ctx, "_n_probe", 0)
    u, v, w = (np.empty(n, np.float32) for _ in range(3))
    _check(_city_lib().qesmc_get_probes(ctx._h, _f(u), _f(v), _f(w)), "qesmc_get_probes")
    return u, v, w


UNET_IN_CHANNELS = 8


class UNetConfig(C.Structure):
    _fields_ = [("levels", C.c_int), ("base_channels", C.c_int), ("ncx", C.c_int), ("ncy", C.c_int),
                ("ncz", C.c_int), ("seed", C.c_ulonglong), ("target_scale", C.c_float)]


_UNET_SIGS = {
    "qesmc_unet_create": [C.POINTER(C.c_void_p), C.c_int, C.POINTER(UNetConfig)],
    "qesmc_unet_destroy": [C.c_void_p],
    "qesmc_unet_num_params": [C.c_void_p, C.POINTER(C.c_longlong)],
    "qesmc_unet_get_params": [C.c_void_p, _fp, C.c_longlong],
    "qesmc_unet_set_params": [C.c_void_p, _fp, C.c_longlong],
    "qesmc_unet_dims": [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "qesmc_unet_forward": [C.c_void_p, _fp, _fp],
    "qesmc_unet_backward": [C.c_void_p, _fp, _fp],
    "qesmc_unet_add_sample": [C.c_void_p, C.c_void_p],
    "qesmc_unet_num_samples": [C.c_void_p, C.POINTER(C.c_int)],
    "qesmc_unet_clear_samples": [C.c_void_p],
    "qesmc_unet_train": [C.c_void_p, C.c_int, C.c_float, C.c_int, C.c_int, C.c_ulonglong,
                         C.POINTER(C.c_float), C.POINTER(C.c_double)],
    "qesmc_unet_eval": [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_float)],
    "qesmc_unet_predict": [C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)],
    "qesmc_unet_set_loss": [C.c_void_p, C.c_int, C.c_float],
    "qesmc_unet_target_scale": [C.c_void_p, C.POINTER(C.c_float)],
}
EXPORTS = EXPORTS + tuple(_UNET_SIGS)


def _unet_lib():
    L = lib()
    if not getattr(L, "_unet_sigs_set", False):
        for name, args in _UNET_SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L._unet_sigs_set = True
    return L


class UNet:
    """3-D U-Net warm start (qesmc_unet_*): BF16 MFMA convolutions on the GPU."""

    def __init__(self, ncx, ncy, ncz, levels=4, base_channels=16, seed=0, device=0, target_scale=0.0):
        L = _unet_lib()
        self._h = C.c_void_p()
        cfg = UNetConfig(levels, base_channels, ncx, ncy, ncz, seed, target_scale)
        _check(L.qesmc_unet_create(C.byref(self._h), int(device), C.byref(cfg)), "qesmc_unet_create")
        n = C.c_longlong()
        _check(L.qesmc_unet_num_params(self._h, C.byref(n)), "qesmc_unet_num_params")
        self.nparams = n.value
        dx, dy, dz = C.c_int(), C.c_int(), C.c_int()
        _check(L.qesmc_unet_dims(self._h, C.byref(dx), C.byref(dy), C.byref(dz)), "qesmc_unet_dims")
        self.dims = (dx.value, dy.value, dz.value)
        self.levels, self.base_channels = levels, base_channels
        self.ncx, self.ncy, self.ncz = int(ncx), int(ncy), int(ncz)

    def close(self):
        if self._h:
            _unet_lib().qesmc_unet_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def target_scale(self):
        s = C.c_float()
        _check(_unet_lib().qesmc_unet_target_scale(self._h, C.byref(s)), "qesmc_unet_target_scale")
        return s.value

    def get_params(self):
        out = np.empty(self.nparams, np.float32)
        _check(_unet_lib().qesmc_unet_get_params(self._h, _f(out), self.nparams), "qesmc_unet_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(_unet_lib().qesmc_unet_set_params(self._h, _f(p), p.size), "qesmc_unet_set_params")

    def forward(self, x):
        """x: [dz, dy, dx, 8] fp32 -> [dz, dy, dx] fp32."""
        dx, dy, dz = self.dims
        x = np.ascontiguousarray(x, np.float32).reshape(dz, dy, dx, UNET_IN_CHANNELS)
        y = np.empty((dz, dy, dx), np.float32)
        _check(_unet_lib().qesmc_unet_forward(self._h, _f(x), _f(y)), "qesmc_unet_forward")
        return y

    def backward(self, grad_out):
        g = np.ascontiguousarray(grad_out, np.float32)
        gp = np.empty(self.nparams, np.float32)
        _check(_unet_lib().qesmc_unet_backward(self._h, _f(g), _f(gp)), "qesmc_unet_backward")
        return gp

    def add_sample(self, ctx: "Context"):
        _check(_unet_lib().qesmc_unet_add_sample(self._h, ctx._h), "qesmc_unet_add_sample")

    def num_samples(self):
        n = C.c_int()
        _check(_unet_lib().qesmc_unet_num_samples(self._h, C.byref(n)), "qesmc_unet_num_samples")
        return n.value

    def clear_samples(self):
        _check(_unet_lib().qesmc_unet_clear_samples(self._h), "qesmc_unet_clear_samples")

    def train(self, steps, lr, first, last, seed=0):
        loss, ms = C.c_float(), C.c_double()
        _check(_unet_lib().qesmc_unet_train(self._h, int(steps), float(lr), int(first), int(last), int(seed),
                                            C.byref(loss), C.byref(ms)), "qesmc_unet_train")
        return loss.value, ms.value

    def set_loss(self, block, weight):
        """Add weight x mean((block-mean error)^2) over block^3 blocks to the loss."""
        _check(_unet_lib().qesmc_unet_set_loss(self._h, int(block), float(weight)), "qesmc_unet_set_loss")

    def eval(self, first, last):
        loss = C.c_float()
        _check(_unet_lib().qesmc_unet_eval(self._h, int(first), int(last), C.byref(loss)), "qesmc_unet_eval")
        return loss.value

    def predict(self, ctx: "Context"):
        """Install lambda0 for ctx's resident problem; returns (ms, TFLOP/s)."""
        ms, tf = C.c_double(), C.c_double()
        _check(_unet_lib().qesmc_unet_predict(self._h, ctx._h, C.byref(ms), C.byref(tf)), "qesmc_unet_predict")
        return ms.value, tf.value


def gpu_available() -> bool:
    """True when a HIP device can be opened through the library."""
    try:
        h = C.c_void_p()
        rc = lib().qesmc_create(C.byref(h), 0)
        if rc == OK:
            lib().qesmc_destroy(h)
            return True
        return False
    except Exception:
        return False

"""Writes tests/golden/city1000_fp64.npz: sampled FP64 converged velocity
fields of the 1000 x 1000 x 150 city, the checker of the large-domain
accuracy bar (the GPU field within 1e-4 m/s of the FP64 solution).

The city: 1000 x 1000 x 150 cells, dx = dy = 1 m, dz geometric from 1 m at
ratio 1.02, 3000 seeded buildings (footprints U(6,16) m, heights lognormal,
median 20 m, sigma 0.6, capped at 120 m), log law u* = 0.5 m/s, z0 = 0.5 m.
Its inputs are built by the oracle's restatement of the reference's builders
(oracle.build_city), from the seeded building list that is stored in the
fixture.

The FP64 solution of each basis direction (wind from 270 and from 180 deg) is
the fixed point of the oracle's FP64 red-black SOR (oracle_sor_rb_f64, the
loop of src/winds/Solver_CPU_RB.cpp:95-184 in double).  The fixed point of
red-black SOR does not depend on the relaxation factor or on the start
(lambda S = sum c_nb lambda_nb - R holds for any omega != 0), so it is reached
here with omega near the optimum for this grid (1.9934: about 350
iterations per decade instead of ~1e4 at 1.78; measured from the slowest
mode's rate at omega 1.99, 920 iterations per decade, 1 - rho_Jacobi ~ 5.5e-6)
from lambda = 0 (or from a checkpoint of an earlier run, which the fixed point
does not depend on), until max|dlambda| falls below `--tol` or stops
falling.  Then it is confirmed as the fixed point of the
reference's own setting: `--check` further iterations at omega = 1.78f must
each move lambda by at most 1e-9 (the reference's stop rule, tol 1e-9).

Every other wind direction's FP64 field is cos(t) F_270 + sin(t) F_180 with
t = (270 - dir) deg: the discrete problem is linear in (u0, v0, w0) and the
single-sensor inflow is cos(t) and sin(t) times one profile (up to the fp32
rounding of the profile values, ~1e-7 relative; tests/test_oracle.py checks
the superposition on a small city).

Samples (float32 of the FP64 values; rounding <= 6e-8 relative):
  xz  u, v, w on the faces of rows j in XZ_ROWS, layers k < KS (x-z slices)
  yz  u, v, w on the faces of columns i in YZ_COLS, layers k < KS
  pts u, v, w at NPTS seeded random faces of the whole domain

omega 1.99 from lambda = 0 until max|dlambda| < 3e-5, then 1.9934.

Run (CPU, about 45 min per direction with 6 threads of this container):
  OMP_NUM_THREADS=6 python tests/golden/make_city1000_fp64.py
"""
import argparse
import ctypes as C
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "qes-public_amd"))

NCX = NCY = 1000
NCZ = 150
SEED = 555
NBUILD = 3000
XZ_ROWS = (300, 700)
YZ_COLS = (300, 700)
KS = 100
NPTS = 131072
OMEGA_REF = float(np.float32(1.78))


def city_spec():
    """Grid, buildings and inflow of the city (the same generator calls as
    bench.large_city)."""
    import qesmc as Q

    dz = Q.dz_geometric(NCZ, 1.0, 1.02)
    blds = Q.random_buildings(SEED, NCX, NCY, 1.0, NBUILD, 6.0, 16.0, 20.0, 0.6, lognormal=True, hcap=120.0,
                              margin=2)
    u_ref = 0.5 / 0.4 * math.log((10.0 + 0.5) / 0.5)
    return dz, np.array(blds, np.float32), u_ref


def solve_fp64(O, p, omega, tol, check, scratch, tag, chunk=100, cap=40000, omega_early=None, switch_at=0.0):
    L = O.lib()
    R = O.divergence_f64(p)
    ck = os.path.join(scratch, f"lambda_{tag}.npy")
    ck_its = ck + ".its"
    lam = np.load(ck) if os.path.exists(ck) else np.zeros(p.ncell, np.float64)
    old = np.empty_like(lam)
    me = C.c_double(0.0)
    its = int(open(ck_its).read()) if os.path.exists(ck) and os.path.exists(ck_its) else 0
    hist, t0 = [], time.time()
    cap += its
    stall = 0
    early = omega_early is not None and its == 0
    while its < cap:
        w = omega_early if early else omega
        n = L.oracle_sor_rb_f64(p.nx, p.ny, p.nz, *p.coeffs, R, lam, old, w, -1.0, chunk, C.byref(me))
        its += n
        hist.append(me.value)
        print(f"[{tag}] omega {w} it {its} max|dlambda| {me.value:.3e} ({time.time() - t0:.0f} s)", flush=True)
        if early and me.value < switch_at:
            early = False  # the fast modes are gone: the slowest one wants omega
        if its % 1000 == 0:
            np.save(ck, lam)
            with open(ck_its, "w") as fh:
                fh.write(str(its))
        if me.value <= tol:
            break
        if len(hist) > 5 and me.value < 1e-9 and me.value >= 0.5 * hist[-6]:
            stall += 1
            if stall >= 3:
                break
        else:
            stall = 0
    np.save(ck, lam)
    with open(ck_its, "w") as fh:
        fh.write(str(its))
    # the reference's own setting: every further iteration moves lambda by <= 1e-9
    checks = []
    for _ in range(check):
        L.oracle_sor_rb_f64(p.nx, p.ny, p.nz, *p.coeffs, R, lam, old, OMEGA_REF, -1.0, 1, C.byref(me))
        checks.append(me.value)
    print(f"[{tag}] omega 1.78 check: max|dlambda| per iteration max {max(checks):.3e}", flush=True)
    assert max(checks) <= 1e-9, checks
    return lam, its, hist[-1], max(checks)


def samples(p, u, v, w, rng_idx):
    nx, ny, nz = p.nx, p.ny, p.nz
    U, V, W = (a.reshape(nz, ny, nx) for a in (u, v, w))
    xz = np.stack([np.stack([A[:KS, j, :] for A in (U, V, W)]) for j in XZ_ROWS]).astype(np.float32)
    yz = np.stack([np.stack([A[:KS, :, i] for A in (U, V, W)]) for i in YZ_COLS]).astype(np.float32)
    pts = np.stack([a[rng_idx] for a in (u, v, w)]).astype(np.float32)
    return xz, yz, pts


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--omega", type=float, default=1.9934)
    ap.add_argument("--omega-early", type=float, default=1.99,
                    help="omega from lambda = 0 until max|dlambda| < --switch-at (the fast modes converge at "
                         "omega - 1 per iteration above their own optimum, so a smaller omega suits them)")
    ap.add_argument("--switch-at", type=float, default=3e-5)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--check", type=int, default=10)
    ap.add_argument("--scratch", default="/tmp/qesmc_city1000_fp64")
    ap.add_argument("--out", default=os.path.join(HERE, "city1000_fp64.npz"))
    a = ap.parse_args()
    os.makedirs(a.scratch, exist_ok=True)
    import oracle as O

    dz, blds, u_ref = city_spec()
    nface = (NCX + 1) * (NCY + 1) * (NCZ + 2)
    rng_idx = np.sort(np.random.default_rng(20261203).choice(nface, NPTS, replace=False)).astype(np.int64)
    out = dict(dz_array=dz, buildings=blds, u_ref=np.float32(u_ref), z0=np.float32(0.5), z_ref=np.float32(10.0),
               xz_rows=np.array(XZ_ROWS, np.int32), yz_cols=np.array(YZ_COLS, np.int32), ks=np.int32(KS),
               pts=rng_idx.astype(np.int32), omega_solve=np.float64(a.omega),
               omega_early=np.float64(a.omega_early), omega_switch_at=np.float64(a.switch_at))
    for wd in (270.0, 180.0):
        tag = f"{int(wd)}"
        t0 = time.time()
        p = O.build_city(NCX, NCY, 1.0, dz, [tuple(float(x) for x in b) for b in blds], (1, u_ref, 10.0, 0.5, wd))
        lam, its, me, chk = solve_fp64(O, p, a.omega, a.tol, a.check, a.scratch, tag, omega_early=a.omega_early,
                                       switch_at=a.switch_at)
        u, v, w = O.velocity_f64(p, lam)
        xz, yz, pts = samples(p, u, v, w, rng_idx)
        out[f"xz_{tag}"], out[f"yz_{tag}"], out[f"pts_{tag}"] = xz, yz, pts
        out[f"iterations_{tag}"] = np.int32(its)
        out[f"final_dlambda_{tag}"] = np.float64(me)
        out[f"check_dlambda_{tag}"] = np.float64(chk)
        out[f"lambda_absmax_{tag}"] = np.float64(np.abs(lam).max())
        print(f"[{tag}] done in {time.time() - t0:.0f} s: {its} iterations, |lambda|max {np.abs(lam).max():.3f}, "
              f"|u|max {np.abs(u).max():.3f}", flush=True)
        del p, lam, u, v, w
    np.savez_compressed(a.out, **out)
    print("wrote", a.out, os.path.getsize(a.out), "bytes")


if __name__ == "__main__":
    main()

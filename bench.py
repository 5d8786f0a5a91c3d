"""Benchmark of the MI355X mass-consistency solve (one GPU).

A step is one complete Solver::solve of the reference schedule on the
1000 x 1000 x 150 synthetic city (151 M cells, 3000 lognormal buildings,
wind from 270 deg): stencil encoding, lambda reset, divergence R, 500
red-black SOR iterations, velocity correction and solid zeroing, with every
input already resident in HBM.  `value` is GLUPS (cell updates per second /
1e9) over the timed steps.

  python bench.py [--steps K] [--warmup W] [--gpus 1]

Further legs reported in the same JSON line:
  cpu_baseline  the oracle's fp32 red-black restatement on the host cores,
                a bounded sample of the same 1000 x 1000 x 150 solve
                (OMP_PROC_BIND=close, OMP_PLACES=cores).
  city1000_tol  time to tolerance of the 8-direction ensemble on the same
                city against the FP64 fixture tests/golden/city1000_fp64.npz:
                iterations and device ms until the sampled velocity is within
                1e-4 m/s of the FP64 solution, cold and from the
                superposition of the two basis directions.
  city400       the reference schedule on the 400 x 400 x 120 city.
  time_to_tol   128 x 128 x 64 held-out cities (the warm-start workload):
                iterations and device ms until the velocity field is within
                1e-4 m/s (L_inf) of the FP64 converged oracle, cold start
                versus the trained U-Net's lambda0; U-Net MFMA throughput.
  batch         64 cities of uneven size (64-256 x 64-256 x 64-96 cells),
                wind directions 5.625 deg apart, 500 SOR iterations each:
                one batched launch per colour pass (HIP graph replay) versus
                the instances solved one after another.

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "qes-public_amd"))
# CPU baseline: one OpenMP thread per physical core, packed; the OpenMP
# runtime reads these when the oracle library is first loaded (and then binds
# this thread to one core, so the CPUs the process may use are noted first)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")
try:
    START_CPUS = sorted(os.sched_getaffinity(0))
except AttributeError:
    START_CPUS = list(range(os.cpu_count() or 1))

import numpy as np  # noqa: E402

import qesmc as Q  # noqa: E402

METRIC = "time-to-tol per domain (ms) + GLUPS; U-Net warm-start vs plain SOR, same GPU"
HBM_PEAK_GBS = 8000.0
# over-relaxation factor of the reference solver (src/winds/Solver.h:83)
REFERENCE_OMEGA = 1.78


def city_400x400x120():
    """400 x 400 x 120 cells, dx = dy = 2 m, dz geometric from 1 m at ratio
    1.02, 200 seeded buildings (footprints U(10,40) m, heights U(10,60) m),
    log law with u* = 0.5 m/s, z0 = 0.5 m (U_ref at 10 m chosen to give that
    u*), wind from 217 deg."""
    ncx, ncy, ncz, dx = 400, 400, 120, 2.0
    dz = Q.dz_geometric(ncz, 1.0, 1.02)
    blds = Q.random_buildings(2024, ncx, ncy, dx, 200, 10.0, 40.0, 10.0, 60.0, margin=2)
    u_ref = 0.5 / 0.4 * math.log((10.0 + 0.5) / 0.5)
    city = Q.city_build(ncx, ncy, dx, dz, blds, (1, u_ref, 10.0, 0.5, 217.0))
    desc = ("400x400x120 cells, dx=dy=2 m, dz 1 m x1.02 geometric, %d seeded buildings, "
            "log law u*=0.5 m/s z0=0.5 m from 217 deg; R + 500 red-black SOR iterations + correction"
            % len(blds))
    return city, desc


def bytes_per_lup(res):
    """Algorithmic HBM bytes per lattice update of the dominant SOR kernel
    (a colour pass updates half the cells; a fused launch all of them):
      split, wall mask : own lambda r+w 8 + R 4 + mask 1 + other-colour lambda 4 = 17
      split, full coef : 8 + 4 + 24 + 4 = 40
      fused, wall mask : own lambda r+w 8 + R 4 + mask 1 = 13 (each lambda is
                         read once per iteration: a wave's march keeps the
                         rows around the ones it relaxes in registers)
      plain            : coefficient and R lines of both colours are fetched by
                         each pass: 2*(24+4) + lambda r (4, neighbours shared) + w 4 = 64"""
    if res["path"] == Q.PATH_PLAIN:
        return 64.0
    if res["path"] == Q.PATH_FUSED:
        return 13.0
    return 17.0 if res["compressed"] else 40.0


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    return O


def fp64_reference_fields(cities):
    """FP64 converged velocity fields (oracle red-black SOR to max|dlambda| <=
    1e-9 in double) of the given cities, which the time_to_tol leg compares against."""
    O = _oracle()
    out = {}
    for seed, city in cities.items():
        u, v, w, _, its, _ = O.solve_rb_f64(O.Problem(**city.arrays()), tol=1e-9)
        out[seed] = (u, v, w, its)
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def physical_cores(cpus=None):
    """Physical cores among the CPUs this process may run on (sibling
    hyperthreads counted once), from sysfs topology.  `cpus` defaults to the
    process's affinity; bench.py records it at start-up, before the OpenMP
    runtime binds the main thread to one place."""
    seen = set()
    if cpus is None:
        try:
            cpus = sorted(os.sched_getaffinity(0))
        except AttributeError:
            cpus = list(range(os.cpu_count() or 1))
    for c in cpus:
        base = "/sys/devices/system/cpu/cpu%d/topology/" % c
        try:
            with open(base + "core_id") as fh:
                core = fh.read().strip()
            with open(base + "physical_package_id") as fh:
                pkg = fh.read().strip()
        except OSError:
            core, pkg = str(c), "0"
        seen.add((pkg, core))
    return len(seen), len(cpus)


def cpu_baseline_resident(ctx, desc, samples=5, per_sample=4, serial_iters=1):
    """The oracle's fp32 red-black restatement of Solver_CPU_RB on the host
    cores (OpenMP; OMP_PROC_BIND=close, OMP_PLACES=cores) on the resident
    city of ctx, copied back from the device: R once, then `samples` timed
    samples of `per_sample` SOR iterations; `value` is the median sample's
    rate (per-iteration cost is the same at every iteration, so the rate
    extrapolates to the 500-iteration schedule).  Also the serial
    lexicographic Solver_CPU restatement on one thread."""
    city = ctx.get_inputs()
    O = _oracle()
    p = O.Problem(**city.arrays())
    del city
    t0 = time.perf_counter()
    R = O.divergence_f32(p)
    r_s = time.perf_counter() - t0
    lam, rates = None, []
    for _ in range(samples):
        t0 = time.perf_counter()
        lam, it, _ = O.sor_rb_f32(p, R, lam, itermax=per_sample, tol=-1.0)
        rates.append(p.interior_lups() * it / (time.perf_counter() - t0) / 1e9)
    threads = O.num_threads()
    O.set_num_threads(1)
    try:
        t0 = time.perf_counter()
        _, it_s, _ = O.sor_serial_f32(p, R, None, itermax=serial_iters, tol=-1.0)
        serial = p.interior_lups() * it_s / (time.perf_counter() - t0) / 1e9
    finally:
        O.set_num_threads(threads)
    phys, logical = physical_cores(START_CPUS)
    glups = float(np.median(rates))
    its = 500
    t_ref_ms = (r_s + p.interior_lups() * its / (glups * 1e9)) * 1e3
    return {"value": round(glups, 4), "unit": "GLUPS", "cores": threads, "kind": "port",
            "physical_cores_available": phys, "logical_cpus_available": logical,
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND"), "omp_places": os.environ.get("OMP_PLACES"),
            "samples_glups": [round(r, 4) for r in rates], "cpu_model": _cpu_model(),
            "serial_single_thread_glups": round(serial, 4),
            "t_ref_ms_extrapolated": round(t_ref_ms, 1),
            "sample": "%s: R once (%.2f s), then the median of %d samples of %d SOR iterations "
                      "(oracle/qesmc_oracle.c Solver_CPU_RB restatement, OpenMP, %d threads); t_ref_ms_extrapolated "
                      "= R + 500 iterations at that rate (extrapolated, not run); serial Solver_CPU restatement "
                      "on 1 thread over %d iteration(s)" % (desc, r_s, samples, per_sample, threads, serial_iters)}


WEIGHTS = os.path.join(ROOT, "qes-public_amd", "qesmc", "weights", "unet_128x128x64.npz")
BF16_PEAK_TFLOPS = 2500.0


def _linf(a, b):
    return float(np.abs(a.astype(np.float64) - b).max())


# refresh interval of the FP64 residual (qesmc_params.refine_every) in the
# 128 x 128 x 64 stop-rule runs at the estimated omega
REFINE_EVERY = 100
# (name, qesmc_params.stop_extend) at tol 1e-4
STOP_RULES = (("reference_tol_1e-4", 0.0), ("production_tol_1e-4_extend_0.3", 0.3))


def time_to_tol_leg(seeds, fp64_ref, tol_ms=1e-4, chunk=25, cap=6000, omega_steps=100):
    """Cold versus U-Net warm start on held-out 128 x 128 x 64 cities, at the
    reference's omega 1.78 and at the omega qesmc_estimate_omega gives for
    each city (the estimate depends only on the geometry, so one estimate
    serves every wind direction and time step of a city; its cost is listed
    on its own), and cold with the cyclic Chebyshev schedule from the
    estimated rho_J (qesmc_params.chebyshev_rho).

    Pass 1 (untimed) runs SOR in chunks of `chunk` iterations and reads the
    velocity back after each, to find the first checkpoint whose L_inf error
    against the FP64 field is <= tol_ms (and the error after the reference's
    500 iterations).  Pass 2 times that many iterations from a fresh start:
    R + SOR + correction on the device (qesmc_result.ms_total), plus the
    U-Net inference for the warm start.  Without FP64 fields the stop rule
    max|dlambda| <= 1e-4 is used instead.  With them, the device's stop
    rules are also run at tol 1e-4 (STOP_RULES: the reference's, and the
    production rule that runs 30 % more iterations after it), recording
    where each stops, its device ms and its distance to the FP64 field."""
    from qesmc import warmstart as W

    if not os.path.exists(WEIGHTS):
        return {"skipped": "trained weights not found: " + WEIGHTS}
    net = W.load(WEIGHTS)
    rows = []
    runs = (("cold", "ref"), ("warm", "ref"), ("cold", "est"), ("warm", "est"), ("cold", "cheb"))
    suffix = {"ref": "", "est": "_omega_est", "cheb": "_chebyshev"}
    with Q.Context(0) as ctx:
        for seed in seeds:
            city = W.seeded_city(seed)
            W.load_problem(ctx, city)
            t0 = time.perf_counter()
            rho, om_est = ctx.estimate_omega(omega_steps)
            est_ms = (time.perf_counter() - t0) * 1e3
            row = {"seed": seed, "omega_estimate": {"rho_jacobi": rho, "omega": om_est, "ms": round(est_ms, 3)}}
            for mode, which in runs:
                omega = REFERENCE_OMEGA if which == "ref" else om_est
                key = mode + suffix[which]

                def sched(start):
                    """the cyclic Chebyshev schedule from the estimated rho_J,
                    continued across the chunked calls"""
                    return dict(chebyshev_rho=rho, chebyshev_start=start) if which == "cheb" else {}

                ctx.set_lambda(None)
                if mode == "warm":
                    net.predict(ctx)
                it, it_tol, err500, err_tol = 0, None, None, None
                if fp64_ref is not None:
                    u64, v64, w64, _ = fp64_ref[seed]
                    while it < cap:
                        ctx.run(itermax=chunk, tol=-1.0, omega=omega, **sched(it))
                        it += chunk
                        u, v, w = ctx.get_velocity()
                        err = max(_linf(u, u64), _linf(v, v64), _linf(w, w64))
                        if it == 500:
                            err500 = err
                        if it_tol is None and err <= tol_ms:
                            it_tol, err_tol = it, err
                        if it_tol is not None and it >= 500:
                            break
                # pass 2: timed from a fresh start
                ctx.set_lambda(None)
                unet_ms, tflops = net.predict(ctx) if mode == "warm" else (0.0, 0.0)
                if fp64_ref is not None and it_tol is not None:
                    r = ctx.run(itermax=it_tol, tol=-1.0, omega=omega, **sched(0))
                else:
                    r = ctx.run(itermax=20000, tol=1e-4, omega=omega, **sched(0))
                row[key] = {"omega": omega if which != "cheb" else "cyclic Chebyshev, rho_J %.8f" % rho,
                            "iterations": r["iterations"], "ms": round(r["ms_total"] + unet_ms, 3),
                            "sor_ms": round(r["ms_sor"], 3), "err_at_tol": err_tol, "err_at_500": err500}
                if mode == "warm":
                    row[key].update(unet_ms=round(unet_ms, 3), unet_tflops=round(tflops, 1))
                if fp64_ref is not None:
                    # the device's stop rules at tol 1e-4: the reference's
                    # (max|dlambda| <= tol) and the production one (then 30 %
                    # more iterations): where they stop, what it costs, and how
                    # far from the FP64 field that is
                    # at the estimated omega, also with the residual re-formed
                    # in FP64 every REFINE_EVERY iterations (the plain fp32
                    # iteration's max|dlambda| settles on its rounding noise
                    # near 1e-4 there)
                    row[key]["stop_rule"] = {}
                    variants = [(n_, e_, 0) for n_, e_ in STOP_RULES]
                    if which != "ref":
                        variants += [(n_ + "_refine_every_%d" % REFINE_EVERY, e_, REFINE_EVERY) for n_, e_ in STOP_RULES]
                    for name, ext, ref_every in variants:
                        ctx.set_lambda(None)
                        pred_ms = net.predict(ctx)[0] if mode == "warm" else 0.0
                        r = ctx.run(itermax=cap, tol=1e-4, stop_extend=ext, omega=omega, refine_every=ref_every,
                                    **sched(0))
                        u, v, w = ctx.get_velocity()
                        row[key]["stop_rule"][name] = {
                            "iterations": r["iterations"], "ms": round(r["ms_total"] + pred_ms, 3),
                            "err": max(_linf(u, u64), _linf(v, v64), _linf(w, w64))}
            rows.append(row)
    net.close()

    def mean(key, field):
        vals = [r[key][field] for r in rows if r[key].get(field) is not None]
        return float(np.mean(vals)) if vals else None

    keys = [m + suffix[w] for m, w in runs]
    stop = None
    if fp64_ref is not None:
        stop = {}
        for key in keys:
            stop[key] = {}
            for name in rows[0][key]["stop_rule"]:
                v = [r[key]["stop_rule"][name] for r in rows]
                stop[key][name] = {"mean_iterations": float(np.mean([x["iterations"] for x in v])),
                                   "mean_ms": round(float(np.mean([x["ms"] for x in v])), 3),
                                   "max_err": max(x["err"] for x in v),
                                   "cities_within_tol_ms": sum(x["err"] <= tol_ms for x in v)}
        prod = STOP_RULES[1][0]
        stop["speedup_warm_vs_cold_production_rule"] = round(stop["cold"][prod]["mean_ms"]
                                                             / stop["warm"][prod]["mean_ms"], 3)
        stop["speedup_omega_est_vs_reference_production_rule"] = round(
            stop["cold"][prod]["mean_ms"] / stop["cold_omega_est"][prod]["mean_ms"], 3)
        prod_ref = prod + "_refine_every_%d" % REFINE_EVERY
        stop["speedup_omega_est_refined_vs_reference_production_rule"] = round(
            stop["cold"][prod]["mean_ms"] / min(stop[k][prod_ref]["mean_ms"]
                                                for k in ("cold_omega_est", "warm_omega_est")), 3)
        stop["speedup_chebyshev_refined_vs_reference_production_rule"] = round(
            stop["cold"][prod]["mean_ms"] / stop["cold_chebyshev"][prod_ref]["mean_ms"], 3)
    summary = {}
    for key in keys:
        summary[key] = {"iterations": mean(key, "iterations"), "ms": mean(key, "ms"),
                        "err_at_500": mean(key, "err_at_500")}
        if key.startswith("warm"):
            summary[key]["unet_ms"] = mean(key, "unet_ms")

    def ratio(a_, b_):
        x, y = summary[a_]["ms"], summary[b_]["ms"]
        return round(x / y, 3) if x and y else None

    tfl = mean("warm", "unet_tflops")
    return {
        "workload": "%d held-out seeded cities, 128x128x64 cells, dx=2 m, dz 1 m x1.02, 12-30 buildings, log law, "
                    "random direction; U-Net 4 levels 16/32/64/128 trained on 2000 solver-generated cities" % len(rows),
        "criterion": ("L_inf(u,v,w) vs FP64 oracle <= %g m/s, checked every %d iterations" % (tol_ms, chunk)
                      if fp64_ref is not None else "stop rule max|dlambda| <= 1e-4 (no FP64 fields)"),
        **summary,
        "speedup_warm_vs_cold": ratio("cold", "warm"),
        "speedup_warm_vs_cold_omega_est": ratio("cold_omega_est", "warm_omega_est"),
        "speedup_omega_est_vs_reference": ratio("cold", "cold_omega_est"),
        "speedup_chebyshev_vs_omega_est": ratio("cold_omega_est", "cold_chebyshev"),
        "speedup_best_vs_reference_cold": round(summary["cold"]["ms"] / min(summary[k]["ms"] for k in keys), 3),
        "omega_estimate": {"steps_cap": omega_steps,
                           "mean_omega": float(np.mean([r["omega_estimate"]["omega"] for r in rows])),
                           "mean_ms": round(float(np.mean([r["omega_estimate"]["ms"] for r in rows])), 3),
                           "note": "once per city geometry, shared by every wind direction and time step"},
        "stop_rule_vs_fp64": stop,
        "unet_mfma": {"tflops": tfl, "peak_bf16_dense": BF16_PEAK_TFLOPS,
                      "frac": round(tfl / BF16_PEAK_TFLOPS, 4) if tfl else None},
        "per_city": rows,
    }


def batch_cities(n=64, seed=3):
    """Uneven ensemble: (ncx, ncy) uniform in 64..256, ncz in 64..96, 20-120
    requested buildings (footprints U(10,40) m, heights U(10,60) m; fewer are
    placed where they do not fit), dx = 2 m, dz 1 m x1.02, log law u* = 0.5 m/s
    z0 = 0.5 m, wind directions q * 5.625 deg."""
    rng = np.random.default_rng(seed)
    u_ref = 0.5 / 0.4 * math.log((10.0 + 0.5) / 0.5)
    out = []
    for q in range(n):
        ncx, ncy = (int(v) for v in rng.integers(64, 257, 2))
        ncz = int(rng.integers(64, 97))
        nb = int(rng.integers(20, 121))
        blds = Q.random_buildings(5000 + q, ncx, ncy, 2.0, nb, 10.0, 40.0, 10.0, 60.0, margin=2)
        out.append((ncx, ncy, Q.dz_geometric(ncz, 1.0, 1.02), blds, (1, u_ref, 10.0, 0.5, q * 5.625)))
    return out


def batch_leg(n=64, itermax=500, reps=2):
    specs = batch_cities(n)
    ctxs = [Q.Context(0) for _ in specs]
    try:
        t0 = time.perf_counter()
        for c, (ncx, ncy, dz, blds, inflow) in zip(ctxs, specs):
            c.set_city(ncx, ncy, 2.0, dz, blds, inflow)
        build_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        for c in ctxs:
            Q.autotune(c, iterations=20)
        tune_s = time.perf_counter() - t0
        lups = sum(c.city.interior_lups() for c in ctxs) * itermax
        best = {}
        for g in (True, False):
            ms = []
            for _ in range(reps):
                for c in ctxs:
                    c.set_lambda(None)
                res, ms_loop = Q.run_batch(ctxs, itermax=itermax, tol=1e-9, check_every=50, use_graph=g)
                ms.append(ms_loop)
                batch_tiling = (res[0]["tile_k"], res[0]["tile_v"], res[0]["wg_per_cu"])
            best["graph" if g else "eager"] = min(ms)
        seq = []
        for _ in range(reps):
            tot = 0.0
            for c in ctxs:
                c.set_lambda(None)
                tot += c.run(itermax=itermax, tol=1e-9)["ms_sor"]
            seq.append(tot)
        seq_ms = min(seq)
    finally:
        for c in ctxs:
            c.close()
    return {
        "workload": "%d cities, 64-256 x 64-256 x 64-96 cells (%.1f M cells in total), %d buildings placed, "
                    "directions 5.625 deg apart, %d SOR iterations each"
                    % (n, sum(c[2].size * c[0] * c[1] for c in specs) / 1e6, sum(len(c[3]) for c in specs), itermax),
        "batched_graph_ms": round(best["graph"], 3), "batched_eager_ms": round(best["eager"], 3),
        "sequential_ms": round(seq_ms, 3),
        "glups_batched": round(lups / best["graph"] / 1e6, 2), "glups_sequential": round(lups / seq_ms / 1e6, 2),
        "speedup_batched_vs_sequential": round(seq_ms / best["graph"], 3),
        "device_city_build_s": round(build_s, 3), "autotune_s": round(tune_s, 3),
        "tiling": "batched: one tiling for all cities (tile_k %d, tile_v %d, %d work-groups per CU); "
                  "sequential: each city's autotuned tiling" % batch_tiling,
    }


def large_city(dirs=8):
    """1000 x 1000 x 150 cells, dx = dy = 1 m, dz geometric from 1 m at ratio
    1.02 (top at 925 m), 3000 seeded buildings with heights lognormal (median
    20 m, sigma 0.6) capped at 120 m on footprints U(6,16) m (3000 buildings
    of the 400 x 400 city's 10-40 m footprints do not fit on 1 km^2), log law
    u* = 0.5 m/s z0 = 0.5 m, `dirs` wind directions 360/dirs deg apart."""
    ncx = ncy = 1000
    dz = Q.dz_geometric(150, 1.0, 1.02)
    blds = Q.random_buildings(555, ncx, ncy, 1.0, 3000, 6.0, 16.0, 20.0, 0.6, lognormal=True, hcap=120.0, margin=2)
    u_ref = 0.5 / 0.4 * math.log((10.0 + 0.5) / 0.5)
    winds = [q * 360.0 / dirs for q in range(dirs)]
    return ncx, ncy, dz, blds, u_ref, winds


FIXTURE_C5 = os.path.join(ROOT, "tests", "golden", "city1000_fp64.npz")


def c5_fixture_probes(F, nx, ny):
    """Face indices (i + j nx + k nx ny) of the FP64 fixture's samples of the
    1000 x 1000 x 150 city (x-z slices of rows xz_rows, y-z slices of columns
    yz_cols, layers k < ks, and random faces pts) and the FP64 u, v, w of the
    two basis directions there: {270: (3, n), 180: (3, n)} float64."""
    ks = int(F["ks"])
    faces, vals = [], {270: [], 180: []}
    K, I = np.meshgrid(np.arange(ks, dtype=np.int64), np.arange(nx, dtype=np.int64), indexing="ij")
    for r, j in enumerate(F["xz_rows"]):
        faces.append((I + int(j) * nx + K * nx * ny).ravel())
        for tag in vals:
            vals[tag].append(F["xz_%d" % tag][r].reshape(3, -1))
    K, J = np.meshgrid(np.arange(ks, dtype=np.int64), np.arange(ny, dtype=np.int64), indexing="ij")
    for r, i in enumerate(F["yz_cols"]):
        faces.append((int(i) + J * nx + K * nx * ny).ravel())
        for tag in vals:
            vals[tag].append(F["yz_%d" % tag][r].reshape(3, -1))
    faces.append(F["pts"].astype(np.int64))
    for tag in vals:
        vals[tag].append(F["pts_%d" % tag])
    return np.concatenate(faces), {t: np.concatenate(v, axis=1).astype(np.float64) for t, v in vals.items()}


def c5_reference(ref, wind_deg):
    """FP64 sampled field of any direction: the problem is linear in the
    inflow and a single-sensor inflow from `wind_deg` is cos(t) x the 270 deg
    inflow + sin(t) x the 180 deg one, t = 270 - wind_deg."""
    th = math.radians(270.0 - wind_deg)
    return math.cos(th) * ref[270] + math.sin(th) * ref[180]


def probe_error(ctx, ref_dir):
    """Sampled L_inf (u, v, w) distance of ctx's corrected field to ref_dir."""
    got = np.stack(Q.get_probes(ctx)).astype(np.float64)
    d = np.abs(got - ref_dir)
    return float(np.inf) if not np.all(np.isfinite(d)) else float(d.max())


def tol_search(work, save, ref_dir, tol_ms, coarse, fine, cap, omega=1.78):
    """Iterations from work's current lambda until its sampled velocity is
    within tol_ms of ref_dir: chunks of `coarse` iterations (lambda saved in
    `save` before each), then, from the last chunk's start, chunks of `fine`.
    Returns (iterations or None, SOR ms per iteration, R + correction ms,
    sampled error at the stop, errors at the coarse checkpoints)."""
    its, run_its, ms_sor, extra, hist = 0, 0, 0.0, 0.0, []
    r = work.run(itermax=0, tol=-1.0)
    extra = r["ms_total"]
    e = probe_error(work, ref_dir)
    hist.append((0, e, None))
    if e <= tol_ms:
        return 0, 0.0, extra, e, hist
    while its < cap:
        Q.lambda_combine(save, [work], np.array([1.0], np.float32))
        r = work.run(itermax=coarse, tol=-1.0, omega=omega)
        its += coarse
        run_its += coarse
        ms_sor += r["ms_sor"]
        extra = r["ms_total"] - r["ms_sor"]
        e = probe_error(work, ref_dir)
        hist.append((its, e, r["final_max_dlambda"]))
        print(f"  tol search: {its} its, sampled error {e:.3e} m/s, max|dlambda| {r['final_max_dlambda']:.3e}",
              file=sys.stderr, flush=True)
        if e <= tol_ms:
            Q.lambda_combine(work, [save], np.array([1.0], np.float32))
            its -= coarse
            while True:
                r = work.run(itermax=fine, tol=-1.0, omega=omega)
                its += fine
                run_its += fine
                ms_sor += r["ms_sor"]
                e = probe_error(work, ref_dir)
                if e <= tol_ms:
                    return its, ms_sor / run_its, extra, e, hist
    return None, ms_sor / max(run_its, 1), extra, e, hist


def refined_search(work, ref_dir, tol_ms, basis_tol, omega, chunk, cap):
    """Chunk-iteration refinements (qesmc_refine) from work's current lambda,
    one after another, with the sampled error against ref_dir after each:
    bit for bit the iterates of one run with refine_every = chunk.  Returns
    (iterations to tol_ms, iterations to basis_tol, checkpoints)."""
    its, it_tol, it_basis, hist = 0, None, None, []
    while its < cap:
        r = work.refine(itermax=chunk, tol=-1.0, omega=omega)
        its += chunk
        e = probe_error(work, ref_dir)
        hist.append((its, e, r["final_max_dlambda"]))
        if it_tol is None and e <= tol_ms:
            it_tol = its
        if e <= basis_tol:
            it_basis = its
            break
    return it_tol, it_basis, hist


def c5_tol_leg(tol_ms=1e-4, dirs=8, cap=20000, cpu_ms_per_it=None, omega_steps=100, chunk=125,
               basis_tol=6.5e-5):
    """Time to tolerance of the `dirs`-direction ensemble on the 1000 x 1000 x
    150 city, against the FP64 fixture (sampled: two x-z and two y-z slices of
    the lowest 100 layers and 131072 random faces; the fixture's FP64 fields
    are the fixed point of the oracle's FP64 red-black SOR).  T_tol = device
    ms of R + SOR + correction (+ the U-Net call) until the sampled velocity
    is within tol_ms of the FP64 solution.
      reference   wind from 270 deg from lambda = 0 at the reference's omega
                  1.78 with the plain fp32 iteration: the red-black SOR
                  yardstick (checked every 2000 iterations, then every 100
                  from the last checkpoint above);
      omega       qesmc_estimate_omega (`omega_steps` Lanczos steps);
      basis       270 and 180 deg at that omega with refine_every = `chunk`
                  (the residual re-formed in FP64 every `chunk` iterations,
                  so that the fp32 rounding noise, which SOR amplifies near
                  omega = 2, sets no floor), from lambda = 0 and from the
                  U-Net's lambda0: T_tol, and the iterations to `basis_tol`,
                  to which the faster start solves the basis so that the
                  combinations below stay within tol_ms: at every sampled
                  face a combination's error is at most |cos t| e_270 +
                  |sin t| e_180 <= sqrt(2) basis_tol (9.2e-5 m/s for the
                  default 6.5e-5), and each is measured;
      superposed  every other direction from cos(t) lambda_270 + sin(t)
                  lambda_180, refined in chunks of 25 iterations if needed.
    Iteration counts come from chained refinements with checks in between
    (bit for bit the iterates of one run); the times from one timed run of
    that many iterations from a fresh start.  The ensemble total (omega
    estimate, both basis solves, the superposed directions) is measured; the
    cold ensemble at the reference's omega is 8 x the reference solve
    (extrapolated, stated as such), as are the classic-layout path and the
    CPU at their rates."""
    if not os.path.exists(FIXTURE_C5):
        return {"skipped": "FP64 fixture not found: " + os.path.relpath(FIXTURE_C5, ROOT)}
    F = np.load(FIXTURE_C5)
    ncx, ncy, dz, blds, u_ref, winds = large_city(dirs)
    if not (np.array_equal(F["dz_array"], dz) and np.array_equal(F["buildings"], np.array(blds, np.float32))):
        return {"skipped": "fixture city differs from the generator's"}
    faces, ref = c5_fixture_probes(F, ncx + 1, ncy + 1)
    out = {"workload": "1000x1000x150 cells (151 M), dx=dy=1 m, dz 1 m x1.02, %d lognormal buildings, log law "
                       "u*=0.5 m/s z0=0.5 m, ensemble of %d directions %.0f deg apart" % (len(blds), dirs, 360.0 / dirs),
           "criterion": "sampled L_inf(u,v,w) <= %g m/s of the FP64 solution (%d sampled faces x 3 components: "
                        "2 x-z and 2 y-z slices of layers k < %d, %d random faces)"
                        % (tol_ms, faces.size, int(F["ks"]), F["pts"].size),
           "fp64_fixture": {"iterations": [int(F["iterations_270"]), int(F["iterations_180"])],
                            "final_max_dlambda": [float(F["final_dlambda_270"]), float(F["final_dlambda_180"])],
                            "check_max_dlambda_omega_1.78": [float(F["check_dlambda_270"]),
                                                             float(F["check_dlambda_180"])]}}
    from qesmc import warmstart as W

    net = W.load(WEIGHTS, size=(ncx, ncy, dz.size - 1)) if os.path.exists(WEIGHTS) else None
    ctx = [Q.Context(0) for _ in range(4)]  # work, save, basis 270, basis 180
    work, save, bx, by = ctx
    try:
        for c in ctx:
            c.set_city(ncx, ncy, 1.0, dz, blds, (1, u_ref, 10.0, 0.5, 270.0))
        Q.set_probes(work, faces)
        lups = work.city.interior_lups()
        r = work.run(itermax=100, tol=-1.0, path=Q.PATH_PLAIN)
        plain_ms_it = r["ms_sor"] / 100

        def start(wd, mode):
            work.set_city_inflow((1, u_ref, 10.0, 0.5, wd))
            work.set_lambda(None)
            return net.predict(work) if mode == "warm" else (0.0, 0.0)

        # the yardstick: plain fp32 SOR at the reference's omega from lambda = 0
        start(270.0, "cold")
        its, ms_it, extra, err, hist = tol_search(work, save, c5_reference(ref, 270.0), tol_ms, 2000, 100, 120000,
                                                  omega=REFERENCE_OMEGA)
        reference = {"wind_deg": 270.0, "start": "cold", "omega": REFERENCE_OMEGA, "iterations": its,
                     "sampled_err": err, "ms": round(its * ms_it + extra, 1) if its is not None else None,
                     "sor_ms_per_iteration": round(ms_it, 4),
                     "err_at_checkpoints": [[i, round(e, 7), dl] for i, e, dl in hist]}
        print(f"c5 tol: reference {its} its {reference['ms']} ms err {err:.2e}", file=sys.stderr, flush=True)
        out["reference"] = reference
        t0 = time.perf_counter()
        rho, omega = work.estimate_omega(omega_steps)
        est_ms = (time.perf_counter() - t0) * 1e3
        out["omega_estimate"] = {"rho_jacobi": rho, "omega": omega, "lanczos_steps": omega_steps,
                                 "ms": round(est_ms, 1)}

        def timed(wd, mode, n):
            """One run of n iterations with refine_every = chunk from a fresh start."""
            pred_ms, tflops = start(wd, mode)
            rr = work.run(itermax=n, tol=-1.0, omega=omega, refine_every=chunk)
            return rr["ms_total"] + pred_ms, probe_error(work, c5_reference(ref, wd)), pred_ms, tflops

        basis_rows, ens_ms = [], est_ms
        for wd, basis in ((270.0, bx), (180.0, by)):
            rows = {}
            for mode in (("cold", "warm") if net is not None else ("cold",)):
                start(wd, mode)
                it_tol, it_basis, hist = refined_search(work, c5_reference(ref, wd), tol_ms, basis_tol, omega,
                                                        chunk, cap)
                row = {"wind_deg": wd, "start": mode, "omega": round(omega, 5), "refine_every": chunk,
                       "iterations": it_tol, "iterations_to_basis_tol": it_basis,
                       "err_at_checkpoints": [[i, round(e, 7), dl] for i, e, dl in hist]}
                if it_tol is not None:
                    ms, err, pred_ms, tflops = timed(wd, mode, it_tol)
                    row.update(ms=round(ms, 1), sampled_err=err)
                    if mode == "warm":
                        row.update(unet_ms=round(pred_ms, 1), unet_tflops=round(tflops, 1))
                if it_basis is not None:
                    ms_b, err_b, _, _ = timed(wd, mode, it_basis)
                    row.update(basis_ms=round(ms_b, 1), basis_err=err_b)
                rows[mode] = row
                print(f"c5 tol: {wd:5.1f} deg {mode} omega {omega:.4f}: T_tol {it_tol} its {row.get('ms')} ms, "
                      f"basis {it_basis} its {row.get('basis_ms')} ms", file=sys.stderr, flush=True)
            best = min(rows.values(), key=lambda r_: r_.get("basis_ms", 1e300))
            # the chosen start once more, into the basis
            ms_b, err_b, _, _ = timed(wd, best["start"], best["iterations_to_basis_tol"])
            Q.lambda_combine(basis, [work], np.array([1.0], np.float32))
            ens_ms += ms_b
            basis_rows.append({"wind_deg": wd, "starts": rows, "chosen": best["start"],
                               "iterations": best["iterations_to_basis_tol"], "ms": round(ms_b, 1),
                               "sampled_err": err_b})
        out["basis"] = basis_rows
        sup = []
        for wd in winds:
            if wd in (270.0, 180.0):
                continue
            th = math.radians(270.0 - wd)
            work.set_city_inflow((1, u_ref, 10.0, 0.5, wd))
            tc = time.perf_counter()
            Q.lambda_combine(work, [bx, by], np.array([math.cos(th), math.sin(th)], np.float32))
            comb_ms = (time.perf_counter() - tc) * 1e3
            r = work.run(itermax=0)
            ms = comb_ms + r["ms_total"]
            err = probe_error(work, c5_reference(ref, wd))
            its = 0
            while err > tol_ms and its < 2000:
                r = work.refine(itermax=25, tol=-1.0, omega=omega)
                its += 25
                ms += r["ms_total"]
                err = probe_error(work, c5_reference(ref, wd))
            ok = err <= tol_ms
            sup.append({"wind_deg": wd, "iterations": its if ok else None, "sampled_err": err,
                        "ms": round(ms, 1) if ok else None})
            if ok:
                ens_ms += ms
            print(f"c5 tol: {wd:5.1f} deg superposed {its} its {ms:.1f} ms err {err:.2e}", file=sys.stderr, flush=True)
        out["superposed"] = sup
        if reference["ms"] is not None and all(x["ms"] is not None for x in sup):
            cold_ms, cold_its = reference["ms"], reference["iterations"]
            ens = {"ms": round(ens_ms, 1), "measured": True,
                   "method": "omega from qesmc_estimate_omega (%.4f); basis directions with refine_every %d from "
                             "the faster start to %g m/s; the other %d directions by superposition"
                             % (omega, chunk, basis_tol, len(sup)),
                   "reference_cold_ms_extrapolated": round(dirs * cold_ms, 1),
                   "speedup_vs_reference_cold": round(dirs * cold_ms / ens_ms, 2),
                   "plain_path_cold_ms_extrapolated": round(dirs * cold_its * plain_ms_it, 1),
                   "speedup_vs_plain_path_cold": round(dirs * cold_its * plain_ms_it / ens_ms, 2),
                   "note": "the cold totals are %d x the measured cold solve at omega 1.78 (270 deg) on the "
                           "library's default path (the one every solve here runs on), and at the classic-layout "
                           "path's measured time per iteration" % dirs}
            if cpu_ms_per_it:
                ens["cpu_cold_ms_extrapolated"] = round(dirs * cold_its * cpu_ms_per_it, 0)
                ens["speedup_vs_cpu"] = round(dirs * cold_its * cpu_ms_per_it / ens_ms, 1)
            out["ensemble"] = ens
        # one direction alone: the faster start's T_tol at the estimated omega
        one = [min((r_ for r_ in b["starts"].values() if r_.get("ms") is not None), key=lambda r_: r_["ms"],
                   default=None) for b in basis_rows]
        if reference["ms"] is not None and all(one):
            out["single_direction"] = {
                "ms": [o["ms"] for o in one], "start": [o["start"] for o in one],
                "speedup_vs_reference_cold": [round(reference["ms"] / o["ms"], 2) for o in one]}
        out["plain_ms_per_iteration"] = round(plain_ms_it, 4)
        out["plain_glups"] = round(lups / plain_ms_it / 1e6, 1)
    finally:
        for c in ctx:
            c.close()
        if net is not None:
            net.close()
    return out


def tile_sweep_leg():
    """Per-iteration SOR time of every tiling (tile_k, tile_v, wg_per_cu) in a
    grid, on four domains: the 400 x 400 x 120 city, seeded 128 x 128 x 64 and
    256 x 256 x 96 cities, and the 1000 x 1000 x 150 city (wind from 270 deg).
    Best of 3 (2 on the large city) runs of the same iterations; the default
    tiling is timed the same way.  Returns {domain: {"default": [tk, tv, wg,
    us], "rows": [[us, tk, tv, wg], ...] sorted}}."""
    from qesmc import warmstart as W

    def c400(ctx):
        city, _ = city_400x400x120()
        ctx.set_geometry(city)
        ctx.set_inflow(city.u0, city.v0, city.w0)

    def seeded(n, nz):
        return lambda ctx: W.load_problem(ctx, W.seeded_city(77, n, n, nz))

    def c1000(ctx):
        ncx, ncy, dz, blds, u_ref, _ = large_city(1)
        ctx.set_city(ncx, ncy, 1.0, dz, blds, (1, u_ref, 10.0, 0.5, 270.0))

    tk_all = (2, 3, 4, 5, 6, 7, 8, 10, 12, 16)
    plan = [("400x400x120", c400, 500, 3, (1, 2, 4), tk_all, (3, 4, 5, 6, 7, 8)),
            ("128x128x64", seeded(128, 64), 500, 3, (1, 2), (1, 2, 3, 4, 6, 8), (4, 6, 8, 10, 12)),
            ("256x256x96", seeded(256, 96), 500, 3, (1, 2), (2, 3, 4, 6, 8, 10), (4, 5, 6, 7, 8)),
            ("1000x1000x150", c1000, 100, 2, (1, 2), (4, 6, 8, 10, 12), (4, 5, 6, 7, 8))]
    out = {}
    for name, setup, its, reps, tvs, tks, wgs in plan:
        def timed(**tile):
            best = 1e300
            for _ in range(reps):
                ctx.set_lambda(None)
                best = min(best, ctx.run(Q.default_params(itermax=its, tol=-1.0, **tile))["ms_sor"])
            return round(best / its * 1e3, 2)

        with Q.Context(0) as ctx:
            setup(ctx)
            d = ctx.run(Q.default_params(itermax=its, tol=-1.0))
            rows = [[timed(tile_k=tk, tile_v=tv, wg_per_cu=wg), tk, tv, wg]
                    for tv in tvs for tk in tks for wg in wgs]
            out[name] = {"default": [d["tile_k"], d["tile_v"], d["wg_per_cu"], timed()],
                         "rows": sorted(rows)}
        print("tile sweep %s: default %s, best %s" % (name, out[name]["default"], out[name]["rows"][0]),
              file=sys.stderr, flush=True)
    return out


def reference_schedule(ctx, params, steps, warmup):
    """`warmup` untimed, then `steps` timed solves of the reference schedule
    on the resident city: lambda reset, R, SOR, correction, zeroing (and the
    stencil encoding, which params asks for).  Returns (wall s, results)."""
    for _ in range(warmup):
        ctx.set_lambda(None)
        ctx.run(params)
    results = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.set_lambda(None)
        results.append(ctx.run(params))
    return time.perf_counter() - t0, results


def roofline_of(results, lups_per_iteration, workload_key):
    """Roofline object of the dominant SOR kernel from the run's own HIP-event
    timings: achieved = algorithmic bytes per launch / mean launch time."""
    last = results[-1]
    its = last["iterations"]
    fused = last["path"] == Q.PATH_FUSED
    launches = its if fused else 2 * its
    ms_sor = float(np.mean([r["ms_sor"] for r in results]))
    launch_ms = ms_sor / launches
    bpl = bytes_per_lup(last)
    per_launch_updates = lups_per_iteration / (1 if fused else 2)
    achieved = bpl * per_launch_updates / (launch_ms * 1e-3) / 1e9
    if last["path"] == Q.PATH_PLAIN:
        kernel = "k_rb_plain"
    elif fused:
        kernel = "k_sor_march<%d>" % (8 if last["tile_k"] >= 8 else 4)
    else:
        kernel = "k_sor_split<%s, %d>" % ("true" if last["compressed"] else "false", last["tile_v"])
    traffic, source = load_traffic(kernel, workload_key)
    ref36 = 36.0 * per_launch_updates / (launch_ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": source, "bytes_per_lup": bpl, "avg_launch_us": round(launch_ms * 1e3, 2),
            "updates_per_launch": int(per_launch_updates),
            "at_36_bytes_per_lup": {"achieved": round(ref36, 1), "frac": round(ref36 / HBM_PEAK_GBS, 4),
                                    "note": "six fp32 coefficients + R + lambda r/w per update, the reference "
                                            "layout's algorithmic bytes; this kernel reads a 1-byte wall mask "
                                            "instead of the coefficients, bitwise equivalent"}}, ms_sor


def load_traffic(kernel, workload_key):
    """HBM bytes per dispatch of `kernel` on `workload_key` from the committed
    counter summary (profiles/traffic.json: rocprofv3 FETCH_SIZE x the copy
    calibration factor + WRITE_SIZE)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None, None
    ent = t.get(workload_key, {}).get(kernel)
    if ent is None:
        return None, None
    return ent["bytes"], "profiles/" + ent["summary"] + " (not measured in this run)"


def city400_leg(path, tiles, steps=5, warmup=1):
    """The reference schedule on the 400 x 400 x 120 city."""
    city, desc = city_400x400x120()
    params = Q.default_params(itermax=500, tol=1e-9, path=path, flags=Q.RUN_REENCODE, **tiles)
    with Q.Context(0) as ctx:
        ctx.set_geometry(city)
        ctx.set_inflow(city.u0, city.v0, city.w0)
        wall, results = reference_schedule(ctx, params, steps, warmup)
        div_ref = ctx.divergence_linf(False)
        grad_ms = min(ctx.velocity_gradient(uniform_z=False, copy=False)[1] for _ in range(3))
    its = results[-1]["iterations"]
    roof, ms_sor = roofline_of(results, city.interior_lups(), "city400x400x120")
    return {"workload": desc, "steps": steps, "t_ref_ms": round(wall / steps * 1e3, 3), "sor_ms": round(ms_sor, 3),
            "glups": round(city.interior_lups() * its * steps / wall / 1e9, 3), "div_linf_ref_operator": div_ref,
            "velocity_gradient": {"ms": round(grad_ms, 4), "bytes_per_cell": 52,
                                  "achieved_gbs": round(52.0 * city.ncell / (grad_ms * 1e-3) / 1e9, 1)},
            "roofline": roof}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--itermax", type=int, default=500)
    ap.add_argument("--path", choices=["auto", "plain", "split", "fused"], default="auto")
    ap.add_argument("--tile-k", type=int, default=0, help="k planes per SOR work-group (0: library choice)")
    ap.add_argument("--tile-v", type=int, default=0, help="split path: half-cells per lane (0: library choice)")
    ap.add_argument("--wg-per-cu", type=int, default=0, help="split path: work-groups per CU (0: library choice)")
    ap.add_argument("--slab-planes", type=int, default=0,
                    help="split path: k-slab of the interleaved red/black schedule (0: library choice, -1: whole "
                         "passes)")
    ap.add_argument("--legs", default="cpu,city1000_tol,city400,ttol,batch",
                    help="extra legs after the headline (comma list of cpu, city1000_tol, city400, ttol, batch; "
                         "'none' for the headline alone)")
    ap.add_argument("--skip-headline", action="store_true",
                    help="run only the legs named by --legs (no headline, no bench line; for profiling a leg)")
    ap.add_argument("--ttol-cities", type=int, default=8, help="time_to_tol leg: held-out cities")
    ap.add_argument("--batch", type=int, default=64, help="batch leg: instances")
    ap.add_argument("--tile-sweep", action="store_true",
                    help="only time a grid of SOR tilings on four domain sizes (one JSON line, about a minute)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run 5 plain device copies of this many MiB (HBM counter calibration under rocprofv3)")
    a = ap.parse_args()

    if a.gpus != 1:
        print(f"bench.py: --gpus {a.gpus} refused: this build targets one GPU; "
              "sharding across GPUs is not built", file=sys.stderr)
        sys.exit(2)

    if a.tile_sweep:
        print(json.dumps({"tile_sweep_us_per_iteration": tile_sweep_leg()}))
        return

    if a.copy_calibration_mib:
        # A copy of known size through torch's device copy kernel: its
        # FETCH_SIZE / WRITE_SIZE per dispatch calibrate the counters.
        import torch

        n = a.copy_calibration_mib * (1 << 20) // 4
        src = torch.arange(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4}))
        return

    legs = [] if a.legs == "none" else [x for x in a.legs.split(",") if x]
    path = {"auto": Q.PATH_AUTO, "plain": Q.PATH_PLAIN, "split": Q.PATH_SPLIT, "fused": Q.PATH_FUSED}[a.path]
    tiles = dict(tile_k=a.tile_k, tile_v=a.tile_v, wg_per_cu=a.wg_per_cu, slab_planes=a.slab_planes)
    params = Q.default_params(itermax=a.itermax, tol=1e-9, path=path, flags=Q.RUN_REENCODE, **tiles)
    ncx, ncy, dz, blds, u_ref, _ = large_city()
    desc = ("1000x1000x150 cells (151 M), dx=dy=1 m, dz 1 m x1.02 geometric, %d lognormal buildings (median 20 m, "
            "sigma 0.6, cap 120 m, footprints U(6,16) m), log law u*=0.5 m/s z0=0.5 m from 270 deg; R + %d red-black "
            "SOR iterations + correction" % (len(blds), a.itermax))
    if a.skip_headline:
        out = {"metric": METRIC, "legs_only": legs}
        run_legs(a, legs, out, path, tiles, None)
        print(json.dumps(out))
        return
    cpu = None
    with Q.Context(0) as ctx:
        shape = ctx.set_city(ncx, ncy, 1.0, dz, blds, (1, u_ref, 10.0, 0.5, 270.0))
        wall, results = reference_schedule(ctx, params, a.steps, a.warmup)
        div_ref = ctx.divergence_linf(False)
        div_phys = ctx.divergence_linf(True)
        if "cpu" in legs:
            cpu = cpu_baseline_resident(ctx, "1000x1000x150 city, wind 270 deg")
    lups_it = shape.interior_lups()
    its = results[-1]["iterations"]
    last = results[-1]
    roof, ms_sor = roofline_of(results, lups_it, "city1000x1000x150")
    glups = lups_it * its * a.steps / wall / 1e9
    out = {
        "metric": METRIC,
        "value": round(glups, 3),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded city generator, inputs resident in HBM)",
        "config": {"workload": desc, "cells": shape.ncell, "updates_per_iteration": lups_it,
                   "iterations": its, "omega": 1.78, "path": last["path"],
                   "wall_mask_stencil": bool(last["compressed"]),
                   "tile": [last["tile_k"], last["tile_v"], last["wg_per_cu"]]},
        "t_ref_ms": round(wall / a.steps * 1e3, 3),
        "sor_ms": round(ms_sor, 3),
        "div_linf_ref_operator": div_ref,
        "div_linf_physical": div_phys,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(f"headline: {out['value']} GLUPS, {out['ms_per_step']} ms per solve", file=sys.stderr, flush=True)
    run_legs(a, legs, out, path, tiles, (lups_it / (cpu["value"] * 1e6)) if cpu else None)
    print(json.dumps(out))


def run_legs(a, legs, out, path, tiles, cpu_ms_per_it):
    """The extra legs named in `legs`, each into its key of `out`."""
    if "city1000_tol" in legs:
        out["city1000_tol"] = c5_tol_leg(cpu_ms_per_it=cpu_ms_per_it)
    if "city400" in legs:
        out["city400"] = city400_leg(path, tiles)
    if "ttol" in legs and a.ttol_cities > 0:
        from qesmc import warmstart as W

        seeds = list(range(1_000_000, 1_000_000 + a.ttol_cities))
        fp64_ref = fp64_reference_fields({s_: W.seeded_city(s_) for s_ in seeds})
        out["time_to_tol"] = time_to_tol_leg(seeds, fp64_ref)
        out["time_to_tol"]["fp64_oracle_iterations"] = [fp64_ref[s_][3] for s_ in seeds]
    if "batch" in legs and a.batch > 0:
        out["batch"] = batch_leg(a.batch)


if __name__ == "__main__":
    main()

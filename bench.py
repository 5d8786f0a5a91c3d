"""Benchmark of the MI355X mass-consistency solve (one GPU).

A step is one complete Solver::solve of the reference schedule on the
400 x 400 x 120 synthetic city: stencil encoding, lambda
reset, divergence R, 500 red-black SOR iterations, velocity correction and
solid zeroing, with every input already resident in HBM.  `value` is GLUPS
(cell updates per second / 1e9) over the timed steps.

  python bench.py [--steps K] [--warmup W] [--gpus 1]

Further legs reported in the same JSON line:
  time_to_tol  128 x 128 x 64 held-out cities (the warm-start workload):
               iterations and device ms until the velocity field is within
               1e-4 m/s (L_inf) of the FP64 converged oracle, cold start
               versus the trained U-Net's lambda0; U-Net MFMA throughput.
  batch        64 cities of uneven size (64-256 x 64-256 x 64-96 cells),
               wind directions 5.625 deg apart, 500 SOR iterations each:
               one batched launch per colour pass (HIP graph replay) versus
               the instances solved one after another.
  cpu_baseline the oracle's fp32 red-black restatement on the host cores
               (bounded sample); the same leg runs the FP64 oracle that the
               time_to_tol leg checks the GPU fields against.

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

import numpy as np  # noqa: E402

import qesmc as Q  # noqa: E402

METRIC = "time-to-tol per domain (ms) + GLUPS; U-Net warm-start vs plain SOR, same GPU"
HBM_PEAK_GBS = 8000.0


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
                         read once per iteration: the black pass takes its red
                         neighbours from the work-group's LDS)
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


def cpu_baseline(city, budget_s):
    """The oracle's fp32 red-black restatement of Solver_CPU_RB on the host
    cores (OpenMP, all threads the runtime gives it): R once, then five timed
    samples of 100 SOR iterations (fewer when the budget would be exceeded);
    `value` is the median sample's rate.  Also the serial lexicographic
    Solver_CPU restatement on one thread over a short sample."""
    O = _oracle()
    p = O.Problem(**city.arrays())
    R = O.divergence_f32(p)
    per_sample = max(10, min(100, int(100 * budget_s / 12.0)))
    lam, rates = None, []
    t_all = time.perf_counter()
    for _ in range(5):
        t0 = time.perf_counter()
        lam, it, _ = O.sor_rb_f32(p, R, lam, itermax=per_sample, tol=-1.0)
        rates.append(p.interior_lups() * it / (time.perf_counter() - t0) / 1e9)
        if time.perf_counter() - t_all > budget_s:
            break
    threads = O.num_threads()
    O.set_num_threads(1)
    try:
        t0 = time.perf_counter()
        _, it_s, _ = O.sor_serial_f32(p, R, None, itermax=5, tol=-1.0)
        serial = p.interior_lups() * it_s / (time.perf_counter() - t0) / 1e9
    finally:
        O.set_num_threads(threads)
    return {"value": round(float(np.median(rates)), 4), "unit": "GLUPS", "cores": threads, "kind": "port",
            "samples_glups": [round(r, 4) for r in rates], "cpu_model": _cpu_model(),
            "serial_single_thread_glups": round(serial, 4),
            "sample": f"400x400x120 domain, median of {len(rates)} samples of {per_sample} SOR iterations "
                      f"(oracle/qesmc_oracle.c Solver_CPU_RB restatement, OpenMP, {threads} threads); serial "
                      f"Solver_CPU restatement on 1 thread over 5 iterations"}


WEIGHTS = os.path.join(ROOT, "qes-public_amd", "qesmc", "weights", "unet_128x128x64.npz")
BF16_PEAK_TFLOPS = 2500.0


def _linf(a, b):
    return float(np.abs(a.astype(np.float64) - b).max())


STOP_TOLS = (1e-3, 3e-4, 1e-4, 6e-5, 5e-5, 4e-5, 3e-5)


def time_to_tol_leg(seeds, fp64_ref, tol_ms=1e-4, chunk=25, cap=6000):
    """Cold versus U-Net warm start on held-out 128 x 128 x 64 cities.

    Pass 1 (untimed) runs SOR in chunks of `chunk` iterations and reads the
    velocity back after each, to find the first checkpoint whose L_inf error
    against the FP64 field is <= tol_ms (and the error after the reference's
    500 iterations).  Pass 2 times that many iterations from a fresh start:
    R + SOR + correction on the device (qesmc_result.ms_total), plus the
    U-Net inference for the warm start.  Without FP64 fields the stop rule
    max|dlambda| <= 1e-4 is used instead.  With them, the library's stop rule
    (max|dlambda| <= tol) is also run at each of STOP_TOLS, recording where
    it stops and its distance to the FP64 field: which tol a production run
    needs for the 1e-4 m/s criterion."""
    from qesmc import warmstart as W

    if not os.path.exists(WEIGHTS):
        return {"skipped": "trained weights not found: " + WEIGHTS}
    net = W.load(WEIGHTS)
    rows = []
    with Q.Context(0) as ctx:
        for seed in seeds:
            city = W.seeded_city(seed)
            W.load_problem(ctx, city)
            row = {"seed": seed}
            for mode in ("cold", "warm"):
                ctx.set_lambda(None)
                if mode == "warm":
                    net.predict(ctx)
                it, it_tol, err500, err_tol = 0, None, None, None
                if fp64_ref is not None:
                    u64, v64, w64, _ = fp64_ref[seed]
                    while it < cap:
                        ctx.run(itermax=chunk, tol=-1.0)
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
                    r = ctx.run(itermax=it_tol, tol=-1.0)
                else:
                    r = ctx.run(itermax=20000, tol=1e-4)
                row[mode] = {"iterations": r["iterations"], "ms": round(r["ms_total"] + unet_ms, 3),
                             "sor_ms": round(r["ms_sor"], 3), "err_at_tol": err_tol, "err_at_500": err500}
                if mode == "warm":
                    row[mode].update(unet_ms=round(unet_ms, 3), unet_tflops=round(tflops, 1))
                if fp64_ref is not None:
                    # the library's own stop rule (max|dlambda| <= tol, as the
                    # reference's) at several tol: where it stops, and how far
                    # from the FP64 field that is
                    row[mode]["stop_rule"] = {}
                    for stol in STOP_TOLS:
                        ctx.set_lambda(None)
                        if mode == "warm":
                            net.predict(ctx)
                        r = ctx.run(itermax=cap, tol=stol)
                        u, v, w = ctx.get_velocity()
                        row[mode]["stop_rule"][str(stol)] = {
                            "iterations": r["iterations"],
                            "err": max(_linf(u, u64), _linf(v, v64), _linf(w, w64))}
            rows.append(row)
    net.close()

    def mean(mode, key):
        vals = [r[mode][key] for r in rows if r[mode].get(key) is not None]
        return float(np.mean(vals)) if vals else None

    cold_ms, warm_ms = mean("cold", "ms"), mean("warm", "ms")
    stop = None
    if fp64_ref is not None:
        stop = {}
        for mode in ("cold", "warm"):
            stop[mode] = {}
            for stol in STOP_TOLS:
                v = [r[mode]["stop_rule"][str(stol)] for r in rows]
                stop[mode][str(stol)] = {"mean_iterations": float(np.mean([x["iterations"] for x in v])),
                                         "max_err": max(x["err"] for x in v),
                                         "cities_within_tol_ms": sum(x["err"] <= tol_ms for x in v)}
    tfl = mean("warm", "unet_tflops")
    return {
        "workload": "%d held-out seeded cities, 128x128x64 cells, dx=2 m, dz 1 m x1.02, 12-30 buildings, log law, "
                    "random direction; U-Net 4 levels 16/32/64/128 trained on 2000 solver-generated cities" % len(rows),
        "criterion": ("L_inf(u,v,w) vs FP64 oracle <= %g m/s, checked every %d iterations" % (tol_ms, chunk)
                      if fp64_ref is not None else "stop rule max|dlambda| <= 1e-4 (no FP64 fields)"),
        "cold": {"iterations": mean("cold", "iterations"), "ms": cold_ms, "err_at_500": mean("cold", "err_at_500")},
        "warm": {"iterations": mean("warm", "iterations"), "ms": warm_ms, "err_at_500": mean("warm", "err_at_500"),
                 "unet_ms": mean("warm", "unet_ms")},
        "speedup_warm_vs_cold": round(cold_ms / warm_ms, 3) if cold_ms and warm_ms else None,
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


def _tol_run(ctx, ref, chunk, cap, tol_ms):
    """Run SOR on ctx in chunks from its current lambda until its velocity is
    within tol_ms (L_inf) of ref's; returns (iterations, device ms).  The ms
    are the chunks' SOR time plus one R + correction (the chunk runs repeat
    those only to read the field back)."""
    its, ms_sor, extra = 0, 0.0, None
    while its < cap:
        r = ctx.run(itermax=chunk, tol=-1.0)
        its += chunk
        ms_sor += r["ms_sor"]
        if extra is None:
            extra = r["ms_total"] - r["ms_sor"]
        if Q.velocity_linf_diff(ctx, ref) <= tol_ms:
            return its, ms_sor + extra
    return None, ms_sor + (extra or 0.0)


def large_leg(dirs, measured, ref_iters, cpu_iters, tol_ms=1e-4):
    """Time-to-tolerance of a wind-direction ensemble on the 1000 x 1000 x 150
    city.  An FP64 CPU solve of 151 M cells is out of reach (hours), so each
    measured direction's reference is the GPU's own fp32 field after
    ref_iters iterations: at this size the field stops changing (beyond fp32
    noise of ~6e-5 m/s between checkpoints 10000 iterations apart) after
    about 60000 iterations.  T_tol = device ms until the velocity is within
    tol_ms (L_inf, u, v, w) of the reference:
      cold        from lambda = 0 (checked every 2500 iterations);
      superposed  directions 270 and 180 deg (wind along +x and +y) cold; any
                  other direction from cos(theta) lambda_270 + sin(theta)
                  lambda_180 (checked every 25 iterations).
    Only the `measured` directions are solved (each costs two long runs); the
    `dirs`-direction ensemble total is extrapolated from them and reported as
    such.  Directions run one after another: at 151 M cells one problem fills
    the GPU.  Also the classic-layout path's time per iteration (the
    reference GPU algorithm) and a CPU sample of the oracle."""
    ncx, ncy, dz, blds, u_ref, _ = large_city(dirs)
    order = [270.0, 180.0] + [w for w in measured if w not in (270.0, 180.0)]
    out = {"workload": "1000x1000x150 cells (151 M), dx=dy=1 m, dz 1 m x1.02, %d lognormal buildings "
                       "(median 20 m, sigma 0.6, cap 120 m, footprints U(6,16) m), log law u*=0.5 m/s z0=0.5 m, "
                       "ensemble of %d directions %.0f deg apart" % (len(blds), dirs, 360.0 / dirs),
           "criterion": "L_inf(u,v,w) <= %g m/s of the direction's fp32 field after %d iterations"
                        % (tol_ms, ref_iters)}
    ctx = [Q.Context(0) for _ in range(4)]  # reference, work, basis x, basis y
    ref, work, bx, by = ctx
    try:
        t0 = time.perf_counter()
        for c in ctx:
            c.set_city(ncx, ncy, 1.0, dz, blds, (1, u_ref, 10.0, 0.5, 270.0))
        out["device_city_build_s_per_context"] = round((time.perf_counter() - t0) / 4, 3)
        lups = ref.city.interior_lups()
        # tiling chosen by the autotuner for this domain (kept per context)
        t0 = time.perf_counter()
        best = None
        for c in ctx:
            best = Q.autotune(c, iterations=20)
        out["autotuned_tile"] = [best.tile_k, best.tile_v, best.wg_per_cu]
        out["autotune_s"] = round(time.perf_counter() - t0, 2)
        r = work.run(itermax=100, tol=-1.0, path=Q.PATH_PLAIN)
        plain_ms_it = r["ms_sor"] / 100
        work.set_lambda(None)
        r = work.run(itermax=100, tol=-1.0)
        split_ms_it = r["ms_sor"] / 100
        out["ms_per_iteration"] = {"plain": round(plain_ms_it, 4), "split": round(split_ms_it, 4)}
        out["glups"] = {"plain": round(lups / plain_ms_it / 1e6, 1), "split": round(lups / split_ms_it / 1e6, 1)}
        print(f"large: plain {plain_ms_it:.3f} ms/it, split {split_ms_it:.3f} ms/it", file=sys.stderr, flush=True)
        rows = []
        for wd in order:
            inflow = (1, u_ref, 10.0, 0.5, wd)
            ref.set_city_inflow(inflow)
            ref.set_lambda(None)
            ref.run(itermax=ref_iters, tol=-1.0)
            work.set_city_inflow(inflow)
            work.set_lambda(None)
            c_its, c_ms = _tol_run(work, ref, 2500, ref_iters, tol_ms)
            row = {"wind_deg": wd, "cold_iterations": c_its, "cold_ms": round(c_ms, 1)}
            if wd in (270.0, 180.0):
                Q.lambda_combine(bx if wd == 270.0 else by, [work], np.array([1.0], np.float32))
                row["role"] = "basis"
            else:
                th = math.radians(270.0 - wd)
                tc = time.perf_counter()
                Q.lambda_combine(work, [bx, by], np.array([math.cos(th), math.sin(th)], np.float32))
                comb_ms = (time.perf_counter() - tc) * 1e3
                w_its, w_ms = _tol_run(work, ref, 25, 20000, tol_ms)
                row.update(role="superposed", superposed_iterations=w_its, superposed_ms=round(w_ms + comb_ms, 1))
            rows.append(row)
            print(f"large: {wd:5.1f} deg cold {c_its} its {c_ms:.0f} ms, superposed "
                  f"{row.get('superposed_iterations')} its {row.get('superposed_ms')} ms", file=sys.stderr, flush=True)
        basis = [r_ for r_ in rows if r_["role"] == "basis"]
        others = [r_ for r_ in rows if r_["role"] == "superposed"]
        cold_mean = float(np.mean([r_["cold_ms"] for r_ in rows]))
        cold_its_mean = float(np.mean([r_["cold_iterations"] or ref_iters for r_ in rows]))
        sup_mean = float(np.mean([r_["superposed_ms"] for r_ in others])) if others else 0.0
        cold_ens = dirs * cold_mean
        sup_ens = sum(r_["cold_ms"] for r_ in basis) + (dirs - 2) * sup_mean
        out["per_direction"] = rows
        out["ensemble_extrapolated"] = {
            "note": "%d measured directions; the %d-direction totals use the mean cold time of the measured ones and "
                    "the mean superposed time of the measured non-basis ones" % (len(rows), dirs),
            "cold_ms": round(cold_ens, 1), "superposed_ms": round(sup_ens, 1),
            "plain_path_cold_ms": round(dirs * cold_its_mean * plain_ms_it, 1),
            "speedup_superposed_vs_cold": round(cold_ens / sup_ens, 2),
            "speedup_superposed_vs_plain_cold": round(dirs * cold_its_mean * plain_ms_it / sup_ens, 2)}
        if cpu_iters > 0:
            city = ref.get_inputs()
            O = _oracle()
            p = O.Problem(**city.arrays())
            del city
            t0 = time.perf_counter()
            R = O.divergence_f32(p)
            O.sor_rb_f32(p, R, None, itermax=cpu_iters, tol=-1.0)
            el = time.perf_counter() - t0
            cpu_ms_it = el * 1e3 / cpu_iters
            cpu_ens = dirs * cold_its_mean * cpu_ms_it
            out["cpu_baseline"] = {"ms_per_iteration": round(cpu_ms_it, 2), "cores": O.num_threads(), "kind": "port",
                                   "glups": round(lups / cpu_ms_it / 1e6, 3),
                                   "sample": "R + %d SOR iterations (%.1f s), extrapolated to the ensemble's cold "
                                             "iteration counts" % (cpu_iters, el),
                                   "ensemble_cold_ms_extrapolated": round(cpu_ens, 0)}
            out["ensemble_extrapolated"]["speedup_superposed_vs_cpu"] = round(cpu_ens / sup_ens, 1)
    finally:
        for c in ctx:
            c.close()
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


def load_traffic(workload_key):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            return json.load(fh).get(workload_key)
    except (OSError, ValueError):
        return None


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
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU baseline sampling (0: skip)")
    ap.add_argument("--ttol-cities", type=int, default=8, help="time_to_tol leg: held-out cities (0: skip)")
    ap.add_argument("--batch", type=int, default=64, help="batch leg: instances (0: skip)")
    ap.add_argument("--workload", choices=["city400", "city1000"], default="city400",
                    help="city1000: only the 1000x1000x150 8-direction ensemble (minutes of GPU time)")
    ap.add_argument("--large-dirs", type=int, default=8, help="ensemble size of the large workload")
    ap.add_argument("--large-measure", default="270,180,45,135", help="directions solved (deg); 270 and 180 are the basis")
    ap.add_argument("--large-ref-iters", type=int, default=100000, help="iterations of each direction's reference")
    ap.add_argument("--large-cpu-iters", type=int, default=20)
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

    if a.workload == "city1000":
        measured = [float(v) for v in a.large_measure.split(",")]
        out = large_leg(a.large_dirs, measured, a.large_ref_iters, a.large_cpu_iters)
        print(json.dumps({"metric": METRIC, "workload": "city1000", "n_gpus": 1, "city1000": out}))
        return

    city, desc = city_400x400x120()
    path = {"auto": Q.PATH_AUTO, "plain": Q.PATH_PLAIN, "split": Q.PATH_SPLIT, "fused": Q.PATH_FUSED}[a.path]
    params = Q.default_params(itermax=a.itermax, tol=1e-9, path=path, flags=Q.RUN_REENCODE, tile_k=a.tile_k,
                              tile_v=a.tile_v, wg_per_cu=a.wg_per_cu)
    with Q.Context(0) as ctx:
        ctx.set_geometry(city)
        ctx.set_inflow(city.u0, city.v0, city.w0)
        for _ in range(a.warmup):
            ctx.set_lambda(None)
            ctx.run(params)
        results = []
        t0 = time.perf_counter()
        for _ in range(a.steps):
            ctx.set_lambda(None)
            results.append(ctx.run(params))
        wall = time.perf_counter() - t0
        div_ref = ctx.divergence_linf(False)
        # velocity gradient tensor of the corrected field (QES-Turb derivative step)
        grad_ms = min(ctx.velocity_gradient(uniform_z=False, copy=False)[1] for _ in range(3))

    its = results[-1]["iterations"]
    lups = city.interior_lups() * its
    glups = lups * a.steps / wall / 1e9
    ms_step = wall / a.steps * 1e3
    ms_sor = float(np.mean([r["ms_sor"] for r in results]))
    last = results[-1]
    fused = last["path"] == Q.PATH_FUSED
    launches = its if fused else 2 * its
    launch_ms = ms_sor / launches
    bpl = bytes_per_lup(last)
    launch_bytes = bpl * city.interior_lups() / (1 if fused else 2)
    achieved = launch_bytes / (launch_ms * 1e-3) / 1e9
    if last["path"] == Q.PATH_PLAIN:
        kernel = "k_rb_plain"
    elif fused:
        kernel = "k_sor_fused"
    else:
        kernel = "k_sor_split<%s, %d>" % ("true" if last["compressed"] else "false", last["tile_v"])
    out = {
        "metric": METRIC,
        "value": round(glups, 3),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded city generator, inputs resident in HBM)",
        "config": {"workload": desc, "cells": city.ncell, "updates_per_iteration": city.interior_lups(),
                   "iterations": its, "omega": 1.78, "path": results[-1]["path"],
                   "wall_mask_stencil": bool(last["compressed"]),
                   "tile": [last["tile_k"], last["tile_v"], last["wg_per_cu"]]},
        "t_ref_ms": round(ms_step, 3),
        "sor_ms": round(ms_sor, 3),
        "div_linf_ref_operator": div_ref,
        "velocity_gradient": {"ms": round(grad_ms, 4), "bytes_per_cell": 52,
                              "achieved_gbs": round(52.0 * city.ncell / (grad_ms * 1e-3) / 1e9, 1)},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": load_traffic(kernel), "bytes_per_lup": bpl,
                     "avg_launch_us": round(launch_ms * 1e3, 2)},
        "cpu_baseline": None,
    }
    fp64_ref = None
    seeds = list(range(1_000_000, 1_000_000 + a.ttol_cities))
    if a.cpu_budget > 0:
        out["cpu_baseline"] = cpu_baseline(city, a.cpu_budget)
        if seeds:
            from qesmc import warmstart as W

            fp64_ref = fp64_reference_fields({s_: W.seeded_city(s_) for s_ in seeds})
    if seeds:
        out["time_to_tol"] = time_to_tol_leg(seeds, fp64_ref)
        if fp64_ref is not None:
            out["time_to_tol"]["fp64_oracle_iterations"] = [fp64_ref[s_][3] for s_ in seeds]
    if a.batch > 0:
        out["batch"] = batch_leg(a.batch)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

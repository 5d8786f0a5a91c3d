"""Benchmark of the MI355X mass-consistency solve (one GPU).

A step is one complete Solver::solve of the reference schedule on the
400 x 400 x 120 synthetic city (configuration 2): stencil encoding, lambda
reset, divergence R, 500 red-black SOR iterations, velocity correction and
solid zeroing, with every input already resident in HBM.  `value` is GLUPS
(cell updates per second / 1e9) over the timed steps.

  python bench.py [--steps K] [--warmup W] [--gpus 1]

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


def config2_city():
    """400 x 400 x 120 cells, dx = dy = 2 m, dz geometric from 1 m at ratio
    1.02, 200 seeded buildings (footprints U(10,40) m, heights U(10,60) m),
    log law with u* = 0.5 m/s, z0 = 0.5 m (U_ref at 10 m chosen to give that
    u*), wind from 217 deg."""
    ncx, ncy, ncz, dx = 400, 400, 120, 2.0
    dz = Q.dz_geometric(ncz, 1.0, 1.02)
    blds = Q.random_buildings(2024, ncx, ncy, dx, 200, 10.0, 40.0, 10.0, 60.0, margin=2)
    u_ref = 0.5 / 0.4 * math.log((10.0 + 0.5) / 0.5)
    city = Q.city_build(ncx, ncy, dx, dz, blds, (1, u_ref, 10.0, 0.5, 217.0))
    desc = ("config2: 400x400x120 cells, dx=dy=2 m, dz 1 m x1.02 geometric, %d seeded buildings, "
            "log law u*=0.5 m/s z0=0.5 m from 217 deg; R + 500 red-black SOR iterations + correction"
            % len(blds))
    return city, desc


def bytes_per_lup(res):
    """Algorithmic HBM bytes per lattice update of the dominant SOR kernel
    (one colour pass updates half the cells):
      split, wall mask : own lambda r+w 8 + R 4 + mask 1 + other-colour lambda 4 = 17
      split, full coef : 8 + 4 + 24 + 4 = 40
      plain            : coefficient and R lines of both colours are fetched by
                         each pass: 2*(24+4) + lambda r (4, neighbours shared) + w 4 = 64"""
    if res["path"] == Q.PATH_PLAIN:
        return 64.0
    return 17.0 if res["compressed"] else 40.0


def cpu_baseline(city, budget_s):
    """The oracle's fp32 red-black restatement of Solver_CPU_RB on the host
    cores: R once, then chunks of 10 SOR iterations until `budget_s` has
    elapsed (a bounded sample of the 500-iteration schedule)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    p = O.Problem(**city.arrays())
    t0 = time.perf_counter()
    R = O.divergence_f32(p)
    lam, its = None, 0
    while True:
        lam, it, _ = O.sor_rb_f32(p, R, lam, itermax=10, tol=0.0)
        its += it
        el = time.perf_counter() - t0
        if el >= budget_s or its >= 500:
            break
    glups = p.interior_lups() * its / el / 1e9
    return {"value": round(glups, 4), "unit": "GLUPS", "cores": O.num_threads(), "kind": "port",
            "sample": f"config2 domain, R + {its} of 500 SOR iterations ({el:.1f} s), oracle/qesmc_oracle.c "
                      f"Solver_CPU_RB restatement, OpenMP"}


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
    ap.add_argument("--path", choices=["auto", "plain", "split"], default="auto")
    ap.add_argument("--tile-k", type=int, default=0, help="k planes per SOR work-group (0: library choice)")
    ap.add_argument("--tile-v", type=int, default=0, help="split path: half-cells per lane (0: library choice)")
    ap.add_argument("--wg-per-cu", type=int, default=0, help="split path: work-groups per CU (0: library choice)")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU baseline sampling (0: skip)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run 5 plain device copies of this many MiB (HBM counter calibration under rocprofv3)")
    a = ap.parse_args()

    if a.gpus != 1:
        print(f"bench.py: --gpus {a.gpus} refused: this build targets one GPU; "
              "multi-GPU sharding is described in DESIGN.md but not built", file=sys.stderr)
        sys.exit(2)

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

    city, desc = config2_city()
    path = {"auto": Q.PATH_AUTO, "plain": Q.PATH_PLAIN, "split": Q.PATH_SPLIT}[a.path]
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

    its = results[-1]["iterations"]
    lups = city.interior_lups() * its
    glups = lups * a.steps / wall / 1e9
    ms_step = wall / a.steps * 1e3
    ms_sor = float(np.mean([r["ms_sor"] for r in results]))
    launches = 2 * its
    launch_ms = ms_sor / launches
    bpl = bytes_per_lup(results[-1])
    launch_bytes = bpl * city.interior_lups() / 2
    achieved = launch_bytes / (launch_ms * 1e-3) / 1e9
    last = results[-1]
    if last["path"] == Q.PATH_PLAIN:
        kernel = "k_rb_plain"
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
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": load_traffic(kernel), "bytes_per_lup": bpl,
                     "avg_launch_us": round(launch_ms * 1e3, 2)},
        "cpu_baseline": None,
    }
    if a.cpu_budget > 0:
        out["cpu_baseline"] = cpu_baseline(city, a.cpu_budget)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""Learned warm start pipeline on top of the C ABI (qesmc_unet_*).

  cities      seeded flat 128 x 128 x 64-cell cities (12-30 buildings, log law)
  dataset     converged SOR lambda of each city, captured on the device
  training    Adam on the device, BF16 MFMA convolutions
  comparison  SOR from lambda = 0 (cold) versus SOR from the U-Net's lambda0
              (warm), both stopped by the same rule max|dlambda| <= tol

Everything runs through libqesmc; nothing here computes the solve.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import Context, UNet, city_build, default_params, dz_geometric, random_buildings


def seeded_city(seed: int, ncx=128, ncy=128, ncz=64, dx=2.0):
    """A seeded city: 12-30 non-overlapping rectangles (footprints
    U(8, 30) m, heights U(8, 40) m), dz geometric from 1 m at ratio 1.02,
    log-law inflow u* = 0.5 m/s, z0 = 0.5 m, direction uniform in 0-360 deg."""
    rng = np.random.default_rng(seed)
    nb = int(rng.integers(12, 31))
    blds = random_buildings(seed, ncx, ncy, dx, nb, 8.0, 30.0, 8.0, 40.0, margin=2)
    u_ref = 0.5 / 0.4 * math.log((10.0 + 0.5) / 0.5)
    wind = float(rng.uniform(0.0, 360.0))
    return city_build(ncx, ncy, dx, dz_geometric(ncz, 1.0, 1.02), blds, (1, u_ref, 10.0, 0.5, wind))


def load_problem(ctx: Context, city):
    ctx.set_geometry(city)
    ctx.set_inflow(city.u0, city.v0, city.w0)


def build_dataset(net: UNet, ctx: Context, seeds, iters=3000, tol=1e-7, log=None):
    """Solve each city to `tol` (at most `iters` iterations) and capture
    (features, lambda) as one training sample.  Returns seconds spent."""
    t0 = time.perf_counter()
    params = default_params(itermax=iters, tol=tol)
    for q, s in enumerate(seeds):
        city = seeded_city(s, *net_size(net))
        load_problem(ctx, city)
        ctx.set_lambda(None)
        ctx.run(params)
        net.add_sample(ctx)
        if log and (q + 1) % 100 == 0:
            log(f"dataset {q + 1}/{len(seeds)} ({time.perf_counter() - t0:.1f} s)")
    return time.perf_counter() - t0


def net_size(net: UNet):
    return net.ncx, net.ncy, net.ncz


def train(net: UNet, epochs: float, lr=2e-3, lr_min=2e-4, chunk=200, seed=0, log=None):
    """Adam over all stored samples, cosine-decayed learning rate, in chunks
    of `chunk` steps.  Returns (final chunk mean loss, device ms)."""
    n = net.num_samples()
    steps = int(epochs * n)
    done, ms_total, loss = 0, 0.0, float("nan")
    while done < steps:
        k = min(chunk, steps - done)
        frac = (done + 0.5 * k) / max(steps, 1)
        lr_now = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * frac))
        loss, ms = net.train(k, lr_now, 0, n, seed=seed + done)
        ms_total += ms
        done += k
        if log:
            log(f"train step {done}/{steps} lr {lr_now:.2e} loss {loss:.4e}")
    return loss, ms_total


def save(net: UNet, path: str, **meta):
    np.savez_compressed(path, params=net.get_params(), target_scale=np.float32(net.target_scale()),
                        levels=net.levels, base_channels=net.base_channels,
                        size=np.array([net.ncx, net.ncy, net.ncz]), **meta)


def load(path: str, device=0, size=None) -> UNet:
    """Trained network from `path`; `size` = (ncx, ncy, ncz) instantiates the
    same (fully convolutional) weights for another domain size."""
    z = np.load(path)
    ncx, ncy, ncz = (int(v) for v in (z["size"] if size is None else size))
    net = UNet(ncx, ncy, ncz, levels=int(z["levels"]), base_channels=int(z["base_channels"]), device=device,
               target_scale=float(z["target_scale"]))
    net.set_params(z["params"])
    return net


def compare(net: UNet, ctx: Context, seeds, tol, itermax=20000, check_every=16):
    """Cold versus warm start on held-out cities, same stop rule.
    Returns a list of dicts with iterations and device milliseconds."""
    out = []
    params = default_params(itermax=itermax, tol=tol, check_every=check_every)
    for s in seeds:
        city = seeded_city(s, *net_size(net))
        load_problem(ctx, city)
        ctx.set_lambda(None)
        cold = ctx.run(params)
        ms_unet, tflops = net.predict(ctx)
        warm = ctx.run(params)
        out.append({"seed": s, "cold_iters": cold["iterations"], "cold_ms": cold["ms_total"],
                    "warm_iters": warm["iterations"], "warm_ms": warm["ms_total"] + ms_unet,
                    "unet_ms": ms_unet, "unet_tflops": tflops})
    return out


def _main(argv=None):
    import argparse
    import json
    import sys

    ap = argparse.ArgumentParser(prog="python -m qesmc.warmstart")
    sub = ap.add_subparsers(dest="cmd", required=True)
    tr = sub.add_parser("train", help="build the dataset with the solver and train the U-Net on the GPU")
    tr.add_argument("--samples", type=int, default=2000)
    tr.add_argument("--epochs", type=float, default=4.0)
    tr.add_argument("--lr", type=float, default=2e-3)
    tr.add_argument("--lr-min", type=float, default=2e-4, help="end of the cosine learning-rate schedule")
    tr.add_argument("--init", default=None, help="start from these weights (a file written by this command)")
    tr.add_argument("--size", type=int, nargs=3, default=[128, 128, 64])
    tr.add_argument("--solve-iters", type=int, default=2500)
    tr.add_argument("--smooth-block", type=int, default=8, help="block edge of the block-mean loss term")
    tr.add_argument("--smooth-weight", type=float, default=0.0, help="weight of the block-mean loss term (0: off)")
    tr.add_argument("--out", required=True)
    cp = sub.add_parser("compare", help="cold versus warm start on held-out cities")
    cp.add_argument("--weights", required=True)
    cp.add_argument("--cities", type=int, default=8)
    cp.add_argument("--tol", type=float, default=1e-4)
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    if a.cmd == "train":
        ncx, ncy, ncz = a.size
        z = np.load(a.init) if a.init else None
        # continuing from saved weights keeps their output unit
        net = UNet(ncx, ncy, ncz, levels=4, base_channels=16, seed=1,
                   target_scale=float(z["target_scale"]) if z is not None else 0.0)
        if z is not None:
            net.set_params(z["params"])
        if a.smooth_weight > 0:
            net.set_loss(a.smooth_block, a.smooth_weight)
        with Context(0) as ctx:
            t = build_dataset(net, ctx, range(a.samples), iters=a.solve_iters, tol=1e-9, log=log)
            log(f"dataset: {a.samples} cities in {t:.1f} s")
            loss, ms = train(net, a.epochs, lr=a.lr, lr_min=a.lr_min, log=log)
            log(f"training: {int(a.epochs * a.samples)} steps, {ms / 1e3:.1f} s on the device, last loss {loss:.4e}")
        save(net, a.out, samples=a.samples, epochs=a.epochs, final_loss=loss, train_device_s=ms / 1e3,
             smooth_block=a.smooth_block, smooth_weight=a.smooth_weight, lr=a.lr, lr_min=a.lr_min,
             init=a.init or "")
        log(f"saved {a.out}")
        return 0
    net = load(a.weights)
    with Context(0) as ctx:
        res = compare(net, ctx, range(1_000_000, 1_000_000 + a.cities), a.tol)
    summary = {k: float(np.mean([r[k] for r in res])) for k in res[0] if k != "seed"}
    json.dump({"tol": a.tol, "cities": a.cities, "mean": summary, "per_city": res}, sys.stdout)
    print()
    return 0


if __name__ == "__main__":
    raise SystemExit(_main())

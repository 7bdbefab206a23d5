"""Headline benchmark: D3Q19 BGK MLUPS and fraction of the MI355X HBM roofline.

Workload (N=1): 256^3 periodic box, D3Q19 BGK, tau = 0.6, FP32 populations
(stored shifted, f - w), Taylor-Green vortex u0 = 0.05 plus seeded Gaussian
noise sigma = 1e-3, rho = 1.  A step is one fused collide+stream launch over
all 16.8M cells.  W warm-up steps, then K timed steps bracketed by
synchronisation; the per-launch time of the step kernel comes from HIP
events recorded on the stream the kernel is launched on.

Algorithmic bytes per lattice update: 2 * 19 * 4 = 152 B (one read and one
write of every population); achieved = 152 B * cells / launch time.
The CPU baseline is the project's FP64 oracle (OpenMP) on a bounded sample
of the same workload (64^3 cells, a few steps), run on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def cpu_baseline(seconds_target=12.0):
    import numpy as np

    from oracle import oracle as O

    n = 64
    shape = (n, n, n)
    rho0, u0 = O.taylor_green_noise(shape, 0.05, 1e-3, seed=1234)
    p = O.Problem(shape, omega=1 / 0.6)
    p.init(rho0, u0)
    p.step(2)
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < seconds_target:
        p.step(5)
        steps += 5
    dt = time.perf_counter() - t0
    mlups = n ** 3 * steps / dt / 1e6
    return {"value": round(mlups, 3), "unit": "MLUPS", "cores": O.num_threads(), "kind": "port",
            "sample": f"FP64 OpenMP oracle, D3Q19 periodic TGV {n}^3, {steps} steps ({dt:.1f} s)"}


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", name)
    if os.path.exists(path):
        with open(path) as f:
            d = json.load(f)
        return d.get("hbm_bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "bf16", "f64"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="time only the device copy kernel over the step kernel's algorithmic bytes")
    args = ap.parse_args()
    if args.copy_calibration:
        from lbm_amd import capi

        nbytes = 2 * 19 * capi.BYTES[args.storage] * args.n ** 3 // 2  # read + write = 2x this
        ms = capi.copy_bench(nbytes, args.steps)
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "ms_per_launch": ms,
                          "GBps": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import numpy as np

    from lbm_amd import capi

    n = args.n
    shape = (n, n, n)
    # Taylor-Green + noise initial field, generated identically to the oracle's
    from oracle.oracle import taylor_green_noise

    rho0, u0 = taylor_green_noise(shape, 0.05, 1e-3, seed=1234)
    lat = capi.Lattice(shape, storage=args.storage, omega=1 / 0.6)
    lat.init(rho0, u0)
    del rho0, u0
    lat.step(args.warmup)
    lat.sync()
    t0 = time.perf_counter()
    ms = lat.time_steps(args.steps)  # HIP events on the kernel's stream
    lat.sync()
    wall = time.perf_counter() - t0
    cells = n ** 3
    ms_per_step = ms / args.steps
    mlups = cells / (ms_per_step * 1e-3) / 1e6
    bpp = capi.BYTES[args.storage]
    alg_bytes = 2 * 19 * bpp * cells
    achieved = alg_bytes / (ms_per_step * 1e-3) / 1e9
    s = lat.reduce()
    lat.close()
    traffic = load_traffic(f"traffic_d3q19_{args.storage}_{n}.json")
    line = {
        "metric": "D3Q19 MLUPS + % HBM roofline; time-to-1e-6 residual vs plain FP32 GPU",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.storage,
        "data": "synthetic (Taylor-Green vortex u0=0.05 + Gaussian noise sigma=1e-3, seed 1234)",
        "config": {"workload": f"D3Q19 BGK periodic box {n}^3, tau=0.6, {args.storage} populations (shifted f-w), "
                               f"fused pull collide+stream", "cells": cells, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "alg_bytes_per_launch": alg_bytes},
        "wall_s_timed": round(wall, 4),
        "mass_after": s.mass,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

"""Headline benchmark: D3Q19 BGK MLUPS and fraction of the MI355X HBM roofline.

Workload (N=1): 256^3 periodic box, D3Q19 BGK, tau = 0.6, FP32 populations
(stored shifted, f - w), Taylor-Green vortex u0 = 0.05 plus seeded Gaussian
noise sigma = 1e-3, rho = 1.  A step is one fused collide+stream launch over
all 16.8M cells.  W warm-up steps, untimed settling groups until the step
time is stable (reported in "settle"), then K timed steps bracketed by
synchronisation, run without a break as five repetitions with HIP events
recorded between them on the stream the kernel is launched on; the reported
time per step is the median over the repetitions (min, max and the mean
over all K steps beside it).

Algorithmic bytes per lattice update: 2 * 19 * 4 = 152 B (one read and one
write of every population); achieved = 152 B * cells / launch time.
The CPU baseline is the project's FP64 oracle (OpenMP) on a bounded sample
of the same workload (64^3 cells, a few steps), run on this host.

After the headline, "time_to_residual" holds the metric's second half:
the sphere channel to the relative residual 1e-6 in FP32, and the learned
warm start against the plain FP32 start (wall-time ratio to 1e-6) on unseen
obstacle channels, with a network trained in the same run (--ttr 0 skips it).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
# arithmetic type of the step for each population storage form: FP32 and FP64
# storage collide in FP64, the 16-bit forms in FP32
COMPUTE = {"f32": "f64", "f64": "f64", "f16": "f32", "bf16": "f32"}


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


def repetitions(k, r=5):
    """Split K timed steps into r repetitions (all K steps are timed as one
    uninterrupted run; the median per-step time over the repetitions is
    reported).  Equal when r divides K, otherwise the last K mod r
    repetitions take one step more."""
    r = max(1, min(r, k))
    base, extra = divmod(k, r)
    return [base + (1 if i >= r - extra else 0) for i in range(r)]


def settle(lat, group=10, max_groups=30, tol=0.02):
    """Untimed groups of steps after the warm-up until the step time is
    stable: three consecutive groups within tol of each other.  On MI355X
    the first ~10 ms of full HBM load run up to 12 % slower than the
    sustained rate (a kernel trace shows steps 6-25 of a fresh run at
    464-489 us against 434 us before and after), so a short timed window
    right after a 5-step warm-up measures that transient, not the kernel."""
    hist = []
    for _ in range(max_groups):
        hist.append(float(lat.time_step_groups([group])[0]) / group)
        if len(hist) >= 3 and max(hist[-3:]) <= min(hist[-3:]) * (1 + tol):
            break
    return group * len(hist), hist


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", name)
    if os.path.exists(path):
        with open(path) as f:
            d = json.load(f)
        return d.get("hbm_bytes_per_launch")
    return None


def run_batch(args):
    """Config 4: 2048 uneven backward-facing-step channels (D2Q9) advanced in one
    persistent launch per step, 200-step cycles captured in a hipGraph, each
    domain retired once its relative residual < 1e-6."""
    import numpy as np

    from lbm_amd import capi
    from lbm_amd import workloads as W

    doms = W.backward_step_batch(n=args.ndom, seed=2024)
    b = capi.Batch(doms, storage=args.storage)
    rho, u = zip(*[W.plug_start(d, 0.03) for d in doms])
    rho, u = np.concatenate(rho), np.concatenate(u)
    # the tuner times steps of the flow it will run (it advances the state,
    # so the fields are set again afterwards)
    b.init(rho, u)
    t0 = time.perf_counter()
    tx, blocks, one, tms = b.tune(40)
    tune_s = time.perf_counter() - t0
    b.init(rho, u)
    ms, done_steps = 0.0, 0
    while done_steps < args.max_steps:
        chunk = min(20000, args.max_steps - done_steps)
        steps, res, mass, dms = b.run(200, 1e-6, chunk)
        ms += dms
        done_steps += chunk
        nconv = int((steps != -1).sum())
        print(f"# batch {done_steps} steps: {nconv}/{len(doms)} retired, {ms / 1e3:.1f} s", flush=True)
        if nconv == len(doms):
            break
    info = b.info()
    b.close()
    cells = np.array([d["shape"][0] * d["shape"][1] for d in doms])
    # solid cells are never updated: updates and bytes count fluid cells, each
    # update reading and writing its 9 populations once, +1 B for the flag
    # field these domains carry (all-cell rates reported beside)
    fluid = np.array([int((np.asarray(d["solid"]) == 0).sum()) for d in doms])
    done = steps.copy()
    done[done < 0] = args.max_steps
    updates = float((done * fluid).sum())
    glups = updates / (ms * 1e-3) / 1e9
    glups_all = float((done * cells).sum()) / (ms * 1e-3) / 1e9
    bpp = capi.BYTES[args.storage]
    achieved = updates * (2 * 9 * bpp + 1) / (ms * 1e-3) / 1e9
    line = {"metric": "batched D2Q9 step channels: wall time to residual 1e-6 (all domains)",
            "value": round(ms / 1e3, 3), "unit": "s", "higher_is_better": False, "dtype": COMPUTE[args.storage],
            "storage": args.storage,
            "config": {"workload": f"{args.ndom} backward-facing-step channels L 64..512, H 32..128, step 0.3-0.6 H, "
                                   f"tau 0.55..1.0, inlet u=0.03, seed 2024", "cells": int(cells.sum()),
                       "fluid_cells": int(fluid.sum())},
            "converged": int((steps > 0).sum()), "diverged": int((steps == -2).sum()),
            "steps_min": int(steps[steps > 0].min()) if (steps > 0).any() else None,
            "steps_max": int(steps.max()), "updates": updates, "GLUPS_effective": round(glups, 2),
            "GLUPS_all_cells": round(glups_all, 2),
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": round(achieved / PEAK_HBM_GBS, 4)},
            "tiling": {"tile_x": tx, "blocks": blocks, "one_launch": one, "tuning_s": round(tune_s, 2),
                       "ms_per_step_full": tms},
            "tiles": info["tiles"]}
    print(json.dumps(line), flush=True)


def sphere_channel(n=(512, 128, 128), radius=16.0, centre=(128.0, 64.0, 64.0), u_in=0.04, re=100.0):
    import numpy as np

    nx, ny, nz = n
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    solid = (((x + 0.5 - centre[0]) ** 2 + (y + 0.5 - centre[1]) ** 2 + (z + 0.5 - centre[2]) ** 2)
             < radius * radius).reshape(-1)
    nu = u_in * 2 * radius / re
    tau = 3 * nu + 0.5
    faces = {k: ("wall", None, 1.0) for k in ["-y", "+y", "-z", "+z"]}
    faces["-x"] = ("velocity_neem", (u_in, 0.0, 0.0), 1.0)
    faces["+x"] = ("pressure_neem", None, 1.0)
    return solid, tau, faces


def sphere_solve(storage, max_steps, halo="l2", tol=1e-6):
    """The sphere channel from plug flow (rho = 1, u = 0.04 in the fluid) to
    the relative residual `tol` in one storage form: checks every 500 steps,
    wall time from the first launch to the synchronised last check."""
    import numpy as np

    from lbm_amd import capi

    solid, tau, faces = sphere_channel()
    N = solid.size
    lat = capi.Lattice((512, 128, 128), storage=storage, omega=1 / tau, faces=faces, solid=solid)
    if halo == "lds":
        lat.set_halo(1)
    rho = np.ones(N)
    u = np.zeros((N, 3))
    u[~solid, 0] = 0.04  # plug-flow start
    lat.init(rho, u)
    lat.sync()
    t0 = time.perf_counter()
    steps, res, hist, rc = 0, 1.0, [], 0
    K = 500
    while steps < max_steps:
        lat.step(K - 2)
        s, rc = lat.step_check()
        steps += K
        res = s.res_rel
        hist.append((steps, res))
        if rc != 0 or not (res >= tol):
            break
        if len(hist) % 20 == 0:
            print(f"# sphere {storage} step {steps} res {res:.3e} ({time.perf_counter() - t0:.1f} s)", flush=True)
    lat.sync()
    wall = time.perf_counter() - t0
    _, uu = lat.macros()
    lat.close()
    print(f"# sphere {storage}: {steps} steps, res {res:.3e}, {wall:.2f} s", flush=True)
    return {"steps": steps, "res_rel": res, "wall_s": round(wall, 3), "converged": bool(res < tol),
            "diverged": rc != 0, "ms_per_step": round(wall / steps * 1e3, 4), "u": uu}


def run_sphere(args):
    """Config 3: 512x128x128 D3Q19 channel past a sphere (r = 16 at (128,64,64)),
    NEEM velocity inlet u = 0.04 and NEEM pressure outlet, Re = 100; wall time
    to relative residual 1e-6 for each storage form.  (At this tau = 0.5384
    the Zou-He inlet goes unstable once the sphere's pressure wave reaches
    it, the anti-bounce-back outlet is linearly unstable and the Zou-He
    pressure outlet keeps an undamped period-2 mode.)"""
    import numpy as np

    _, tau, _ = sphere_channel()
    out = {}
    for storage in args.forms.split(","):
        out[storage] = sphere_solve(storage, args.max_steps, args.halo)
    ref = out[args.forms.split(",")[0]]["u"]
    line = {"metric": "D3Q19 sphere channel 512x128x128 Re=100: wall time to residual 1e-6",
            "unit": "s", "higher_is_better": False,
            "config": {"workload": "sphere r=16 at (128,64,64), NEEM velocity inlet u=0.04, NEEM pressure outlet, BB walls",
                       "tau": tau, "gather": args.halo},
            "forms": {}}
    for k, v in out.items():
        e = float(np.linalg.norm(v["u"] - ref) / np.linalg.norm(ref))
        line["forms"][k] = {kk: vv for kk, vv in v.items() if kk != "u"}
        line["forms"][k]["u_rel_l2_vs_first_form"] = e
    line["value"] = line["forms"][args.forms.split(",")[0]]["wall_s"]
    print(json.dumps(line), flush=True)


def run_warmstart(args):
    print(json.dumps(warmstart_eval(args)), flush=True)


def warmstart_eval(args):
    """Config 5: learned warm start on 256x128 D2Q9 obstacle channels
    (1-4 circles/rectangles, Re 20..200 on the channel height, u_in = 0.05).
    Training fields come from the batched FP32 solver (relative residual
    1e-5); the surrogate (bf16 MFMA U-Net, ~2M parameters) is trained on
    them; on unseen channels the plain FP32 start (rho = 1, inlet velocity in
    every fluid cell) and the warm start both run to the relative residual
    1e-6, and their velocities are compared with FP64 solutions converged to
    1e-10."""
    import numpy as np

    from lbm_amd import surrogate as S
    from lbm_amd import warmstart as WS

    def log(msg):
        print(msg, flush=True)

    if os.environ.get("ROCPROF_OUTPUT_PATH"):
        # Under rocprofv3 --kernel-trace this workload's first lbm_batch_run
        # ended in a host SIGSEGV inside the HIP runtime (twice; the same
        # call runs clean without the profiler and the batch workload runs
        # clean under it).  The cause is not known, so the combination is
        # refused; profile the network with --workload train and the batched
        # solver with --workload batch.
        sys.exit("bench.py --workload warmstart does not run under rocprofv3 (see warmstart_eval)")
    t0 = time.perf_counter()
    sur = S.Surrogate(batch=16, height=WS.NY, width=WS.NX, seed=7)
    if args.load_model:
        norm = WS.load_model(sur, args.load_model)
        wall_gen = wall_train = 0.0
        losses = [float("nan")]
    else:
        doms, x = WS.make_set(args.ntrain, seed=11)
        steps, res, wall_gen, _, fields = WS.solve(doms, 1e-5, args.max_steps, log=log)
        ok = steps > 0
        log(f"# training set: {int(ok.sum())}/{len(doms)} channels converged to 1e-5 in {wall_gen:.1f} s "
            f"(steps {int(steps[ok].min())}..{int(steps[ok].max())})")
        keep = np.nonzero(ok)[0]
        t, m, norm = WS.targets([doms[i] for i in keep], [fields[i] for i in keep])
        del fields
        xa, ta, ma = WS.mirror(x[keep], t, m) if args.mirror else (x[keep], t, m)
        losses, wall_train = WS.train(sur, xa, ta, ma, args.epochs, args.lr, args.lr / 50, log=log)
        del xa, ta, ma
        del x, t, m
        if args.save_model:
            WS.save_model(sur, norm, args.save_model)
    tdoms, tx = WS.make_set(args.ntest, seed=12)
    pred, wall_pred = WS.predict_fields(sur, tx, tdoms, norm)
    pred_per_case = wall_pred / len(tdoms)
    sur.close()
    # each unseen channel solved alone (a batch of one domain: graph-captured
    # cycles of 100 steps), as a user would solve one problem
    res = {"cold": [], "warm": [], "two": []}
    for i, d in enumerate(tdoms):
        sc, _, wc, _, fc = WS.solve([d], 1e-6, args.max_steps)
        sw, _, ww, _, fw = WS.solve([d], 1e-6, args.max_steps, fields=[pred[i]])
        s1, s2, w2, f2 = WS.solve_two_phase([d], [pred[i]], args.tol16, 1e-6, args.max_steps)
        res["cold"].append((int(sc[0]), wc, fc[0]))
        res["warm"].append((int(sw[0]), ww + pred_per_case, fw[0]))
        res["two"].append((int(s1[0]) + int(s2[0]) if s1[0] > 0 and s2[0] > 0 else -1, w2 + pred_per_case, f2[0]))
    ref_steps, _, _, _, ref_fields = WS.solve(tdoms, 1e-10, 4 * args.max_steps, storage="f64",
                                              fields=[r[2] for r in res["cold"]])
    ok = np.array([res["cold"][i][0] > 0 and res["warm"][i][0] > 0 and res["two"][i][0] > 0 and ref_steps[i] > 0
                   for i in range(len(tdoms))])
    idx = np.nonzero(ok)[0]

    def stats(key):
        sp = np.array([res["cold"][i][1] / res[key][i][1] for i in idx])
        st = np.array([res["cold"][i][0] / max(1, res[key][i][0]) for i in idx])
        err = [WS.rel_l2(res[key][i][2][1], ref_fields[i][1], tdoms[i]["solid"]) for i in idx]
        return {"speedup_median": round(float(np.median(sp)), 3),
                "speedup_geomean": round(float(np.exp(np.mean(np.log(sp)))), 3),
                "steps_ratio_median": round(float(np.median(st)), 3),
                "wall_s_median": round(float(np.median([res[key][i][1] for i in idx])), 4),
                "steps_median": float(np.median([res[key][i][0] for i in idx])),
                "u_rel_l2_vs_fp64_max": float(max(err)), "u_rel_l2_vs_fp64_median": float(np.median(err))}

    cold = stats("cold")
    warm = stats("warm")
    two = stats("two")
    e_pred = [WS.rel_l2(pred[i][1], ref_fields[i][1], tdoms[i]["solid"]) for i in idx]
    line = {"metric": "time-to-1e-6 residual vs plain FP32 GPU (learned warm start)",
            "value": warm["speedup_median"], "unit": "x", "higher_is_better": True,
            "dtype": "f64 solver arithmetic on f32 storage, bf16 MFMA surrogate", "data": "synthetic seeded obstacle channels",
            "config": {"workload": f"D2Q9 256x128 obstacle channels, Re 20..200, "
                                   + (f"network loaded from {os.path.basename(args.load_model)}" if args.load_model else
                                      f"{args.ntrain} training ({'+ mirrored, ' if args.mirror else ''}"
                                      f"{args.epochs} epochs)")
                                   + f" / {args.ntest} test channels, each test channel solved alone"},
            "cold_f32": cold, "warm_f32": warm, "warm_f16_then_f32": dict(two, tol_f16=args.tol16),
            "converged_cases": int(ok.sum()), "prediction_u_rel_l2_median": float(np.median(e_pred)),
            "predict_s_per_case": round(pred_per_case, 4),
            "training": {"fields_s": round(wall_gen, 1), "train_s": round(wall_train, 1),
                         "loss_first": losses[0], "loss_last": losses[-1]},
            "total_s": round(time.perf_counter() - t0, 1)}
    return line


def run_poisson(args):
    """The Poisson step (include/lbm_poisson.h) on an N x N D2Q9 window
    (default 8192^2, FP64), Dirichlet faces, Debye-Hueckel source: W
    untimed launches, then K launches timed with host clocks around a
    synchronising check launch; algorithmic bytes per cell update are one
    read and one write of the nine FP64 populations plus the potential's
    write (152 B)."""
    import numpy as np

    from lbm_amd import poisson as P

    n = args.n if args.n != 256 else 8192
    tau = 1.0
    dt = 1.0 / (n - 1)
    src = dt * (1.0 / 3.0) * (0.5 - tau) * dt * 27.79 ** 2
    d = P.Poisson(9, n, n, 1.0 / tau, src)
    ones = np.ones(n)
    d.set_faces({f: ("dirichlet", ones, k) for k, f in enumerate(["+x", "+y", "-x", "-y"])})
    phi0 = np.zeros((n, n))
    phi0[0, :] = phi0[-1, :] = phi0[:, 0] = phi0[:, -1] = 1.0
    d.init(phi0)
    del phi0
    d.step(args.warmup)
    d.step_check()
    t0 = time.perf_counter()
    d.step(args.steps - 1)
    res, _ = d.step_check()
    sec = time.perf_counter() - t0
    d.close()
    ms = sec / args.steps * 1e3
    alg = (2 * 9 * 8 + 8) * n * n
    print(json.dumps({"metric": "Poisson D2Q9 step, FP64", "value": round(n * n / (ms * 1e-3) / 1e6, 1), "unit": "MLUPS",
                      "ms_per_step": round(ms, 5), "steps": args.steps, "warmup": args.warmup,
                      "higher_is_better": True, "dtype": "f64", "data": "synthetic (Dirichlet box, phi = 1 on the faces)",
                      "config": {"workload": f"D2Q9 Poisson {n}x{n}, tau 1, Debye-Hueckel kappa 27.79"},
                      "roofline": {"bound": "hbm", "achieved": round(alg / (ms * 1e-3) / 1e9, 1), "peak": PEAK_HBM_GBS,
                                   "unit": "GB/s", "frac": round(alg / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                                   "alg_bytes_per_launch": alg},
                      "residual_last": res}), flush=True)


def run_train(args):
    """Training steps of the warm-start network (16 samples of 256x128,
    forward + backward + Adam) on seeded random fields: W untimed steps, then
    K timed steps, wall time around the synchronous training call."""
    import numpy as np

    from lbm_amd import surrogate as S
    from lbm_amd import warmstart as WS

    b = 16
    sur = S.Surrogate(batch=b, height=WS.NY, width=WS.NX, seed=7)
    rng = np.random.default_rng(1)
    n = 4 * b
    x = rng.standard_normal((n, WS.NY, WS.NX, S.IN_CHANNELS)).astype(np.float32)
    t = rng.standard_normal((n, WS.NY, WS.NX, 3)).astype(np.float32)
    m = np.ones((n, WS.NY, WS.NX), np.uint8)
    sur.set_dataset(x, t, m)
    idx = lambda k: np.resize(np.arange(n), k * b)
    if args.warmup:
        sur.train(idx(args.warmup), args.lr)
    t0 = time.perf_counter()
    losses = sur.train(idx(args.steps), args.lr)
    sec = time.perf_counter() - t0
    sur.close()
    print(json.dumps({"metric": "warm-start network training step", "value": round(sec / args.steps * 1e3, 3),
                      "unit": "ms/step", "higher_is_better": False, "steps": args.steps, "warmup": args.warmup,
                      "samples_per_s": round(b * args.steps / sec, 1), "dtype": "bf16 MFMA, fp32 master weights",
                      "data": "synthetic (seeded random fields)",
                      "config": {"workload": f"U-Net, batch {b} x {WS.NY}x{WS.NX}, {S.IN_CHANNELS} input channels"},
                      "loss_last": float(losses[-1])}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "bf16", "f64"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--settle", type=int, default=1,
                    help="after the warm-up, run untimed groups of 10 steps until the step time is stable")
    ap.add_argument("--workload", default="tgv", choices=["tgv", "batch", "sphere", "warmstart", "train", "poisson"],
                    help="tgv: the headline 256^3 box (default); batch / sphere / warmstart: time-to-residual "
                         "workloads; train: training-step throughput of the warm-start network; poisson: the "
                         "Poisson step on an 8192^2 D2Q9 window")
    ap.add_argument("--ntrain", type=int, default=1024)
    ap.add_argument("--ntest", type=int, default=32)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--tol16", type=float, default=1e-3,
                    help="warm start: residual at which FP16 storage hands over to FP32 (its rounding floor is ~5e-4)")
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--mirror", type=int, default=1, help="train on the mirrored channels too")
    ap.add_argument("--save-model", default=None)
    ap.add_argument("--load-model", default=None)
    ap.add_argument("--ndom", type=int, default=2048)
    ap.add_argument("--max-steps", type=int, default=400000)
    ap.add_argument("--forms", default="f32,f16")
    ap.add_argument("--halo", default="l2", choices=["l2", "lds"],
                    help="tgv / sphere, FP32: neighbour gather straight from L2 (default) or the LDS-staged "
                         "x-halo variant of the step kernel")
    ap.add_argument("--ttr", type=int, default=1,
                    help="default workload: add the time-to-residual object (sphere channel in FP32, warm start "
                         "against plain start on unseen channels with a network trained in the run)")
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
    if args.workload == "batch":
        return run_batch(args)
    if args.workload == "sphere":
        return run_sphere(args)
    if args.workload == "train":
        return run_train(args)
    if args.workload == "warmstart":
        return run_warmstart(args)
    if args.workload == "poisson":
        return run_poisson(args)

    import numpy as np

    from lbm_amd import capi
    from lbm_amd import workloads as W

    n = args.n
    shape = (n, n, n)
    rho0, u0 = W.taylor_green_noise(shape, 0.05, 1e-3, seed=1234)
    lat = capi.Lattice(shape, storage=args.storage, omega=1 / 0.6)
    if args.halo == "lds":
        lat.set_halo(1)
    lat.init(rho0, u0)
    del rho0, u0
    lat.step(args.warmup)
    lat.sync()
    settle_steps, settle_ms = settle(lat) if args.settle else (0, [])
    groups = repetitions(args.steps)
    t0 = time.perf_counter()
    group_ms = lat.time_step_groups(groups)  # HIP events on the kernel's stream between repetitions
    lat.sync()
    wall = time.perf_counter() - t0
    per_step = group_ms / np.asarray(groups)
    cells = n ** 3
    ms_per_step = float(np.median(per_step))
    mlups = cells / (ms_per_step * 1e-3) / 1e6
    bpp = capi.BYTES[args.storage]
    alg_bytes = 2 * 19 * bpp * cells
    achieved = alg_bytes / (ms_per_step * 1e-3) / 1e9
    s = lat.reduce()
    lat.close()
    # the committed counter run covers the default (L2 gather) kernel only
    traffic = load_traffic(f"traffic_d3q19_{args.storage}_{n}.json") if args.halo == "l2" else None
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
        "dtype": COMPUTE[args.storage],
        "storage": args.storage,
        "data": "synthetic (Taylor-Green vortex u0=0.05 + Gaussian noise sigma=1e-3, seed 1234)",
        "config": {"workload": f"D3Q19 BGK periodic box {n}^3, tau=0.6, {args.storage} populations (shifted f-w), "
                               f"fused pull collide+stream", "cells": cells, "parallelism": "single GPU",
                   "gather": args.halo},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "alg_bytes_per_launch": alg_bytes},
        "timing": {"repetitions": len(groups), "steps_per_repetition": groups,
                   "ms_per_step_median": round(ms_per_step, 5), "ms_per_step_min": round(float(per_step.min()), 5),
                   "ms_per_step_max": round(float(per_step.max()), 5),
                   "ms_per_step_mean": round(float(group_ms.sum()) / args.steps, 5)},
        "settle": {"steps": settle_steps, "ms_per_step_of_groups": [round(x, 4) for x in settle_ms],
                   "rule": "untimed groups of 10 steps after the warm-up until three in a row agree within 2 % "
                           "(at most 300 steps)"},
        "wall_s_timed": round(wall, 4),
        "mass_after": s.mass,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    if args.ttr and not os.environ.get("ROCPROF_OUTPUT_PATH"):
        line["time_to_residual"] = time_to_residual()
    print(json.dumps(line), flush=True)


def time_to_residual():
    """The metric's second half, bounded to about two minutes of GPU time:
    the 512x128x128 sphere channel to the relative residual 1e-6 in FP32
    (capped at 80,000 steps), and the learned warm start against the plain
    FP32 start on 8 unseen 256x128 obstacle channels, twice: with a network
    trained in this run on 1024 channels (+ mirrored, 30 epochs) by the
    project's own kernels, and with the network that ships with the package
    (the same kernels, trained offline on 8192 channels + mirrored, 240
    epochs).  Skipped under rocprofv3 (the warm start's batched solves do not
    run under it)."""
    t0 = time.perf_counter()
    sph = sphere_solve("f32", 80000)
    sph.pop("u")
    ws = warmstart_eval(argparse.Namespace(load_model=None, ntrain=1024, ntest=8, epochs=30, lr=5e-4, mirror=1,
                                           max_steps=400000, tol16=1e-3, save_model=None))
    keep = ("cold_f32", "warm_f32", "converged_cases", "prediction_u_rel_l2_median", "predict_s_per_case", "training")
    from lbm_amd import warmstart as WS
    wp = warmstart_eval(argparse.Namespace(load_model=WS.SHIPPED_MODEL, ntrain=8192, ntest=8, epochs=240, lr=5e-4,
                                           mirror=1, max_steps=400000, tol16=1e-3, save_model=None))
    return {"sphere_512x128x128_f32": sph,
            "warm_start": {"speedup_median_vs_plain_f32": ws["value"], "config": ws["config"]["workload"],
                           **{k: ws[k] for k in keep}},
            "warm_start_shipped_network": {"speedup_median_vs_plain_f32": wp["value"],
                                           "config": wp["config"]["workload"],
                                           **{k: wp[k] for k in keep if k != "training"}},
            "wall_s": round(time.perf_counter() - t0, 1)}


if __name__ == "__main__":
    main()

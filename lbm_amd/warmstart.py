"""Learned warm start for steady 2D obstacle channels (256 x 128 D2Q9).

Pipeline: seeded obstacle channels -> steady fields from the batched solver
(include/lbm_batch.h) -> features/targets -> surrogate training on the GPU
(include/lbm_surrogate.h) -> prediction used as the initial field of new
channels, compared with the plain FP32 start (rho = 1, inlet velocity in
every fluid cell) on wall time and steps to the relative residual 1e-6, and
on the velocity error against FP64 solutions converged to 1e-10.
"""
import os
import time

import numpy as np

from . import capi
from . import surrogate as S
from . import workloads as W

U_IN = 0.05
NX, NY = 256, 128


def make_set(n, seed):
    rng = np.random.default_rng(seed)
    doms, feats = [], []
    for _ in range(n):
        dom, re, _ = W.obstacle_channel(rng, NX, NY, U_IN)
        solid = dom["solid"].reshape(NY, NX).astype(bool)
        doms.append(dom)
        feats.append(S.features(solid, re))
    return doms, np.stack(feats)


def _init(b, doms, fields=None):
    if fields is None:
        rho, u = zip(*[W.plug_start(d, U_IN) for d in doms])
    else:
        rho, u = zip(*fields)
    b.init(np.concatenate(rho), np.concatenate(u))


def solve(doms, tol, max_steps, storage="f32", fields=None, interval=100, tiling=None, log=None, chunk=20000):
    """Run a batch to tol; returns (steps[n], res[n], wall seconds, device
    seconds, [(rho, u)]).  With log, the run goes in pieces of `chunk` steps
    and each piece reports how many domains have retired."""
    b = capi.Batch(doms, storage=storage)
    if tiling:
        b.set_tiling(*tiling)
    _init(b, doms, fields)
    t0 = time.perf_counter()
    if log is None:
        steps, res, _, ms = b.run(interval, tol, max_steps)
    else:
        done, ms = 0, 0.0
        piece = max(interval, chunk - chunk % interval)
        while done < max_steps:
            n = min(piece, max_steps - done)
            steps, res, _, dms = b.run(interval, tol, n)
            ms += dms
            done += n
            retired = int((steps != -1).sum())
            log(f"# {done} steps: {retired}/{len(doms)} domains retired, {time.perf_counter() - t0:.1f} s")
            if retired == len(doms):
                break
    wall = time.perf_counter() - t0
    out = [b.macros(k) for k in range(len(doms))]
    b.close()
    return steps, res, wall, ms / 1e3, out


def targets(doms, fields):
    """Encoded targets [n][NY][NX][3], fluid masks, and the per-channel
    (mean, std) over fluid cells used to standardise them."""
    t = np.stack([S.encode_targets(r, u, U_IN).reshape(NY, NX, 3) for r, u in fields])
    m = np.stack([(d["solid"].reshape(NY, NX) == 0) for d in doms]).astype(np.uint8)
    fl = m.astype(bool)
    mu = np.array([t[..., c][fl].mean() for c in range(3)], np.float32)
    mu[1] = 0.0  # u_y: zero mean by the mirror symmetry of the channel family
    sd = np.array([t[..., c][fl].std() for c in range(3)], np.float32)
    return ((t - mu) / sd).astype(np.float32), m, (mu, sd)


def mirror(x, t, m):
    """Append the channels mirrored about the centre line (y -> H-1-y): the
    flow of a mirrored geometry is the mirrored flow with u_y negated.  The
    y feature and the u_y target change sign; both targets are standardised
    with a mean that is zero for u_y by symmetry, so negation keeps them
    standardised."""
    xf = x[:, ::-1].copy()
    xf[..., 1] = -xf[..., 1]
    tf = t[:, ::-1].copy()
    tf[..., 1] = -tf[..., 1]
    return (np.concatenate([x, xf]), np.concatenate([t, tf]), np.concatenate([m, m[:, ::-1]]))


def train(sur, x, t, m, epochs, lr0, lr1, seed=0, log=None):
    """Adam with a cosine learning-rate schedule from lr0 to lr1."""
    sur.set_dataset(x, t, m)
    n = x.shape[0]
    rng = np.random.default_rng(seed)
    steps_per_epoch = n // sur.batch
    total = epochs * steps_per_epoch
    losses = []
    k = 0
    t0 = time.perf_counter()
    for e in range(epochs):
        perm = rng.permutation(n)[:steps_per_epoch * sur.batch]
        # one call per epoch with a fixed learning rate for that epoch
        lr = lr1 + 0.5 * (lr0 - lr1) * (1 + np.cos(np.pi * k / max(1, total)))
        l = sur.train(perm, lr)
        k += steps_per_epoch
        losses.append(float(np.mean(l)))
        if log:
            log(f"# epoch {e + 1}/{epochs} lr {lr:.2e} loss {losses[-1]:.3e} ({time.perf_counter() - t0:.1f} s)")
    return losses, time.perf_counter() - t0


def predict_fields(sur, x, doms, norm):
    mu, sd = norm
    out = []
    t0 = time.perf_counter()
    for i in range(0, x.shape[0], sur.batch):
        y = sur.forward(x[i:i + sur.batch]) * sd + mu
        for k in range(y.shape[0]):
            d = doms[i + k]
            out.append(S.decode_prediction(y[k], d["solid"], U_IN))
    return out, time.perf_counter() - t0


def rel_l2(u, ref, solid):
    f = ~np.asarray(solid, bool).reshape(-1)
    return float(np.linalg.norm(u[f] - ref[f]) / np.linalg.norm(ref[f]))


def solve_two_phase(doms, fields, tol1, tol, max_steps, storage1="f16", interval=100):
    """Reduced storage first (to tol1, above its rounding floor), then FP32 from
    the phase-1 fields to tol.  Returns (steps1, steps2, wall seconds, fields)."""
    s1, _, w1, _, f1 = solve(doms, tol1, max_steps, storage=storage1, fields=fields, interval=interval)
    s2, _, w2, _, f2 = solve(doms, tol, max_steps, storage="f32", fields=f1, interval=interval)
    return s1, s2, w1 + w2, f2


# A network trained by this project's kernels on 8192 channels (seed 11) and
# their mirror images for 240 epochs, stored as its bf16 working weights:
#   bench.py --workload warmstart --ntrain 8192 --epochs 240 --ntest 16 --save-model m.npz
# then save_model_bf16(params, mu, sd) of m.npz.  On its 16 test channels
# (seed 12): median 2.10x less wall time to 1e-6 than the plain start
# (2.12x fewer steps), prediction error 2.5 % (median u rel-L2).
SHIPPED_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models", "warmstart_obstacle_8192.npz")


def save_model(sur, norm, path):
    mu, sd = norm
    np.savez(path, params=sur.params(), mu=mu, sd=sd)


def bf16_bits(p):
    """FP32 -> bf16 bit patterns, rounded to nearest even as the network's
    working copies are (finite values)."""
    u = np.ascontiguousarray(p, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def save_model_bf16(params, mu, sd, path):
    """A trained network as its bf16 working weights (the forward pass uses
    only those, so predictions are unchanged) and the target statistics."""
    np.savez_compressed(path, params_bf16=bf16_bits(params), mu=np.asarray(mu, np.float32),
                        sd=np.asarray(sd, np.float32))


def load_model(sur, path):
    z = np.load(path)
    if "params_bf16" in z:
        p = (z["params_bf16"].astype(np.uint32) << 16).view(np.float32)
    else:
        p = z["params"]
    sur.set_params(p)
    return z["mu"], z["sd"]

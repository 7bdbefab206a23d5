"""Synthetic inputs of the benchmark workloads (all seeded, reproducible).

backward_step_batch: uneven backward-facing-step channels, D2Q9: length
uniform in 64..512, height uniform in 32..128, step height 0.3..0.6 of the
channel height over the first quarter of the length, tau uniform in
0.55..1.0, Zou-He inlet u = 0.03 on the open part of x = 0, anti-bounce-back
pressure outlet rho = 1 at the right end, bounce-back top/bottom walls and step.

obstacle_channel: 256 x 128 D2Q9 channel with 1..4 seeded circles or
rectangles (radius / half-size 4..20) at Re 20..200 (Re on the channel height
with inlet u = 0.05), NEEM (non-equilibrium extrapolation) velocity inlet
and pressure outlet, bounce-back walls.

taylor_green_noise: the periodic-box start field, a Taylor-Green vortex at
cell centres plus seeded Gaussian noise (splitmix64 + Box-Muller).
"""
import numpy as np

_MASK64 = (1 << 64) - 1


def _splitmix64(first, count, seed):
    """Outputs first+1 .. first+count of splitmix64 (Steele, Lea & Flood,
    "Fast splittable pseudorandom number generators", OOPSLA 2014) started
    from seed: state_k = seed + k * golden gamma, then the variant-13 mix."""
    k = np.arange(first + 1, first + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & _MASK64) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _box_muller(first_pair, npairs, seed):
    """Standard normal pairs: outputs 2p+1 and 2p+2 of the generator give
    u1 in (0, 1] and u2 in [0, 1) from their top 53 bits; the pair is
    sqrt(-2 ln u1) (cos 2 pi u2, sin 2 pi u2)."""
    r = _splitmix64(2 * first_pair, 2 * npairs, seed)
    scale = 1.0 / 9007199254740992.0  # 2^-53
    u1 = ((r[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * scale
    u2 = (r[1::2] >> np.uint64(11)).astype(np.float64) * scale
    rad = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * npairs)
    out[0::2] = rad * np.cos(2 * np.pi * u2)
    out[1::2] = rad * np.sin(2 * np.pi * u2)
    return out


def taylor_green_noise(shape, u0, sigma, seed, chunk=1 << 21):
    """Taylor-Green vortex at cell centres x = (i + 1/2), k = 2 pi / n per
    axis: u = u0 (sin kx cos ky cos kz, -cos kx sin ky cos kz, 0), plus
    Gaussian noise of standard deviation sigma on every component, drawn in
    cell-major order (u_x, u_y, u_z of cell 0, then cell 1, ...).  Built in
    chunks of cells to bound the temporaries.  Returns (rho[n], u[n, 3])."""
    nx, ny, nz = shape
    n = nx * ny * nz
    u = np.empty((n, 3))
    for c0 in range(0, n, chunk):
        c1 = min(n, c0 + chunk)
        i = np.arange(c0, c1)
        X = (i % nx + 0.5) * 2 * np.pi / nx
        Y = ((i // nx) % ny + 0.5) * 2 * np.pi / ny
        Z = (i // (nx * ny) + 0.5) * 2 * np.pi / nz
        blk = u[c0:c1]
        blk[:, 0] = u0 * np.sin(X) * np.cos(Y) * np.cos(Z)
        blk[:, 1] = -u0 * np.cos(X) * np.sin(Y) * np.cos(Z)
        blk[:, 2] = 0.0
        if sigma > 0:
            # values 3*c0 .. 3*c1-1 of the normal stream; pairs start at even values
            v0, v1 = 3 * c0, 3 * c1
            p0, p1 = v0 // 2, (v1 + 1) // 2
            g = _box_muller(p0, p1 - p0, seed)[v0 - 2 * p0:v1 - 2 * p0]
            blk += sigma * g.reshape(-1, 3)
    return np.ones(n), u

WALLS2 = {"-y": ("wall", None, 1.0), "+y": ("wall", None, 1.0)}


def channel_faces(u_in, rho_out=1.0):
    """Zou-He velocity inlet, anti-bounce-back pressure outlet (the reference's
    'pressure' boundary), bounce-back walls.  A Zou-He pressure outlet is not
    used here: it keeps an undamped period-2 oscillation of the outlet
    velocity, so the one-step residual stalls near 1e-3."""
    f = dict(WALLS2)
    f["-x"] = ("velocity_zh", (u_in, 0.0), 1.0)
    f["+x"] = ("pressure_abb", None, rho_out)
    return f


def backward_step_batch(n=2048, seed=2024, u_in=0.03):
    rng = np.random.default_rng(seed)
    doms = []
    for _ in range(n):
        L = int(rng.integers(64, 513))
        H = int(rng.integers(32, 129))
        hs = int(round(rng.uniform(0.3, 0.6) * H))
        ls = L // 4
        tau = float(rng.uniform(0.55, 1.0))
        solid = np.zeros((H, L), dtype=np.uint8)
        solid[:hs, :ls] = 1
        doms.append({"shape": (L, H), "omega": 1.0 / tau, "solid": solid.reshape(-1), "faces": channel_faces(u_in),
                     "tau": tau, "step": (ls, hs)})
    return doms


def cold_start(dom, u_in):
    """rho = 1, u = 0, inlet cells at the prescribed velocity (the reference's
    initial condition with the boundary initCnd values)."""
    nx, ny = dom["shape"][:2]
    rho = np.ones(nx * ny)
    u = np.zeros((nx * ny, 2))
    solid = dom.get("solid")
    inlet = (np.arange(nx * ny) % nx) == 0
    if solid is not None:
        inlet &= np.asarray(solid).reshape(-1) == 0
    u[inlet, 0] = u_in
    return rho, u


def plug_start(dom, u_in):
    """rho = 1 and the inlet velocity in every fluid cell: the usual start of a
    channel computation; it avoids the pressure shock of an impulsive inlet,
    which makes BGK blow up at relaxation times near 1/2."""
    nx, ny = dom["shape"][:2]
    nz = dom["shape"][2] if len(dom["shape"]) > 2 else 1
    n = nx * ny * nz
    dim = 2 if nz == 1 else 3
    rho = np.ones(n)
    u = np.zeros((n, dim))
    u[:, 0] = u_in
    solid = dom.get("solid")
    if solid is not None:
        u[np.asarray(solid).reshape(-1) != 0] = 0.0
    return rho, u


def neem_channel_faces(u_in, rho_out=1.0):
    """NEEM velocity inlet and NEEM pressure outlet with bounce-back walls: the
    open faces that stay stable and converge down to tau ~ 0.54 (the Zou-He
    inlet and the anti-bounce-back outlet do not)."""
    f = dict(WALLS2)
    f["-x"] = ("velocity_neem", (u_in, 0.0), 1.0)
    f["+x"] = ("pressure_neem", None, rho_out)
    return f


def obstacle_channel(rng, nx=256, ny=128, u_in=0.05):
    """One seeded obstacle channel (Re on the channel height, NEEM inlet and
    outlet, bounce-back walls); returns (domain dict, Re, obstacle list)."""
    re = float(rng.uniform(20.0, 200.0))
    nu = u_in * ny / re
    tau = 3.0 * nu + 0.5
    yy, xx = np.mgrid[0:ny, 0:nx]
    solid = np.zeros((ny, nx), dtype=bool)
    obs = []
    for _ in range(int(rng.integers(1, 5))):
        r = float(rng.uniform(4.0, 20.0))
        cx = float(rng.uniform(40.0, nx - 60.0))
        cy = float(rng.uniform(r + 4.0, ny - r - 4.0))
        if rng.random() < 0.5:
            m = (xx + 0.5 - cx) ** 2 + (yy + 0.5 - cy) ** 2 < r * r
            obs.append(("circle", cx, cy, r))
        else:
            ry = float(rng.uniform(4.0, 20.0))
            ry = min(ry, cy - 3.0, ny - cy - 3.0)
            m = (np.abs(xx + 0.5 - cx) < r) & (np.abs(yy + 0.5 - cy) < ry)
            obs.append(("rect", cx, cy, r, ry))
        solid |= m
    dom = {"shape": (nx, ny), "omega": 1.0 / tau, "solid": solid.astype(np.uint8).reshape(-1),
           "faces": neem_channel_faces(u_in), "tau": tau, "re": re}
    return dom, re, obs

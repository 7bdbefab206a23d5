"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

Bars: FP32 storage <= 1e-5 relative L2 on u and rho; FP64 storage <= 1e-11
(same arithmetic, different summation/FMA order); reduced 16-bit storage
reports its error against the FP64 oracle and must stay below the stated
bound.  Sizes are small enough for the oracle to finish in seconds.
"""
import numpy as np
import pytest

from lbm_amd import capi
from oracle import oracle as O

pytestmark = pytest.mark.gpu

WALLS2 = {"-y": ("wall", None, 1.0), "+y": ("wall", None, 1.0)}


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def run_pair(shape, steps, storage, omega, faces=None, solid=None, force=(0.0, 0.0, 0.0), init=None, seed=0):
    shape3 = tuple(shape) + ((1,) if len(shape) == 2 else ())
    N = int(np.prod(shape3))
    dim = 2 if len(shape) == 2 else 3
    if init is None:
        rng = np.random.default_rng(seed)
        rho0 = 1 + 0.01 * rng.standard_normal(N)
        u0 = 0.02 * rng.standard_normal((N, dim))
    else:
        rho0, u0 = init
    print(f"case shape={shape} steps={steps} storage={storage} faces={list((faces or {}).keys())}", flush=True)
    lat = capi.Lattice(shape, storage=storage, omega=omega, force=force, faces=faces, solid=solid)
    lat.init(rho0, u0)
    lat.step(steps)
    rho, u = lat.macros()
    ref = O.Problem(shape, faces=faces, solid=solid, omega=omega, force=force)
    ref.init(rho0, u0)
    ref.step(steps)
    r_rho, r_u = ref.macros()
    fl = ~ref.solid
    lat.close()
    return rel(u[fl], r_u[fl]), rel(rho[fl], r_rho[fl]), (rho, u, r_rho, r_u)


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f64", 1e-11)])
def test_d3q19_periodic_tgv(storage, tol):
    shape = (32, 32, 32)
    init = O.taylor_green_noise(shape, 0.05, 1e-3, seed=1234)
    eu, er, _ = run_pair(shape, 100, storage, 1 / 0.6, init=init)
    print(f"d3q19 periodic {storage}: rel-L2 u={eu:.3e} rho={er:.3e}")
    assert eu < tol and er < tol


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f64", 1e-11)])
def test_d3q19_non_multiple_of_64(storage, tol):
    eu, er, _ = run_pair((70, 9, 5), 30, storage, 1 / 0.8)
    assert eu < tol and er < tol


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f64", 1e-11)])
def test_d2q9_force_channel(storage, tol):
    eu, er, _ = run_pair((64, 32), 200, storage, 1 / 0.8, faces=WALLS2, force=(1e-5, 0.0, 0.0))
    assert eu < tol and er < tol


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f64", 1e-11)])
def test_d2q9_couette_moving_wall(storage, tol):
    faces = {"-y": ("wall", None, 1.0), "+y": ("wall", (0.02, 0.0), 1.0)}
    eu, er, _ = run_pair((16, 24), 300, storage, 1.0, faces=faces)
    assert eu < tol and er < tol


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f64", 1e-11)])
def test_d2q9_pressure_abb_channel(storage, tol):
    faces = dict(WALLS2)
    faces["-x"] = ("pressure_abb", None, 1.03)
    faces["+x"] = ("pressure_abb", None, 1.0)
    N = 40 * 16
    eu, er, _ = run_pair((40, 16), 300, storage, 1 / 0.7, faces=faces, init=(np.ones(N), np.zeros((N, 2))))
    assert eu < tol and er < tol


def sphere_solid(shape, c, r):
    nx, ny, nz = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return (((x + 0.5 - c[0]) ** 2 + (y + 0.5 - c[1]) ** 2 + (z + 0.5 - c[2]) ** 2) < r * r).reshape(-1)


def channel_faces3(uin=0.04, pout=1.0):
    f = {k: ("wall", None, 1.0) for k in ["-y", "+y", "-z", "+z"]}
    f["-x"] = ("velocity_zh", (uin, 0.0, 0.0), 1.0)
    f["+x"] = ("pressure_zh", None, pout)
    return f


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f64", 1e-11)])
def test_d3q19_sphere_channel(storage, tol):
    shape = (48, 20, 20)
    solid = sphere_solid(shape, (16, 10, 10), 4.0)
    N = int(np.prod(shape))
    init = (np.ones(N), np.zeros((N, 3)))
    eu, er, _ = run_pair(shape, 150, storage, 1 / 0.56, faces=channel_faces3(), solid=solid, init=init)
    print(f"sphere channel {storage}: rel-L2 u={eu:.3e} rho={er:.3e}")
    assert eu < tol and er < tol


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f64", 1e-11)])
def test_d3q19_moving_lid_and_abb(storage, tol):
    f = {k: ("wall", None, 1.0) for k in ["-x", "+x", "-y"]}
    f["+y"] = ("wall", (0.03, 0.0, 0.01), 1.0)
    f["-z"] = ("pressure_abb", None, 1.0005)
    f["+z"] = ("pressure_abb", None, 1.0)
    eu, er, _ = run_pair((12, 10, 14), 80, storage, 1 / 0.65, faces=f)
    assert eu < tol and er < tol


def test_populations_f64_bitwise_close():
    """Post-collision state after n launches equals the oracle's f_post of step n+1."""
    shape = (20, 12)
    rng = np.random.default_rng(5)
    N = 240
    rho0 = 1 + 0.01 * rng.standard_normal(N)
    u0 = 0.02 * rng.standard_normal((N, 2))
    faces = {"-y": ("wall", (0.01, 0.0), 1.0), "+y": ("wall", None, 1.0)}
    lat = capi.Lattice(shape, storage="f64", omega=1.3, faces=faces, force=(1e-5, 2e-6, 0))
    lat.init(rho0, u0)
    lat.step(7)
    f = lat.populations()
    ref = O.Problem(shape, faces=faces, omega=1.3, force=(1e-5, 2e-6, 0))
    ref.init(rho0, u0)
    ref.step(8)
    assert np.abs(f - ref.post_collision()).max() < 1e-14
    lat.close()


def test_step_check_matches_oracle_residual_f64():
    shape = (24, 16)
    rng = np.random.default_rng(8)
    N = 24 * 16
    rho0 = 1 + 0.01 * rng.standard_normal(N)
    u0 = 0.02 * rng.standard_normal((N, 2))
    lat = capi.Lattice(shape, storage="f64", omega=1.1, faces=WALLS2)
    lat.init(rho0, u0)
    lat.step(5)
    s, rc = lat.step_check()  # steps 6 and 7: change fold^7 - fold^6
    assert rc == 0
    ref = O.Problem(shape, faces=WALLS2, omega=1.1)
    ref.init(rho0, u0)
    ref.step(6)
    r6, u6 = ref.macros()
    ref.step(1)
    r7, u7 = ref.macros()
    v6 = np.concatenate([u6, r6[:, None]], 1).reshape(-1)
    v7 = np.concatenate([u7, r7[:, None]], 1).reshape(-1)
    res = O.residual(2, v7, v6)
    assert np.allclose(s.abs_diff[:3], res[:3], rtol=1e-9)
    assert abs(s.diff_sq - res[3]) < 1e-9 * res[3]
    assert abs(s.norm_sq - res[4]) < 1e-9 * res[4]
    assert abs(s.mass - r7.sum()) < 1e-10
    assert np.allclose(s.momentum[:2], (r7[:, None] * u7).sum(0), atol=1e-12)
    lat.close()


def test_step_check_deterministic():
    shape = (64, 48, 40)
    init = O.taylor_green_noise(shape, 0.05, 1e-3, seed=3)
    out = []
    for _ in range(2):
        lat = capi.Lattice(shape, storage="f32", omega=1 / 0.6)
        lat.init(*init)
        lat.step(3)
        s, rc = lat.step_check()
        out.append(s.as_dict())
        lat.close()
    assert out[0] == out[1]


def test_divergence_reported():
    lat = capi.Lattice((16, 8, 8), storage="f32", omega=1.0)
    N = 16 * 8 * 8
    lat.init(np.ones(N), np.zeros((N, 3)))
    f = lat.populations()
    f[3, 17] = np.nan
    lat.set_populations(f)
    s, rc = lat.step_check()
    assert rc == capi.LBM_DIVERGED and s.nonfinite >= 1
    lat.close()


def test_invalid_arguments_rejected():
    with pytest.raises(capi.LBMError):
        capi.Lattice((2, 8, 8), storage="f32", omega=1.0)
    with pytest.raises(capi.LBMError):
        capi.Lattice((8, 8, 8), storage="f32", omega=2.5)
    lat = capi.Lattice((8, 8, 8), storage="f32", omega=1.0)
    with pytest.raises(capi.LBMError):
        lat.set_faces({"-x": ("wall", None, 1.0)})  # periodic partner missing
    with pytest.raises(capi.LBMError):
        lat.set_launch(96, 2, 0)
    lat.close()


@pytest.mark.parametrize("storage,bound", [("f16", 2e-3), ("bf16", 3e-2)])
def test_reduced_storage_error_stated(storage, bound):
    shape = (32, 32, 32)
    init = O.taylor_green_noise(shape, 0.05, 1e-3, seed=1234)
    eu, er, _ = run_pair(shape, 100, storage, 1 / 0.6, init=init)
    print(f"d3q19 shifted {storage}: rel-L2 vs FP64 oracle u={eu:.3e} rho={er:.3e}")
    assert eu < bound and er < bound


@pytest.mark.parametrize("launch", [(64, 4, 0), (128, 2, 1), (256, 1, 1)])
def test_launch_shapes_identical(launch):
    shape = (128, 16, 8)
    init = O.taylor_green_noise(shape, 0.05, 1e-3, seed=2)
    res = []
    for lz in [(256, 1, 0), launch]:
        lat = capi.Lattice(shape, storage="f32", omega=1 / 0.6)
        lat.set_launch(*lz)
        lat.init(*init)
        lat.step(10)
        res.append(lat.populations())
        lat.close()
    assert np.array_equal(res[0], res[1])


@pytest.mark.parametrize("storage", ["f32", "f64", "f16"])
def test_index_widths_identical(storage):
    """32-bit per-plane offsets (the default) and 64-bit offsets run the same
    arithmetic: identical populations on a sphere channel (every boundary
    branch, solid cells and flag loads), plain and checked steps."""
    shape = (40, 18, 16)
    solid = sphere_solid(shape, (14, 9, 8), 3.5)
    N = int(np.prod(shape))
    rng = np.random.default_rng(11)
    rho0, u0 = 1 + 0.01 * rng.standard_normal(N), 0.02 * rng.standard_normal((N, 3))
    out = []
    for bits in (32, 64):
        print(f"index width {bits} {storage}", flush=True)
        lat = capi.Lattice(shape, storage=storage, omega=1 / 0.6, faces=channel_faces3(), solid=solid)
        lat.set_index_width(bits)
        lat.init(rho0, u0)
        lat.step(7)
        s, rc = lat.step_check()
        lat.step(4)
        out.append((lat.populations(), s.res_rel))
        lat.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


def test_index_width_rejects_bad_values():
    lat = capi.Lattice((8, 8, 8), storage="f32", omega=1.0)
    with pytest.raises(capi.LBMError):
        lat.set_index_width(16)
    lat.close()


def cylinder_solid2(shape, c, r):
    nx, ny = shape
    yy, xx = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    return (((xx + 0.5 - c[0]) ** 2 + (yy + 0.5 - c[1]) ** 2) < r * r).reshape(-1)


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f64", 1e-11)])
def test_d2q9_neem_outlet_cylinder(storage, tol):
    shape = (60, 22)
    solid = cylinder_solid2(shape, (18, 11), 3.5)
    faces = {"-y": ("wall", None, 1.0), "+y": ("wall", (0.01, 0.0), 1.0), "-x": ("velocity_neem", (0.04, 0.003), 1.0),
             "+x": ("pressure_neem", None, 1.002)}
    eu, er, _ = run_pair(shape, 300, storage, 1 / 0.56, faces=faces, solid=solid, seed=3)
    print(f"d2q9 NEEM outlet {storage}: rel-L2 u={eu:.3e} rho={er:.3e}")
    assert eu < tol and er < tol


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f64", 1e-11)])
def test_d3q19_sphere_channel_neem_outlet(storage, tol):
    shape = (40, 16, 14)
    solid = sphere_solid(shape, (12, 8, 7), 3.5)
    f = channel_faces3()
    f["-x"] = ("velocity_neem", (0.04, 0.0, 0.0), 1.0)
    f["+x"] = ("pressure_neem", None, 1.0)
    f["+z"] = ("wall", (0.02, 0.0, 0.0), 1.0)  # moving wall next to the NEEM face
    N = int(np.prod(shape))
    u0 = np.zeros((N, 3))
    u0[~solid, 0] = 0.04
    eu, er, _ = run_pair(shape, 150, storage, 1 / 0.5384, faces=f, solid=solid, init=(np.ones(N), u0))
    print(f"d3q19 sphere NEEM outlet {storage}: rel-L2 u={eu:.3e} rho={er:.3e}")
    assert eu < tol and er < tol


def test_neem_next_to_open_face_rejected():
    lat = capi.Lattice((12, 8), storage="f32", omega=1.0)
    with pytest.raises(capi.LBMError):
        lat.set_faces({"-x": ("velocity_zh", (0.02, 0.0), 1.0), "+x": ("pressure_neem", None, 1.0),
                       "-y": ("pressure_abb", None, 1.0), "+y": ("wall", None, 1.0)})
    lat.close()


def test_set_storage_conversions_exact():
    """lbm_set_storage converts on the device: widening is exact (FP32 ->
    FP64 -> FP32 returns the same bits), narrowing rounds to nearest even
    (FP64 -> FP32 and FP32 -> FP16 match numpy's rounding of h = f - w), and
    stepping after the conversion equals a fresh handle of the new storage
    loaded with the same populations."""
    shape = (20, 12, 10)
    solid = sphere_solid(shape, (8, 6, 5), 2.5)
    N = int(np.prod(shape))
    rng = np.random.default_rng(21)
    rho0, u0 = 1 + 0.01 * rng.standard_normal(N), 0.02 * rng.standard_normal((N, 3))
    faces = channel_faces3()
    faces["+x"] = ("pressure_neem", None, 1.0)
    _, w, _ = O.tables(19)
    w = w[:, None]
    lat = capi.Lattice(shape, storage="f64", omega=1 / 0.6, faces=faces, solid=solid)
    lat.init(rho0, u0)
    lat.step(5)
    p64 = lat.populations()
    print("set_storage f64 -> f32", flush=True)
    lat.set_storage("f32")
    p32 = lat.populations()
    assert np.array_equal(p32, (p64 - w).astype(np.float32).astype(np.float64) + w)
    lat.set_storage("f64")
    assert np.array_equal(lat.populations(), p32)
    lat.set_storage("f32")
    assert np.array_equal(lat.populations(), p32)
    print("set_storage f32 -> f16", flush=True)
    lat.set_storage("f16")
    p16 = lat.populations()
    assert np.array_equal(p16, (p32 - w).astype(np.float32).astype(np.float16).astype(np.float64) + w)
    # stepping continues exactly like a fresh FP16 handle holding the same state
    fresh = capi.Lattice(shape, storage="f16", omega=1 / 0.6, faces=faces, solid=solid)
    fresh.set_populations(p16)
    lat.step(6)
    fresh.step(6)
    assert np.array_equal(lat.populations(), fresh.populations())
    s, rc = lat.step_check()
    assert rc == 0 and np.isfinite(s.res_rel)
    fresh.close()
    lat.close()


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f64", 1e-11)])
@pytest.mark.parametrize("nx", [65, 129])
def test_neem_outlet_at_wave_start(storage, tol, nx):
    """The +x NEEM face cell is the first lane of its wave (nx - 1 a multiple
    of 64), so its inward neighbour is in the previous wave and the cell
    gathers it from memory instead of taking it from the adjacent lane."""
    shape = (nx, 12)
    faces = {"-y": ("wall", None, 1.0), "+y": ("wall", None, 1.0), "-x": ("velocity_neem", (0.03, 0.0), 1.0),
             "+x": ("pressure_neem", None, 0.999)}
    eu, er, _ = run_pair(shape, 200, storage, 1 / 0.6, faces=faces, seed=8)
    assert eu < tol and er < tol


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f64", 1e-11)])
@pytest.mark.parametrize("shape", [(40, 14), (24, 10, 8)])
def test_periodic_pressure_channel(storage, tol, shape):
    """Periodic pair with a density jump driving a walled channel."""
    faces = {"-x": ("periodic_pressure", None, 1.002), "+x": ("periodic_pressure", None, 1.0),
             "-y": ("wall", None, 1.0), "+y": ("wall", (0.01, 0.0, 0.0), 1.0)}
    eu, er, _ = run_pair(shape, 150, storage, 1 / 0.7, faces=faces, seed=12)
    print(f"periodic pressure {shape} {storage}: rel-L2 u={eu:.3e} rho={er:.3e}")
    assert eu < tol and er < tol


# wet-node walls (equilibrium / NEEM / NEBB, src/lbm/bnd/bnd_wall.h:100-479)
WET_CASES = {
    "eq_neem_channel2d": ((48, 20), {"-y": ("wall_eq", None, 1.0), "+y": ("wall_neem", (0.05, 0.0), 1.0)}),
    "nebb_channel2d": ((40, 16), {"-y": ("wall_nebb", (0.02, 0.0), 1.0), "+y": ("wall_nebb", (-0.03, 0.0), 1.0)}),
    "pp_eq_neem2d": ((40, 14), {"-x": ("periodic_pressure", None, 1.002), "+x": ("periodic_pressure", None, 1.0),
                                "-y": ("wall_eq", None, 1.0), "+y": ("wall_neem", None, 1.0)}),
    "pp_nebb2d": ((36, 12), {"-x": ("periodic_pressure", None, 1.002), "+x": ("periodic_pressure", None, 1.0),
                             "-y": ("wall_nebb", None, 1.0), "+y": ("wall_nebb", None, 1.0)}),
    # corners: anti-bounce-back x faces below / above the walls in priority
    "abb_eq_low2d": ((40, 14), {"-x": ("pressure_abb", None, 1.002, 1), "+x": ("pressure_abb", None, 1.0, 1),
                                "-y": ("wall_eq", None, 1.0, 2), "+y": ("wall_eq", None, 1.0, 2)}),
    "abb_eq_mixed2d": ((40, 14), {"-x": ("pressure_abb", None, 1.002, 3), "+x": ("pressure_abb", None, 1.0, 1),
                                  "-y": ("wall_eq", None, 1.0, 4), "+y": ("wall_eq", None, 1.0, 2)}),
    # (a NEEM wall next to bounce-back walls is unstable from a noisy start,
    # on the oracle as on the device: not a parity case)
    "box_bb_eq2d": ((30, 12), {"-x": ("wall", None, 1.0, 3), "+x": ("wall", None, 1.0, 1),
                               "-y": ("wall_eq", None, 1.0, 2), "+y": ("wall", (0.03, 0.0), 1.0, 2)}),
    "eq_neem3d": ((24, 10, 8), {"-y": ("wall_eq", (0.0, 0.0, 0.01), 1.0), "+y": ("wall_neem", (0.04, 0.0, 0.0), 1.0)}),
    "edges3d": ((20, 10, 9), {"-y": ("wall_eq", None, 1.0, 2), "+y": ("wall_eq", (0.02, 0.0, 0.0), 1.0, 2),
                              "-z": ("wall", None, 1.0, 3), "+z": ("wall", None, 1.0, 1)}),
    "pp_neem3d": ((20, 10, 6), {"-x": ("periodic_pressure", None, 1.002), "+x": ("periodic_pressure", None, 1.0),
                                "-y": ("wall_neem", None, 1.0), "+y": ("wall_eq", None, 1.0)}),
}


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f64", 1e-11)])
@pytest.mark.parametrize("case", ["nebb0_channel2d", "nebb0_pp2d"])
def test_nebb_without_velocity(storage, tol, case):
    """The reference's no-velocity NEBB (apply_0): its diagonal formulas read
    the two diagonals entering through the wall, which keep their initial
    value; the wall rows start at rho 1, u 0 (the reference's initial
    condition there), which the device takes as that value."""
    shape = (36, 12)
    faces = {"-y": ("wall_nebb0", None, 1.0, 4), "+y": ("wall_nebb0", None, 1.0, 2)}
    if case == "nebb0_pp2d":
        faces.update({"-x": ("periodic_pressure", None, 1.002, 0), "+x": ("periodic_pressure", None, 1.0, 0)})
    N = shape[0] * shape[1]
    rng = np.random.default_rng(7)
    rho0 = 1 + 0.01 * rng.standard_normal(N)
    u0 = 0.02 * rng.standard_normal((N, 2))
    wall = (np.arange(N) // shape[0] == 0) | (np.arange(N) // shape[0] == shape[1] - 1)
    rho0[wall] = 1.0
    u0[wall] = 0.0
    eu, er, _ = run_pair(shape, 200, storage, 1 / 0.7, faces=faces, init=(rho0, u0))
    print(f"nebb0 {case} {storage}: rel-L2 u={eu:.3e} rho={er:.3e}")
    assert eu < tol and er < tol


@pytest.mark.parametrize("storage,tol", [("f32", 1e-5), ("f64", 1e-11)])
@pytest.mark.parametrize("case", sorted(WET_CASES))
def test_wet_node_walls(storage, tol, case):
    shape, faces = WET_CASES[case]
    eu, er, _ = run_pair(shape, 200, storage, 1 / 0.7, faces=faces, seed=21)
    print(f"wet-node {case} {storage}: rel-L2 u={eu:.3e} rho={er:.3e}")
    assert eu < tol and er < tol


def test_wet_node_wall_validation():
    lat = capi.Lattice((10, 8, 6), storage="f32", omega=1.0)
    with pytest.raises(capi.LBMError):  # NEBB is D2Q9 only
        lat.set_faces({"-y": ("wall_nebb", None, 1.0), "+y": ("wall", None, 1.0)})
    lat.close()
    lat = capi.Lattice((12, 8), storage="f32", omega=1.0)
    with pytest.raises(capi.LBMError):  # NEBB velocity must be tangential
        lat.set_faces({"-y": ("wall_nebb", (0.01, 0.02), 1.0), "+y": ("wall", None, 1.0)})
    with pytest.raises(capi.LBMError):  # NEEM wall next to a Zou-He face
        lat.set_faces({"-x": ("velocity_zh", (0.02, 0.0), 1.0), "+x": ("pressure_zh", None, 1.0),
                       "-y": ("wall_neem", None, 1.0), "+y": ("wall", None, 1.0)})
    lat.set_faces({"-x": ("pressure_abb", None, 1.001), "+x": ("pressure_abb", None, 1.0),
                   "-y": ("wall_neem", None, 1.0), "+y": ("wall_eq", None, 1.0)})
    lat.close()


def _halo_cases():
    """(shape, faces, solid) for the LDS x-halo checks: x extents that end
    inside a 256-cell row tile, periodic and walled x, solids next to the
    tile edges, NEEM faces, D2Q9 and D3Q19."""
    walls3 = {k: ("wall", None, 1.0) for k in ["-y", "+y", "-z", "+z"]}
    neem3 = dict(walls3)
    neem3["-x"] = ("velocity_neem", (0.04, 0.0, 0.0), 1.0)
    neem3["+x"] = ("pressure_neem", None, 1.0)
    xwalls3 = dict(walls3)
    xwalls3["-x"] = ("wall", (0.0, 0.02, 0.0), 1.0)
    xwalls3["+x"] = ("wall", None, 1.0)
    return {
        "box3d_periodic": ((300, 12, 10), None, None),
        "sphere3d_neem": ((290, 16, 14), neem3, sphere_solid((290, 16, 14), (255.5, 8, 7), 5.0)),
        "box3d_xwalls": ((257, 10, 8), xwalls3, sphere_solid((257, 10, 8), (0.5, 5, 4), 3.0)),
        "channel2d_neem": ((520, 24), {"-y": ("wall", None, 1.0), "+y": ("wall", (0.01, 0.0), 1.0),
                                       "-x": ("velocity_neem", (0.03, 0.0), 1.0), "+x": ("pressure_neem", None, 1.0)},
                           cylinder_solid2((520, 24), (256.0, 12.0), 5.0)),
    }


@pytest.mark.parametrize("case", ["box3d_periodic", "sphere3d_neem", "box3d_xwalls", "channel2d_neem"])
def test_lds_halo_gather_bitwise_equal(case):
    """lbm_set_halo(1) (populations with c_x != 0 staged through LDS per
    256-cell row) against the default gather from global memory: identical
    populations after plain, snapshot and check steps, identical sums."""
    shape, faces, solid = _halo_cases()[case]
    dim = len(shape)
    N = int(np.prod(shape))
    rng = np.random.default_rng(4)
    rho0 = 1 + 0.01 * rng.standard_normal(N)
    u0 = 0.02 * rng.standard_normal((N, dim))
    if solid is not None:
        u0[solid] = 0.0
    print(f"lds halo case {case} shape {shape}", flush=True)
    out = []
    for lds in (0, 1):
        lat = capi.Lattice(shape, storage="f32", omega=1 / 0.56, faces=faces, solid=solid)
        lat.set_halo(lds)
        lat.init(rho0, u0)
        lat.step(7)
        s1, rc1 = lat.step_check()
        lat.step(3)
        s2, rc2 = lat.step_ex(3)
        assert rc1 == 0 and rc2 == 0
        out.append((lat.populations(), s1.as_dict(), s2.as_dict()))
        lat.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1] and out[0][2] == out[1][2]


def test_lds_halo_matches_oracle_sphere_channel():
    """The LDS x-halo kernel against the FP64 oracle on a sphere channel with
    NEEM inlet and outlet (FP32 bar)."""
    shape, faces, solid = _halo_cases()["sphere3d_neem"]
    N = int(np.prod(shape))
    rho0 = np.ones(N)
    u0 = np.zeros((N, 3))
    u0[~solid, 0] = 0.04
    lat = capi.Lattice(shape, storage="f32", omega=1 / 0.56, faces=faces, solid=solid)
    lat.set_halo(1)
    lat.init(rho0, u0)
    lat.step(60)
    rho, u = lat.macros()
    lat.close()
    ref = O.Problem(shape, faces=faces, solid=solid, omega=1 / 0.56)
    ref.init(rho0, u0)
    ref.step(60)
    r_rho, r_u = ref.macros()
    assert rel(u, r_u) < 1e-5 and rel(rho, r_rho) < 1e-5


def test_lds_halo_rejected_where_unsupported():
    lat = capi.Lattice((100, 12, 10), storage="f32", omega=1.0)  # 128-wide blocks
    with pytest.raises(capi.LBMError, match="LDS x halo"):
        lat.set_halo(1)
    lat.close()
    lat = capi.Lattice((300, 12, 10), storage="f64", omega=1.0)
    with pytest.raises(capi.LBMError, match="LDS x halo"):
        lat.set_halo(1)
    lat.close()

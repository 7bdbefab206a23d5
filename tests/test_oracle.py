"""Pins the FP64 oracle (oracle/lbm_oracle.c) before anything is compared to it.

Each check is chosen so that a plausible slip (a wrong table entry, a sign,
a transposed index, a dropped term) fails: reference fixtures
(tests/golden), the reference's own Hilbert index built into oracle/_ref,
closed-form moments, exact streaming permutations, conservation laws and
analytic flows.
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("Q,name", [(9, "D2Q9"), (19, "D3Q19")])
def test_tables_match_reference_fixture(Q, name):
    ref = json.load(open(os.path.join(GOLD, "lattice_tables.json")))["tables"][name]
    c, w, opp = O.tables(Q)
    dim = 2 if Q == 9 else 3
    assert np.array_equal(c[:, :dim], np.array(ref["c"]))
    assert np.allclose(w, [float(Fraction(s)) for s in ref["w"]], rtol=0, atol=1e-17)
    assert list(opp) == ref["opp"]


@pytest.mark.parametrize("Q", [9, 19])
def test_lattice_invariants(Q):
    c, w, opp = O.tables(Q)
    dim = 2 if Q == 9 else 3
    assert np.all(c[Q - 1] == 0), "rest population must be last"
    assert abs(w.sum() - 1) < 1e-15
    assert np.abs(w @ c).max() < 1e-15
    m2 = np.einsum("i,ia,ib->ab", w, c, c)[:dim, :dim]
    assert np.allclose(m2, np.eye(dim) / 3, atol=1e-15)
    m3 = np.einsum("i,ia,ib,ic->abc", w, c, c, c)
    assert np.abs(m3).max() < 1e-15
    m4 = np.einsum("i,ia,ib,ic,id->abcd", w, c, c, c, c)[:dim, :dim, :dim, :dim]
    d = np.eye(dim)
    iso = (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d) + np.einsum("ad,bc->abcd", d, d)) / 9
    assert np.allclose(m4, iso, atol=1e-15)
    assert np.array_equal(opp[opp], np.arange(Q))
    assert np.array_equal(c[opp], -c)


@pytest.mark.parametrize("Q", [9, 19])
def test_equilibrium_moments(Q):
    c, w, _ = O.tables(Q)
    dim = 2 if Q == 9 else 3
    rng = np.random.default_rng(3)
    for _ in range(5):
        rho = 0.9 + 0.2 * rng.random()
        u = 0.1 * (rng.random(dim) - 0.5)
        f = O.feq(Q, rho, u)
        assert abs(f.sum() - rho) < 1e-14
        assert np.allclose(f @ c[:, :dim], rho * u, atol=1e-15)
        P = np.einsum("i,ia,ib->ab", f, c[:, :dim], c[:, :dim])
        assert np.allclose(P, rho * np.outer(u, u) + rho / 3 * np.eye(dim), atol=1e-14)
        # closed form of one population: w rho (1 + 3cu + 4.5cu^2 - 1.5u^2)
        i = 1
        cu = c[i, :dim] @ u
        assert abs(f[i] - w[i] * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * u @ u)) < 1e-15


def test_hilbert_known_answers():
    kat = json.load(open(os.path.join(GOLD, "hilbert_kat.json")))["cases"]
    assert len(kat) > 100
    for k in kat:
        if k["dim"] > 3:
            continue
        assert O.hilbert_index(k["x"], k["level"]) == k["index"], k


def test_hilbert_against_reference_build():
    if O.ref_hilbert_index((0.1, 0.2), 1) is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(0)
    for dim in (2, 3):
        for _ in range(300):
            x = rng.random(dim)
            lvl = int(rng.integers(1, 9))
            assert O.hilbert_index(x, lvl) == O.ref_hilbert_index(x, lvl)


@pytest.mark.parametrize("shape", [(7, 5), (5, 4, 6)])
def test_streaming_is_exact_shift(shape):
    """omega = 0 makes collision the identity, so one step moves every
    population by c_i exactly (periodic box): brute force over all cells."""
    p = O.Problem(shape, omega=1e-300)
    N = int(np.prod(p.shape))
    rng = np.random.default_rng(1)
    f0 = rng.random(N * p.Q)
    p.init(np.ones(N), np.zeros((N, p.dim)))
    p.fold[:] = f0
    p.omega = 0.0
    p.step(1)
    f0 = f0.reshape(N, p.Q)
    f1 = p.fold.reshape(N, p.Q)
    nx, ny, nz = p.shape
    for cidx in range(N):
        x, y, z = cidx % nx, (cidx // nx) % ny, cidx // (nx * ny)
        for i in range(p.Q):
            sx = (x - int(p.c[i, 0])) % nx
            sy = (y - int(p.c[i, 1])) % ny
            sz = (z - int(p.c[i, 2])) % nz
            assert f1[cidx, i] == f0[sx + nx * (sy + ny * sz), i]


@pytest.mark.parametrize("shape", [(12, 10), (8, 6, 7)])
def test_periodic_conservation(shape):
    p = O.Problem(shape, omega=1 / 0.7)
    N = int(np.prod(p.shape))
    rng = np.random.default_rng(2)
    rho0 = 1 + 0.01 * rng.standard_normal(N)
    u0 = 0.03 * rng.standard_normal((N, p.dim))
    p.init(rho0, u0)
    c = p.c[:, :p.dim]
    mass0 = p.fold.sum()
    mom0 = p.fold.reshape(N, p.Q).sum(0) @ c
    p.step(25)
    assert abs(p.fold.sum() - mass0) < 1e-11
    assert np.allclose(p.fold.reshape(N, p.Q).sum(0) @ c, mom0, atol=1e-12)


def test_guo_force_adds_exact_momentum():
    """Periodic box: every step adds F per cell to the total momentum."""
    F = (1e-5, -2e-5)
    p = O.Problem((10, 8), omega=1.2, force=F)
    N = 80
    p.init(np.ones(N), np.zeros((N, 2)))
    c = p.c[:, :2]
    m0 = p.fold.reshape(N, p.Q).sum(0) @ c
    p.step(7)
    m1 = p.fold.reshape(N, p.Q).sum(0) @ c
    assert np.allclose(m1 - m0, 7 * N * np.array(F), rtol=1e-10, atol=1e-16)


def test_bounce_back_closed_box_rest_and_mass():
    faces = {k: ("wall", None, 1.0) for k in ["-x", "+x", "-y", "+y", "-z", "+z"]}
    p = O.Problem((6, 5, 4), faces=faces, omega=1.1)
    N = p.fold.size // p.Q if hasattr(p, "fold") else 120
    N = 6 * 5 * 4
    rng = np.random.default_rng(5)
    p.init(1 + 0.01 * rng.standard_normal(N), 0.01 * rng.standard_normal((N, 3)))
    mass0 = p.fold.sum()
    p.step(40)
    assert abs(p.fold.sum() - mass0) < 1e-12
    # rest state is a fixed point
    q = O.Problem((6, 5, 4), faces=faces, omega=1.1)
    q.init(np.ones(N), np.zeros((N, 3)))
    f0 = q.fold.copy()
    q.step(10)
    assert np.abs(q.fold - f0).max() < 1e-15


def test_poiseuille_bounce_back_exact_tau():
    """Force-driven channel with halfway bounce-back: at tau = 1/2 + sqrt(3/16)
    the steady BGK solution is the exact parabola with walls halfway between
    nodes (He, Zou, Luo & Dembo 1997, the special case the reference's
    src/lbm/analytical_solutions.h:41 names)."""
    H = 12
    tau = 0.5 + math.sqrt(3.0 / 16.0)
    nu = (2 * tau - 1) / 6
    g = 1e-6
    p = O.Problem((4, H), faces={"-y": ("wall", None, 1.0), "+y": ("wall", None, 1.0)}, omega=1 / tau,
                  force=(g, 0.0))
    N = 4 * H
    p.init(np.ones(N), np.zeros((N, 2)))
    p.step(6000)
    rho, u = p.macros()
    y = np.arange(H) + 0.5
    ua = g / (2 * nu) * y * (H - y)
    uy = u.reshape(H, 4, 2)[:, 0, 0]
    assert np.abs(uy - ua).max() / ua.max() < 1e-9


def test_moving_wall_couette():
    """Ladd moving-wall bounce-back (rho0 = 1): steady Couette profile is linear
    through the halfway walls."""
    H = 10
    U = 0.01
    p = O.Problem((3, H), faces={"-y": ("wall", None, 1.0), "+y": ("wall", (U, 0.0), 1.0)}, omega=1.0)
    N = 3 * H
    p.init(np.ones(N), np.zeros((N, 2)))
    p.step(3000)
    rho, u = p.macros()
    y = np.arange(H) + 0.5
    assert np.abs(u.reshape(H, 3, 2)[:, 1, 0] - U * y / H).max() < 1e-10


@pytest.mark.parametrize("shape,face", [((9, 7), "-x"), ((7, 6, 5), "-x"), ((6, 7, 5), "+y")])
def test_zou_he_velocity_moments(shape, face):
    """After a step, every cell on a Zou-He velocity face carries exactly the
    prescribed velocity (the construction's defining property)."""
    dim = len(shape)
    uin = (0.03, 0.004, -0.002)[:dim]
    faces = {"-x": ("wall", None, 1.0), "+x": ("wall", None, 1.0), "-y": ("wall", None, 1.0),
             "+y": ("wall", None, 1.0)}
    if dim == 3:
        faces.update({"-z": ("wall", None, 1.0), "+z": ("wall", None, 1.0)})
    faces[face] = ("velocity_zh", uin, 1.0)
    p = O.Problem(shape, faces=faces, omega=1.3)
    N = int(np.prod(p.shape))
    rng = np.random.default_rng(9)
    p.init(1 + 0.01 * rng.standard_normal(N), 0.01 * rng.standard_normal((N, dim)))
    p.step(3)
    rho, u = p.macros()
    nx, ny, nz = p.shape
    idx = np.arange(N)
    co = [idx % nx, (idx // nx) % ny, idx // (nx * ny)]
    ax = "xyz".index(face[1])
    on = co[ax] == (0 if face[0] == "-" else p.shape[ax] - 1)
    # interior of the face (edges also touch walls, still prescribed)
    assert np.abs(u[on] - np.array(uin)).max() < 1e-13


def test_zou_he_pressure_density():
    faces = {"-y": ("wall", None, 1.0), "+y": ("wall", None, 1.0), "-x": ("velocity_zh", (0.02, 0.0), 1.0),
             "+x": ("pressure_zh", None, 0.998)}
    p = O.Problem((12, 8), faces=faces, omega=1.4)
    N = 96
    p.init(np.ones(N), np.zeros((N, 2)))
    p.step(50)
    rho, u = p.macros()
    on = (np.arange(N) % 12) == 11
    assert np.abs(rho[on] - 0.998).max() < 1e-13
    assert np.abs(u[on][:, 1]).max() < 1e-13


def test_residual_definition():
    rng = np.random.default_rng(4)
    a = rng.random((50, 4))
    b = rng.random((50, 4))
    r = O.residual(3, a.reshape(-1), b.reshape(-1))
    assert np.allclose(r[:4], np.abs(a - b).sum(0))
    assert abs(r[4] - ((a[:, :3] - b[:, :3]) ** 2).sum()) < 1e-12
    assert abs(r[5] - (a[:, :3] ** 2).sum()) < 1e-12


def _shear_wave_error(ny, tau=0.6, t=200):
    nu = (2 * tau - 1) / 6
    p = O.Problem((4, ny, 4), omega=1 / tau)
    N = 4 * ny * 4
    y = (np.arange(N) // 4) % ny
    u = np.zeros((N, 3))
    u[:, 0] = 0.01 * np.sin(2 * np.pi * (y + 0.5) / ny)
    p.init(np.ones(N), u)
    p.step(t)
    _, uu = p.macros()
    k = 2 * np.pi / ny
    return math.sqrt((uu ** 2).sum() / (u ** 2).sum()) / math.exp(-nu * k * k * t) - 1


def test_shear_wave_viscosity():
    """A shear wave u_x = u0 sin(k y) is an exact linear Navier-Stokes solution
    decaying as exp(-nu k^2 t) with nu = (2 tau - 1)/6 (src/lbm/solver.cpp:108-112);
    the lattice error must be small and shrink at second order."""
    e32, e64 = _shear_wave_error(32), _shear_wave_error(64)
    assert abs(e64) < 2e-3
    assert 3.0 < e32 / e64 < 6.5


def test_quantisers():
    assert O.round_f16(1.0 + 2 ** -11) == 1.0  # tie to even
    assert O.round_f16(1.0 + 3 * 2 ** -11) == 1.0 + 2 ** -9
    assert O.round_f16(65520.0) == math.inf
    assert O.round_f16(2 ** -25) == 0.0 and O.round_f16(3 * 2 ** -26) == 2 ** -24
    assert O.round_bf16(1.0 + 2 ** -8) == 1.0
    assert O.round_bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6


def test_splitmix_reference_values():
    # splitmix64 with seed 0: first output 0xe220a8397b1dcdaf (published test vector)
    assert int(O.splitmix64(1, 0)[0]) == 0xE220A8397B1DCDAF
    g = O.gaussian_splitmix(200000, 11)
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1) < 0.01


@pytest.mark.parametrize("kind", ["pressure_neem", "velocity_neem"])
@pytest.mark.parametrize("shape,periodic_y,solid_cell", [((7, 5), True, None), ((7, 5), False, (4, 2)),
                                                          ((6, 4, 5), True, None), ((6, 5, 4), False, (3, 1, 2))])
def test_neem_face_brute_force(shape, periodic_y, solid_cell, kind):
    """NEEM face on +x, checked cell by cell with omega = 0 (the
    post-collision state equals the pre-step state, so one step is pure
    streaming plus boundaries): every population of a face cell equals the
    post-streaming population g of its inward neighbour n1 (gathered from
    n1 - c_i, bounced back where that cell is missing) plus
    w_i (rho_b P_i(u_b) - rho1 P_i(u1)), P_i(u) = 1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u^2,
    (rho1, u1) the moments of g and (rho_b, u_b) = (rho_face, u1) on a
    pressure face, (rho1, u_face) on a velocity face."""
    dim = len(shape)
    rho_face = 1.013
    u_face = np.array([0.031, -0.007, 0.004][:dim])
    faces = {"-x": ("wall", None, 1.0), "+x": (kind, tuple(u_face), rho_face)}
    if not periodic_y:
        faces.update({"-y": ("wall", None, 1.0), "+y": ("wall", None, 1.0)})
    full = tuple(shape) + ((1,) if dim == 2 else ())
    nx, ny, nz = full
    solid = np.zeros(nx * ny * nz, bool)
    if solid_cell is not None:
        sc = tuple(solid_cell) + ((0,) if dim == 2 else ())
        solid[sc[0] + nx * (sc[1] + ny * sc[2])] = True
    p = O.Problem(shape, faces=faces, solid=solid, omega=0.0)
    N = nx * ny * nz
    rng = np.random.default_rng(4)
    p.init(1 + 0.05 * rng.standard_normal(N), 0.05 * rng.standard_normal((N, dim)))
    Q = p.Q
    f0 = p.fold.reshape(N, Q).copy()
    p.step(1)
    f1 = p.fold.reshape(N, Q)
    c, w, opp = O.tables(Q)
    ext = [nx, ny, nz]
    per = [False, periodic_y, True]

    def gathered(x, y, z):
        cell = x + nx * (y + ny * z)
        g = np.zeros(Q)
        for i in range(Q):
            src = [x - int(c[i, 0]), y - int(c[i, 1]), z - int(c[i, 2])]
            ok = True
            for a in range(3):
                if src[a] < 0 or src[a] >= ext[a]:
                    if per[a]:
                        src[a] %= ext[a]
                    else:
                        ok = False
            s = src[0] + nx * (src[1] + ny * src[2]) if ok else -1
            g[i] = f0[s, i] if ok and not solid[s] else f0[cell, opp[i]]
        return g

    checked = 0
    x = nx - 1
    for z in range(nz):
        for y in range(ny):
            cell = x + nx * (y + ny * z)
            if solid[cell] or solid[cell - 1]:
                continue
            g = gathered(x - 1, y, z)
            rho1 = g.sum()
            u1 = (c[:, :dim] * g[:, None]).sum(axis=0) / rho1

            def P(u):
                cu = c[:, :dim] @ u
                return 1 + 3 * cu + 4.5 * cu * cu - 1.5 * (u @ u)

            rb, ub = (rho_face, u1) if kind == "pressure_neem" else (rho1, u_face)
            expect = g + w * (rb * P(ub) - rho1 * P(u1))
            assert np.allclose(f1[cell], expect, rtol=0, atol=1e-15), (y, z)
            checked += 1
    assert checked > 0


def test_neem_face_moments():
    """The defining moments of the NEEM face: after every step a face cell
    carries exactly the prescribed density and the velocity of its inward
    neighbour."""
    faces = {"-y": ("wall", None, 1.0), "+y": ("wall", None, 1.0), "-x": ("velocity_zh", (0.03, 0.0), 1.0),
             "+x": ("pressure_neem", None, 0.997)}
    p = O.Problem((20, 9), faces=faces, omega=1 / 0.7)
    N = 180
    rng = np.random.default_rng(2)
    p.init(1 + 0.01 * rng.standard_normal(N), 0.01 * rng.standard_normal((N, 2)))
    for _ in range(3):
        p.step(7)
        rho, u = p.macros()
        on = np.arange(N) % 20 == 19
        assert np.abs(rho[on] - 0.997).max() < 1e-14
        assert np.abs(u[on] - u[np.nonzero(on)[0] - 1]).max() < 1e-15


def test_neem_velocity_face_moments():
    """A NEEM velocity face carries exactly the prescribed velocity and the
    density of its inward neighbour after every step."""
    faces = {"-y": ("wall", None, 1.0), "+y": ("wall", None, 1.0), "-x": ("velocity_neem", (0.03, 0.002), 1.0),
             "+x": ("pressure_neem", None, 1.0)}
    p = O.Problem((20, 9), faces=faces, omega=1 / 0.7)
    N = 180
    rng = np.random.default_rng(5)
    p.init(1 + 0.01 * rng.standard_normal(N), 0.01 * rng.standard_normal((N, 2)))
    for _ in range(3):
        p.step(5)
        rho, u = p.macros()
        on = np.arange(N) % 20 == 0
        assert np.abs(u[on] - np.array([0.03, 0.002])).max() < 1e-15
        assert np.abs(rho[on] - rho[np.nonzero(on)[0] + 1]).max() < 1e-14


def test_neem_outlet_converges_behind_cylinder_at_low_viscosity():
    """Cylinder in a walled channel at tau = 0.5384 (the viscosity of the
    Re = 100 sphere benchmark): with the NEEM outlet the one-step relative
    residual falls below 1e-8 and the mass flux is the same through every
    cross-section (to the drift of the slowly relaxing density level).  (The anti-bounce-back outlet diverges at this tau and the
    Zou-He pressure outlet keeps an undamped period-2 mode at the face.)"""
    nx, ny, uin = 96, 24, 0.04
    faces = {"-y": ("wall", None, 1.0), "+y": ("wall", None, 1.0), "-x": ("velocity_zh", (uin, 0.0), 1.0),
             "+x": ("pressure_neem", None, 1.0)}
    yy, xx = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    solid = (((xx + 0.5 - 24) ** 2 + (yy + 0.5 - 12) ** 2) < 9.0).reshape(-1)
    p = O.Problem((nx, ny), faces=faces, solid=solid, omega=1 / 0.5384)
    N = nx * ny
    u0 = np.zeros((N, 2))
    u0[~solid, 0] = uin
    p.init(np.ones(N), u0)
    res = 1.0
    for _ in range(40):
        p.step(999)
        _, ua = p.macros()
        p.step(1)
        rho, ub = p.macros()
        res = np.linalg.norm(ub - ua) / np.linalg.norm(ub)
        if res < 1e-8:
            break
    assert res < 1e-8, res
    # the slow density mode still drifts at ~1e-6 of the flux per section
    flux = (rho * ub[:, 0]).reshape(ny, nx).sum(axis=0)
    assert np.abs(flux / flux[0] - 1).max() < 1e-4


@pytest.mark.parametrize("shape", [(7, 5), (6, 4, 5)])
def test_periodic_pressure_face_brute_force(shape):
    """Periodic pair with a density jump (bnd_periodic.h:102-108), checked with
    omega = 0: a population entering a face cell L across the face equals
    feq(p', u_c) + f(c) - feq(rho_c, u_c) of the cell c straight across on the
    other face, p' the pressure of c's face; walls win at the corners."""
    dim = len(shape)
    p_lo, p_hi = 1.004, 0.997
    faces = {"-x": ("periodic_pressure", None, p_lo), "+x": ("periodic_pressure", None, p_hi),
             "-y": ("wall", None, 1.0), "+y": ("wall", None, 1.0)}
    p = O.Problem(shape, faces=faces, omega=0.0)
    full = tuple(shape) + ((1,) if dim == 2 else ())
    nx, ny, nz = full
    N = nx * ny * nz
    rng = np.random.default_rng(6)
    p.init(1 + 0.05 * rng.standard_normal(N), 0.05 * rng.standard_normal((N, dim)))
    Q = p.Q
    f0 = p.fold.reshape(N, Q).copy()
    p.step(1)
    f1 = p.fold.reshape(N, Q)
    c, w, opp = O.tables(Q)

    def feq(rho, u):
        cu = c[:, :dim] @ u
        return w * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * (u @ u))

    checked = 0
    for z in range(nz):
        for y in range(ny):
            for x, xo, pp in ((nx - 1, 0, p_lo), (0, nx - 1, p_hi)):
                cell = x + nx * (y + ny * z)
                other = xo + nx * (y + ny * z)
                rho_c = f0[other].sum()
                u_c = (c[:, :dim] * f0[other][:, None]).sum(axis=0) / rho_c
                ref = feq(pp, u_c) + f0[other] - feq(rho_c, u_c)
                for o in range(Q - 1):
                    if int(c[o, 0]) != (-1 if x == nx - 1 else 1):
                        continue  # not entering across this x face
                    if not (0 <= y - int(c[o, 1]) < ny):
                        continue  # crosses a wall too: the wall wins
                    assert abs(f1[cell, o] - ref[o]) < 1e-15, (x, y, z, o)
                    checked += 1
    assert checked > 0

"""Steps of a domain with faces or solids choose a path per wave: a wave
whose cells are all interior (no face, wrap or solid within one cell) runs
without the boundary code and the flag loads, the others run the general
path; LBM_STEP_WAVE_PATHS=0 at lbm_create keeps the general path for every
wave.  Both apply the same arithmetic to every cell, so the populations
agree bit for bit, after plain steps, snapshot/check pairs and a change of
the launch shape (which rebuilds the per-wave masks)."""
import os

import numpy as np
import pytest

from lbm_amd import capi
from lbm_amd import workloads as W

pytestmark = pytest.mark.gpu


def make(shape, storage, faces, solid, wave_paths, omega):
    old = os.environ.get("LBM_STEP_WAVE_PATHS")
    os.environ["LBM_STEP_WAVE_PATHS"] = "1" if wave_paths else "0"
    try:
        return capi.Lattice(shape, storage=storage, omega=omega, faces=faces, solid=solid)
    finally:
        if old is None:
            del os.environ["LBM_STEP_WAVE_PATHS"]
        else:
            os.environ["LBM_STEP_WAVE_PATHS"] = old


def sphere_case():
    nx, ny, nz = 96, 40, 36
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    solid = ((x + 0.5 - 30) ** 2 + (y + 0.5 - 20) ** 2 + (z + 0.5 - 18) ** 2 < 36.0).astype(np.uint8).reshape(-1)
    faces = {k: ("wall", None, 1.0) for k in ["-y", "+y", "-z", "+z"]}
    faces["-x"] = ("velocity_neem", (0.04, 0.0, 0.0), 1.0)
    faces["+x"] = ("pressure_neem", None, 1.0)
    return (nx, ny, nz), faces, solid


def step_case():
    nx, ny = 150, 50
    solid = np.zeros((ny, nx), np.uint8)
    solid[:20, :40] = 1
    return (nx, ny), W.channel_faces(0.03), solid.reshape(-1)


def periodic_case():
    # periodic along x with a block next to the wrap (links across the face)
    nx, ny = 128, 48
    solid = np.zeros((ny, nx), np.uint8)
    solid[10:30, 120:128] = 1
    solid[10:30, 0:3] = 1
    faces = {"-y": ("wall", (0.0, 0.0, 0.0), 1.0), "+y": ("wall", (0.02, 0.0, 0.0), 1.0)}
    return (nx, ny), faces, solid.reshape(-1)


CASES = {"sphere": sphere_case, "step": step_case, "periodic": periodic_case}


@pytest.mark.parametrize("case,storage,launch", [("sphere", "f32", None), ("sphere", "f16", None),
                                                 ("sphere", "bf16", None), ("sphere", "f64", None),
                                                 ("step", "f32", None), ("step", "f16", None),
                                                 ("periodic", "f32", None), ("step", "f32", (64, 4)),
                                                 ("sphere", "f32", (128, 2))])
def test_wave_paths_equal_general_path(case, storage, launch):
    shape, faces, solid = CASES[case]()
    n = int(np.prod(shape))
    dim = len(shape)
    rng = np.random.default_rng(5)
    rho = 1 + 0.01 * rng.standard_normal(n)
    u = 0.02 * rng.standard_normal((n, dim))
    print(f"wave paths {case} {storage} launch {launch}", flush=True)
    lats = [make(shape, storage, faces, solid, wp, 1.5) for wp in (True, False)]
    for lat in lats:
        if launch:
            lat.set_launch(*launch)
        lat.init(rho, u)
        lat.step(7)
        s = lat.step_check()
        lat.step(5)
    f0, f1 = lats[0].populations(), lats[1].populations()
    assert np.array_equal(f0, f1)
    r0, u0 = lats[0].macros()
    r1, u1 = lats[1].macros()
    assert np.array_equal(r0, r1) and np.array_equal(u0, u1)
    for lat in lats:
        lat.close()

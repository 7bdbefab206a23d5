"""The "line" postprocessing function: its restatement pinned to textbook
geometry, and the host's cell selection against it on the reference's own
configurations."""
import ctypes as ct
import json
import math
import os

import numpy as np
import pytest

from oracle import post_oracle as P
from tests.test_grid import host, host_grid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFCFG = os.path.join(ROOT, "tests", "golden", "reference_configs")


def test_hyperplane_2d_is_point_line_distance():
    rng = np.random.default_rng(0)
    for _ in range(200):
        a, b = rng.uniform(-2, 2, 2), rng.uniform(-2, 2, 2)
        p = rng.uniform(-3, 3, (5, 2))
        d = b - a
        textbook = np.abs(d[0] * (p[:, 1] - a[1]) - d[1] * (p[:, 0] - a[0])) / math.hypot(*d)
        assert np.allclose(P.distances(p, a, b), textbook, rtol=1e-12, atol=1e-14)


def test_hyperplane_3d_contains_the_line():
    rng = np.random.default_rng(1)
    cases = [rng.uniform(-2, 2, (2, 3)) for _ in range(100)]
    cases.append(np.array([[0.3, 0.2, 0.0], [0.3, 0.2, 1.0]]))  # along z: the other branch
    for a, b in cases:
        n, off = P.line_hyperplane(a, b)
        assert abs(np.linalg.norm(n) - 1) < 1e-14
        assert abs(np.dot(n, b - a)) < 1e-13 * np.linalg.norm(b - a)
        assert P.distances(np.array([a, b, 0.3 * a + 0.7 * b]), a, b).max() < 1e-13


def host_line(cfg, a, b):
    L = host()
    L.lbm_host_grid_line_cells.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int64]
    L.lbm_host_grid_line_cells.restype = ct.c_int64
    g = ct.c_void_p()
    assert L.lbm_host_grid_create(json.dumps(cfg).encode(), ct.byref(g)) == 0
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    n = L.lbm_host_grid_line_cells(g, a.ctypes.data, b.ctypes.data, None, 0)
    out = np.zeros(max(n, 1), dtype=np.int64)
    assert L.lbm_host_grid_line_cells(g, a.ctypes.data, b.ctypes.data, out.ctypes.data, n) == n
    L.lbm_host_grid_destroy(g)
    return out[:n]


def cell_length(cen):
    xs = np.unique(cen[:, 0])
    return float(np.min(np.diff(xs)))


LINE_CFGS = ["poiseuille__poiseuille.json", "couette__couette_bnd_eq.json", "step__step_ns.json"]


@pytest.mark.parametrize("name", LINE_CFGS)
def test_host_line_cells_equal_restatement(name):
    cfg = json.load(open(os.path.join(REFCFG, name)))
    h = host_grid(cfg)
    cen, hl = h["cen"], cell_length(h["cen"])
    lines = [(f["A"], f["B"]) for f in cfg["solver"]["postprocessing"].values()]
    lo, hi = cen.min(axis=0), cen.max(axis=0)
    # plus a diagonal line and a horizontal one through the domain
    lines += [(list(lo), list(hi)), ([lo[0], 0.37 * lo[1] + 0.63 * hi[1]], [hi[0], 0.37 * lo[1] + 0.63 * hi[1]])]
    for a, b in lines:
        got = host_line(cfg, a, b)
        want = P.line_cells(cen, hl, a, b)
        assert len(want) > 0
        assert np.array_equal(got, want), (name, a, b)


def test_axis_parallel_line_is_one_column():
    """x = 0.501 on the Poiseuille grid: exactly the cells with
    |x_c - 0.501| <= h/2, bottom to top."""
    cfg = json.load(open(os.path.join(REFCFG, "poiseuille__poiseuille.json")))
    f = list(cfg["solver"]["postprocessing"].values())[0]
    h = host_grid(cfg)
    cen, hl = h["cen"], cell_length(h["cen"])
    got = host_line(cfg, f["A"], f["B"])
    col = np.nonzero(np.abs(cen[:, 0] - f["A"][0]) <= hl / 2)[0]
    assert len(np.unique(cen[col, 0])) == 1
    assert np.array_equal(got, col[np.argsort(cen[col, 1])])

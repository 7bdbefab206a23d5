"""Grid indexing and neighbour maps of the host layer, bit-exact against the
literal restatement of the reference's generator (oracle/grid_oracle.py).

Cases: the reference's own integration configs (tests/golden/reference_configs)
plus two 3D configs of this build.  Compared: number of cells, cell centres
(bitwise), the full neighbour table including periodic links and diagonals,
boundary-cell flags, and every surface's cell list and normal direction."""
import ctypes as ct
import json
import os

import numpy as np
import pytest

from oracle import grid_oracle as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFCFG = os.path.join(ROOT, "tests", "golden", "reference_configs")
OWNCFG = os.path.join(ROOT, "tests", "golden", "configs")


def host():
    L = ct.CDLL(os.path.join(ROOT, "lbm_amd", "liblbm_host.so"))
    P = ct.c_void_p
    L.lbm_host_grid_create.argtypes = [ct.c_char_p, ct.POINTER(P)]
    L.lbm_host_grid_ncells.argtypes = [P]
    L.lbm_host_grid_ncells.restype = ct.c_int64
    L.lbm_host_grid_dim.argtypes = [P]
    L.lbm_host_grid_stride.argtypes = [P]
    L.lbm_host_grid_centers.argtypes = [P, P]
    L.lbm_host_grid_neighbors.argtypes = [P, P]
    L.lbm_host_grid_bndry.argtypes = [P, P]
    L.lbm_host_grid_structured.argtypes = [P, P, P]
    L.lbm_host_grid_surface_count.argtypes = [P]
    L.lbm_host_grid_surface.argtypes = [P, ct.c_int32, ct.c_char_p, ct.c_int32, ct.POINTER(ct.c_int64), P, P]
    L.lbm_host_grid_destroy.argtypes = [P]
    return L


def host_grid(cfg):
    L = host()
    g = ct.c_void_p()
    assert L.lbm_host_grid_create(json.dumps(cfg).encode(), ct.byref(g)) == 0
    n = L.lbm_host_grid_ncells(g)
    dim = L.lbm_host_grid_dim(g)
    stride = L.lbm_host_grid_stride(g)
    cen = np.zeros(n * dim)
    L.lbm_host_grid_centers(g, cen.ctypes.data)
    nb = np.zeros(n * stride, dtype=np.int64)
    L.lbm_host_grid_neighbors(g, nb.ctypes.data)
    bnd = np.zeros(n, dtype=np.uint8)
    L.lbm_host_grid_bndry(g, bnd.ctypes.data)
    nxyz = np.zeros(3, dtype=np.int64)
    sidx = np.zeros(n, dtype=np.int64)
    L.lbm_host_grid_structured(g, nxyz.ctypes.data, sidx.ctypes.data)
    surf = {}
    for i in range(L.lbm_host_grid_surface_count(g)):
        name = ct.create_string_buffer(128)
        k = ct.c_int64()
        L.lbm_host_grid_surface(g, i, name, 128, ct.byref(k), None, None)
        cells = np.zeros(k.value, dtype=np.int64)
        dirs = np.zeros(k.value, dtype=np.int64)
        L.lbm_host_grid_surface(g, i, name, 128, ct.byref(k), cells.ctypes.data, dirs.ctypes.data)
        surf[name.value.decode()] = (list(cells), list(dirs))
    L.lbm_host_grid_destroy(g)
    return dict(n=n, dim=dim, stride=stride, cen=cen.reshape(n, dim), nb=nb.reshape(n, stride), bnd=bnd,
                nxyz=nxyz, sidx=sidx, surf=surf)


def oracle_grid(cfg):
    """The literal restatement's grid in host_grid's form: cell count,
    centres, neighbour table, boundary flags, surfaces, and the structured
    window (extent per axis and each cell's structured index, read off the
    distinct centre coordinates)."""
    gen, ref = G.build(cfg)
    n, dim = ref.n, gen.dim
    cen = np.array(ref.center, dtype=np.float64).reshape(n, dim)
    nb = np.array(ref.nghbr, dtype=np.int64).reshape(n, -1)
    nxyz = np.ones(3, dtype=np.int64)
    idx = np.zeros((n, 3), dtype=np.int64)
    for a in range(dim):
        vals, inv = np.unique(cen[:, a], return_inverse=True)
        nxyz[a] = vals.size
        idx[:, a] = inv
    sidx = idx[:, 0] + nxyz[0] * (idx[:, 1] + nxyz[1] * idx[:, 2])
    surf = {k: (list(cells), [dirs[c] for c in cells]) for k, (cells, dirs) in ref.surfaces.items()}
    return dict(n=n, dim=dim, stride=nb.shape[1], cen=cen, nb=nb, bnd=np.array(ref.bndry, dtype=np.uint8), nxyz=nxyz,
                sidx=sidx, surf=surf)


def assert_same_grid(h, g):
    """Host grid == restated grid, bit for bit (cell order and centres,
    neighbours incl. periodic links and diagonals, boundary flags, surfaces
    with their normals, structured window)."""
    assert h["n"] == g["n"]
    assert np.array_equal(h["cen"], g["cen"]), "cell centres / cell order differ"
    assert np.array_equal(h["nb"], g["nb"]), "neighbour tables differ"
    assert np.array_equal(h["bnd"].astype(bool), g["bnd"].astype(bool))
    assert h["surf"] == g["surf"]
    assert list(h["nxyz"]) == list(g["nxyz"]) and np.array_equal(h["sidx"], g["sidx"])


CASES = [
    (REFCFG, "couette__couette.json"),
    (REFCFG, "couette__couette_bnd_bbDirichlet.json"),
    (REFCFG, "poiseuille__poiseuille_bnd_pressure.json"),
    (REFCFG, "couette__couette_bnd_eq_aligned.json"),
    (REFCFG, "poiseuille__poiseuille_bnd_NEEM.json"),
    (REFCFG, "step__step_ns.json"),
    (OWNCFG, "box3d_periodic_x.json"),
    (OWNCFG, "sphere3d_channel.json"),
]


@pytest.mark.parametrize("folder,name", CASES)
def test_grid_matches_reference_generator(folder, name):
    cfg = json.load(open(os.path.join(folder, name)))
    gen, ref = G.build(cfg)
    h = host_grid(cfg)
    assert h["n"] == ref.n
    assert np.array_equal(h["cen"], np.array(ref.center)), "cell centres / cell order differ"
    assert np.array_equal(h["nb"], np.array(ref.nghbr)), "neighbour tables differ"
    assert np.array_equal(h["bnd"].astype(bool), np.array(ref.bndry))
    assert set(h["surf"]) == set(ref.surfaces)
    for k, (cells, dirs) in ref.surfaces.items():
        hc, hd = h["surf"][k]
        assert hc == cells, k
        assert hd == [dirs[c] for c in cells], k


@pytest.mark.slow
def test_grid_sphere_reference_config():
    cfg = json.load(open(os.path.join(REFCFG, "sphere__sphere_ns.json")))
    cfg["partitionLevel"] = cfg["uniformLevel"] = cfg["maxRfnmtLvl"] = 6
    gen, ref = G.build(cfg)
    h = host_grid(cfg)
    assert h["n"] == ref.n
    assert np.array_equal(h["cen"], np.array(ref.center))
    assert np.array_equal(h["nb"], np.array(ref.nghbr))


def test_structured_view_is_consistent():
    """Structured index -> neighbour relation: every neighbour link that is not
    periodic goes to the structured neighbour cell."""
    cfg = json.load(open(os.path.join(OWNCFG, "sphere3d_channel.json")))
    h = host_grid(cfg)
    nx, ny, nz = h["nxyz"]
    co = np.stack([h["sidx"] % nx, (h["sidx"] // nx) % ny, h["sidx"] // (nx * ny)], 1)
    for d, v in enumerate(G.DIR_VEC3[:18]):
        ok = h["nb"][:, d] >= 0
        tgt = co[h["nb"][ok, d]]
        assert np.array_equal(tgt, co[ok] + np.array(v))
    assert h["n"] < nx * ny * nz  # sphere cells removed


def test_couette_cell_count_matches_box_derivation():
    """couette.json's box at its level: 20 x 32 cells (the faces sit on the
    outer cell centres), 640 in all."""
    cfg = json.load(open(os.path.join(REFCFG, "couette__couette.json")))
    h = host_grid(cfg)
    assert h["n"] == 640 and list(h["nxyz"]) == [20, 32, 1]

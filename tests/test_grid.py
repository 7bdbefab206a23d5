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
                sidx=sidx, surf=surf, level=np.array(ref.level, dtype=np.int64))


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


# ------------------------------------------------------------ multi-level grids
def host_levels(cfg):
    L = host()
    P = ct.c_void_p
    L.lbm_host_grid_levels.argtypes = [P, P, P]
    g = ct.c_void_p()
    assert L.lbm_host_grid_create(json.dumps(cfg).encode(), ct.byref(g)) == 0
    n = L.lbm_host_grid_ncells(g)
    lvl = np.zeros(n, dtype=np.int32)
    par = np.zeros(n, dtype=np.int64)
    L.lbm_host_grid_levels(g, lvl.ctypes.data, par.ctypes.data)
    L.lbm_host_grid_destroy(g)
    return lvl, par


def leveled(folder, name, p, u, m):
    cfg = json.load(open(os.path.join(folder, name)))
    cfg["partitionLevel"], cfg["uniformLevel"], cfg["maxRfnmtLvl"] = p, u, m
    cfg["maxNoCells"] = max(cfg["maxNoCells"], 4_000_000)
    return cfg


MULTI = [
    (REFCFG, "couette__couette.json", 4, 6, 6),        # two uniform refinements
    (REFCFG, "couette__couette.json", 4, 5, 7),        # uniform, then two boundary refinements
    (REFCFG, "step__step_ns.json", 5, 6, 7),           # several boxes (a step), both kinds
    (REFCFG, "poiseuille__poiseuille_bnd_NEEM.json", 4, 5, 5),  # alignNodesWithSurface on the finest level
    (OWNCFG, "box3d_periodic_x.json", 3, 4, 4),        # 3D, periodic links on every level
    (OWNCFG, "sphere3d_channel.json", 3, 4, 5),        # 3D, a subtracted sphere, boundary refinement
]


@pytest.mark.parametrize("folder,name,p,u,m", MULTI)
def test_multilevel_grid_matches_reference_generator(folder, name, p, u, m):
    """partitionLevel < uniformLevel <= maxRfnmtLvl: every level kept, the
    finer ones appended in parent order with the reference's deletion
    compaction; cell order, centres, levels, parents, neighbour tables
    (periodic links and diagonals within each level), boundary flags (the
    parent rule, each level's cell length) and surfaces bit-equal to the
    literal restatement."""
    cfg = leveled(folder, name, p, u, m)
    gen, ref = G.build(cfg)
    h = host_grid(cfg)
    lvl, par = host_levels(cfg)
    assert h["n"] == ref.n
    assert sorted(set(lvl)) == sorted(set(ref.level)) and len(set(lvl)) > 1
    assert np.array_equal(lvl, np.array(ref.level))
    assert np.array_equal(par, np.array(ref.parent))
    assert np.array_equal(h["cen"], np.array(ref.center)), "cell centres / cell order differ"
    assert np.array_equal(h["nb"], np.array(ref.nghbr)), "neighbour tables differ"
    assert np.array_equal(h["bnd"].astype(bool), np.array(ref.bndry))
    assert set(h["surf"]) == set(ref.surfaces)
    for k, (cells, dirs) in ref.surfaces.items():
        hc, hd = h["surf"][k]
        assert hc == cells, k
        assert hd == [dirs[c] for c in cells], k
    # links never cross levels
    nb = h["nb"]
    ok = nb >= 0
    assert np.all(lvl[np.nonzero(ok)[0]] == lvl[nb[ok]])


def test_multilevel_restatement_invariants():
    """What the tree fixes independently of the restated steps, on a
    uniformly refined box (no child is deleted) and a boundary-refined one:
    every finer cell sits at its parent's centre +- half its own length
    along each axis; with uniform refinement each level has 2^dim times the
    cells of the one below and the children of parent k of a level are
    cells begin + 2^dim k .. +2^dim - 1 in childDir order; the neighbour
    links of every level are exactly the axis neighbours found by brute
    force among that level's centres; boundary refinement refines exactly
    the generator's boundary cells of the level below."""
    cases = [("poisson__poisson2D.json", 3, 5, 5),  # the unit square: every child stays
             ("couette__couette.json", 3, 5, 5), ("couette__couette.json", 3, 4, 5)]
    for (name, p, u, m) in cases:
        cfg = leveled(REFCFG, name, p, u, m)
        cfg["alignNodesWithSurface"] = False  # keep the finest centres on the tree
        cfg["solver"]["boundary"] = {}  # no periodic links: the tables are the tree's own
        gen, ref = G.build(cfg)
        n, dim = ref.n, gen.dim
        cen = np.array(ref.center).reshape(n, dim)
        lvl = np.array(ref.level)
        par = np.array(ref.parent)
        nb = np.array(ref.nghbr).reshape(n, -1)
        for q in range(n):
            if par[q] < 0:
                assert lvl[q] == p
                continue
            half = 0.5 * gen.len[lvl[q]]
            assert lvl[par[q]] == lvl[q] - 1
            assert np.allclose(np.abs(cen[q] - cen[par[q]]), half, rtol=0, atol=1e-12)
        levels = sorted(set(lvl.tolist()))
        counts = [int((lvl == l).sum()) for l in levels]
        if name.startswith("poisson"):
            assert all(counts[i + 1] == (1 << dim) * counts[i] for i in range(len(counts) - 1))
            for l in levels[1:]:
                ids = np.nonzero(lvl == l)[0]
                parents = np.nonzero(lvl == l - 1)[0]
                b0 = ids[0]
                for k, pp in enumerate(parents):
                    kids = np.arange(b0 + (1 << dim) * k, b0 + (1 << dim) * (k + 1))
                    assert np.all(par[kids] == pp)
                    for ch, c in enumerate(kids):
                        sgn = [G.CHILD_DIR[ch][d] for d in range(dim)]
                        assert np.all(np.sign(cen[c] - cen[pp]) == sgn)
        elif u < m:
            top = levels[-1]
            refined = sorted(set(par[lvl == top].tolist()))
            gb = [c for c in range(n) if lvl[c] == top - 1 and gen.bndry[c]]
            assert refined == gb
        for l in levels:
            ids = np.nonzero(lvl == l)[0]
            h = gen.len[l]
            key = {tuple(np.round(cen[c] / h, 6)): c for c in ids}
            for c in ids:
                for d in range(2 * dim):
                    v = np.zeros(dim)
                    v[d // 2] = 1 if d % 2 else -1
                    want = key.get(tuple(np.round(cen[c] / h + v, 6)), -1)
                    assert nb[c, d] == want, (l, c, d)

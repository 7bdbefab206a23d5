"""Batched uneven domains (include/lbm_batch.h) on the GPU.

A domain of a batch follows exactly the single-domain rules, so after the
same number of steps its populations equal those of the same domain run
alone through lbm_step, bit for bit.  The batch runs its steps from a
hipGraph; the graph is replayed several times with host synchronisation and
eager single-domain work in between, and every replay is checked."""
import numpy as np
import pytest

from lbm_amd import capi
from lbm_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def small_domains(seed=5):
    rng = np.random.default_rng(seed)
    doms = []
    for k, (nx, ny) in enumerate([(37, 21), (65, 32), (70, 9), (33, 40), (16, 16), (93, 57)]):
        solid = np.zeros((ny, nx), dtype=np.uint8)
        if k % 2 == 0:
            solid[: ny // 3, : nx // 4] = 1  # a backward-facing step
        # domain 4 fully periodic; 1 and 3 with NEEM inlet/outlet (their x
        # faces exercise both the in-wave neighbour exchange and its
        # gathering fallback, depending on where the face falls in a tile)
        faces = {} if k == 4 else (W.neem_channel_faces(0.03) if k in (1, 3) else W.channel_faces(0.03))
        doms.append({"shape": (nx, ny), "omega": float(rng.uniform(1.0, 1.8)), "solid": solid.reshape(-1),
                     "faces": faces})
    return doms


def random_init(doms, seed=1):
    rng = np.random.default_rng(seed)
    rhos, us = [], []
    for d in doms:
        n = d["shape"][0] * d["shape"][1]
        rhos.append(1 + 0.01 * rng.standard_normal(n))
        us.append(0.02 * rng.standard_normal((n, 2)))
    return rhos, us


@pytest.mark.parametrize("storage,tile,blocks,one", [("f32", 64, 0, 1), ("f64", 64, 0, 1), ("f16", 64, 0, 1),
                                                     ("f32", 32, 0, 1), ("f32", 128, 0, 1), ("f32", 0, 0, 1),
                                                     ("f64", 0, 0, 1), ("f16", 0, 0, 1), ("f32", 0, 24, 1),
                                                     ("f32", 16, 0, 1), ("f32", 32, 0, 0), ("f64", 64, 0, 0),
                                                     ("f32", 0, 0, 0), ("f32", 64, 24, 0), ("f32", 1, 0, 1),
                                                     ("f16", 1, 0, 1), ("f32", 1, 0, 0)])
def test_batch_equals_single_domain_over_graph_replays(storage, tile, blocks, one):
    """tile 0: row runs (a NEEM face column is a tile of its own, so its
    cells gather their inward neighbours instead of taking them from the
    next lane); blocks > 0: a persistent grid; one 0: interior tiles and
    the others in two launches."""
    doms = small_domains()
    rhos, us = random_init(doms)
    print(f"batch vs single {storage} tile {tile} blocks {blocks} one launch {one}", flush=True)
    b = capi.Batch(doms, storage=storage)
    b.set_tiling(tile, blocks)
    b.set_launch(one)
    b.init(np.concatenate(rhos), np.concatenate(us))
    singles = []
    for d, r, u in zip(doms, rhos, us):
        lat = capi.Lattice(d["shape"], storage=storage, omega=d["omega"], faces=d["faces"] or None,
                           solid=d["solid"])
        lat.init(r, u)
        singles.append(lat)
    interval = 6
    for replay in range(3):
        b.run(interval, 0.0, interval)  # tol 0: nobody retires; one graph replay
        for k, lat in enumerate(singles):
            lat.step(interval)  # eager work between replays
            assert np.array_equal(b.populations(k), lat.populations()), (replay, k)
    for lat in singles:
        lat.close()
    b.close()


def interior_domains(dim):
    """Domains with many interior tiles (no face, wrap or solid within one
    cell), which run in the kernel without boundary paths: odd numbers of
    interior tile rows, a solid block that splits columns of tiles, and
    different widths."""
    if dim == 3:
        faces = {"-y": ("wall", (0.0, 0.0, 0.0), 1.0), "+y": ("wall", (0.03, 0.0, 0.0), 1.0)}
        return [{"shape": (150, 30, 6), "omega": 1.4, "solid": None, "faces": faces},
                {"shape": (70, 22, 5), "omega": 1.1, "solid": None, "faces": {}}]
    doms = []
    for k, (nx, ny) in enumerate([(200, 66), (131, 75), (97, 42)]):
        solid = np.zeros((ny, nx), dtype=np.uint8)
        if k == 1:
            solid[30:36, 60:70] = 1
        doms.append({"shape": (nx, ny), "omega": 1.0 + 0.3 * k, "solid": solid.reshape(-1),
                     "faces": W.channel_faces(0.03)})
    return doms


@pytest.mark.parametrize("dim,storage,tile", [(2, "f32", 32), (2, "f32", 64), (2, "f64", 32), (2, "f16", 64),
                                              (2, "bf16", 32), (3, "f32", 32), (3, "f32", 64), (2, "f32", 0),
                                              (2, "f64", 0), (2, "bf16", 0), (3, "f32", 0), (3, "f64", 0),
                                              (2, "f32", 16), (3, "f32", 16), (2, "f32", 1), (3, "f32", 1)])
def test_batch_interior_tiles_equal_single_domain(dim, storage, tile):
    """Interior tiles (their own kernel) and boundary tiles together give
    results bit-identical to the single-domain step, in every storage and in
    3D."""
    doms = interior_domains(dim)
    rng = np.random.default_rng(11)
    rhos, us = [], []
    for d in doms:
        n = int(np.prod(d["shape"]))
        rhos.append(1 + 0.01 * rng.standard_normal(n))
        us.append(0.02 * rng.standard_normal((n, dim)))
    for one in (1, 0):
        print(f"interior tiles {dim}D {storage} tile {tile} one launch {one}", flush=True)
        b = capi.Batch(doms, storage=storage)
        b.set_tiling(tile, 0)
        b.set_launch(one)
        b.init(np.concatenate(rhos), np.concatenate(us))
        singles = []
        for d, r, u in zip(doms, rhos, us):
            lat = capi.Lattice(d["shape"], storage=storage, omega=d["omega"], faces=d["faces"] or None,
                               solid=d["solid"])
            lat.init(r, u)
            singles.append(lat)
        for replay in range(2):
            b.run(8, 0.0, 8)
            for k, lat in enumerate(singles):
                lat.step(8)
                assert np.array_equal(b.populations(k), lat.populations()), (one, replay, k)
        for lat in singles:
            lat.close()
        b.close()


@pytest.mark.parametrize("tile,one", [(16, 1), (32, 1), (64, 1), (16, 0), (32, 0), (1, 1)])
def test_batch_waves_of_solid_cells(tile, one):
    """A tall solid block on the domain's corner: tiles whose waves hold
    only solid cells next to waves with fluid ones (such a wave does no
    update; a wave with solid and fluid lanes takes the flag path), equal to
    the single-domain step."""
    doms = []
    for k, (nx, ny) in enumerate([(96, 40), (80, 48)]):
        solid = np.zeros((ny, nx), dtype=np.uint8)
        solid[:13 + 4 * k, :40] = 1
        doms.append({"shape": (nx, ny), "omega": 1.3, "solid": solid.reshape(-1),
                     "faces": W.channel_faces(0.03)})
    rhos, us = random_init(doms, seed=4)
    print(f"solid waves tile {tile} one launch {one}", flush=True)
    b = capi.Batch(doms, storage="f32")
    b.set_tiling(tile, 0)
    b.set_launch(one)
    b.init(np.concatenate(rhos), np.concatenate(us))
    b.run(6, 0.0, 12)
    for k, d in enumerate(doms):
        lat = capi.Lattice(d["shape"], storage="f32", omega=d["omega"], faces=d["faces"], solid=d["solid"])
        lat.init(rhos[k], us[k])
        lat.step(12)
        assert np.array_equal(b.populations(k), lat.populations()), k
        lat.close()
    b.close()


def test_batch_against_oracle_f64():
    doms = small_domains()[:3]
    rhos, us = random_init(doms, seed=3)
    b = capi.Batch(doms, storage="f64")
    b.init(np.concatenate(rhos), np.concatenate(us))
    b.run(10, 0.0, 50)
    for k, d in enumerate(doms):
        rho, u = b.macros(k)
        p = O.Problem(d["shape"], faces=d["faces"], solid=d["solid"], omega=d["omega"])
        p.init(rhos[k], us[k])
        p.step(50)
        r_rho, r_u = p.macros()
        fl = ~p.solid
        eu = np.linalg.norm(u[fl] - r_u[fl]) / np.linalg.norm(r_u[fl])
        er = np.linalg.norm(rho[fl] - r_rho[fl]) / np.linalg.norm(r_rho[fl])
        assert eu < 1e-11 and er < 1e-11, (k, eu, er)
    b.close()


def test_batch_converges_and_retires_domains():
    doms = W.backward_step_batch(n=12, seed=7)
    for d in doms:  # shrink for a quick test
        pass
    small = []
    rng = np.random.default_rng(0)
    for k in range(12):
        L, H = int(rng.integers(40, 90)), int(rng.integers(16, 32))
        solid = np.zeros((H, L), dtype=np.uint8)
        solid[: H // 3, : L // 4] = 1
        small.append({"shape": (L, H), "omega": 1 / float(rng.uniform(0.7, 1.0)), "solid": solid.reshape(-1),
                      "faces": W.channel_faces(0.03)})
    rho, u = zip(*[W.cold_start(d, 0.03) for d in small])
    b = capi.Batch(small, storage="f32")
    b.init(np.concatenate(rho), np.concatenate(u))
    steps, res, mass, ms = b.run(200, 1e-6, 400000)
    print("converged steps", steps.tolist())
    assert np.all(steps > 0) and np.all(steps % 200 == 0)
    assert np.all(res < 1e-6)
    # a retired domain no longer changes
    f0 = b.populations(0)
    b.run(200, 1e-6, 2000)
    assert np.array_equal(f0, b.populations(0))
    b.close()


def test_autotuner_keeps_results():
    doms = small_domains()
    rhos, us = random_init(doms, seed=9)
    out = []
    for tune in (False, True):
        b = capi.Batch(doms, storage="f32")
        if tune:
            tx, blocks, one, ms = b.tune(4)
            assert tx in (0, 1, 16, 32, 64, 128, 256) and blocks >= 0 and one in (0, 1)
        b.init(np.concatenate(rhos), np.concatenate(us))
        b.run(4, 0.0, 12)
        out.append([b.populations(k) for k in range(len(doms))])
        b.close()
    for a, c in zip(out[0], out[1]):
        assert np.array_equal(a, c)

"""The drop-in host on the GPU: the reference's own test configs through the
`lbm` CLI, and the host's results against the FP64 oracle stepped on the
same reference-ordered grid with the reference's boundary semantics."""
import ctypes as ct
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tests.test_grid import assert_same_grid, host_grid

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LBM = os.path.join(ROOT, "lbm_amd", "bin", "lbm")
REFCFG = os.path.join(ROOT, "tests", "golden", "reference_configs")
OWNCFG = os.path.join(ROOT, "tests", "golden", "configs")


def run_cli(cfg, tmp_path, *args, timeout=300):
    p = tmp_path / "case.json"
    cfg = json.loads(json.dumps(cfg))
    cfg["solver"]["output_dir"] = str(tmp_path / "out")
    p.write_text(json.dumps(cfg))
    print("cli:", LBM, p, args, flush=True)
    r = subprocess.run([LBM, str(p), *args], cwd=tmp_path, capture_output=True, text=True, timeout=timeout)
    return r


@pytest.mark.parametrize("name", ["couette__couette.json", "couette__couette_bnd.json"])
def test_couette_reference_thresholds_f64(name, tmp_path):
    """test/couette/couette.json (periodic links) and couette_bnd.json
    (periodic boundary objects that copy, generateBndry = true): moving top
    wall (tangentialVelocity), reynoldsnumber-based omega; each must pass
    its own errorMax 4e-11 / errorL2 7e-10, unmodified."""
    cfg = json.load(open(os.path.join(REFCFG, name)))
    r = run_cli(cfg, tmp_path, "-p", "f64")
    print(r.stderr[-3000:])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "Reached convergence" in r.stderr
    assert "max. Error" in r.stderr


def test_poiseuille_periodic_pressure_reference_thresholds_f64(tmp_path):
    """test/poiseuille/poiseuille_bnd.json: periodic x with a density jump
    (pressure 1.0000008 / 1.0), bounce-back walls; must pass its own errorMax
    7.9e-6 / errorL2 2.1e-3 against poiseuille2D_1."""
    cfg = json.load(open(os.path.join(REFCFG, "poiseuille__poiseuille_bnd.json")))
    r = run_cli(cfg, tmp_path, "-p", "f64")
    print(r.stderr[-3000:])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "max. Error" in r.stderr


def test_cli_writes_reference_vtp_output(tmp_path):
    """LBMSolver::output (src/lbm/solver.cpp:194-195,323-382): a .vtp file at
    every solution_interval step and at the last step, holding the points and
    U, V, rho of every cell in cell-id order, the values after the 15-decimal
    text round trip; the file bytes equal the writer's restatement."""
    from oracle import vtk_oracle as V

    cfg = json.load(open(os.path.join(REFCFG, "couette__couette.json")))
    s = cfg["solver"]
    s.pop("analyticalSolution", None)
    s["maxSteps"], s["solution_interval"], s["convergence"] = 120, 50, 1e-300
    r = run_cli(cfg, tmp_path, "-p", "f64")
    assert r.returncode == 0, r.stderr[-2000:]
    name = s.get("solution_filename", "solution")
    files = sorted(os.listdir(tmp_path / "out"))
    assert files == sorted(f"{name}_{t}.vtp" for t in (50, 100, 119)), files
    (tmp_path / "again").mkdir()
    h, rho, u, steps = host_run(json.loads(json.dumps(cfg)), tmp_path / "again", "f64")
    data = (tmp_path / "out" / f"{name}_119.vtp").read_bytes()
    pts, vals, headers = V.read_points(data)
    assert np.array_equal(pts[:, :2], h["cen"].astype(np.float32))
    want = {"U": u[:, 0], "V": u[:, 1], "rho": rho}
    for k, v in want.items():
        assert np.array_equal(vals[k], np.array([V.text_round_trip(x) for x in v])), k
    assert data == V.write_points(2, h["cen"].tolist(), [(k, vals[k].tolist()) for k in ("U", "V", "rho")])


def test_cli_line_postprocessing(tmp_path):
    """The reference configs' "line" function at hook atEnd: line.csv in the
    working directory with the centres and the final u of the cells along
    the line (src/postprocess/postprocessing.h:88-116)."""
    from oracle import post_oracle as P
    from tests.test_postprocess import cell_length, host_line

    cfg = json.load(open(os.path.join(REFCFG, "poiseuille__poiseuille_bnd.json")))
    s = cfg["solver"]
    s.pop("analyticalSolution", None)
    s["maxSteps"], s["convergence"] = 300, 1e-300
    r = run_cli(cfg, tmp_path, "-p", "f64")
    assert r.returncode == 0, r.stderr[-2000:]
    assert "Executing postprocessing at hook:atEnd" in r.stderr
    (tmp_path / "again").mkdir()
    h, rho, u, steps = host_run(json.loads(json.dumps(cfg)), tmp_path / "again", "f64")
    f = list(s["postprocessing"].values())[0]
    cells = host_line(cfg, f["A"], f["B"])
    assert np.array_equal(cells, P.line_cells(h["cen"], cell_length(h["cen"]), f["A"], f["B"]))
    text = (tmp_path / "line.csv").read_text()
    assert text == P.line_csv(h["cen"], cells, u[:, 0])
    assert len(text.splitlines()) == len(cells) + 1 and len(cells) >= 32


POISEUILLE_WETNODE_REFCFG = ["poiseuille__poiseuille_bnd_eq.json", "poiseuille__poiseuille_bnd_NEEM.json",
                             "poiseuille__poiseuille_bnd_NEBB.json", "poiseuille__poiseuille_bnd_pressure.json",
                             "poiseuille__poiseuille_bnd_pressure_neem2.json"]


@pytest.mark.parametrize("name", POISEUILLE_WETNODE_REFCFG)
def test_poiseuille_wetnode_reference_thresholds_f64(name, tmp_path):
    """The reference's Poiseuille cases with wet-node walls and pressure
    faces, unmodified, through the CLI in FP64: each passes its own analytic
    error thresholds (exit status 0)."""
    cfg = json.load(open(os.path.join(REFCFG, name)))
    r = run_cli(cfg, tmp_path, "-p", "f64", timeout=600)
    tail = "\n".join(l for l in r.stderr.splitlines() if "Error" in l or "error" in l or "convergence" in l)
    print(name, r.returncode, tail[-1500:])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "max. Error" in r.stderr


def cli_report(stderr):
    """(last check iteration, errorMax, errorL2) printed by the CLI."""
    stop, emax, l2 = None, None, None
    for line in stderr.splitlines():
        head = line.split(":")[0]
        if head.isdigit() and " dU=" in line:
            stop = int(head)
        if line.startswith("max. Error:"):
            emax = float(line.split(":")[1])
        if line.startswith("avg. L2:"):
            l2 = float(line.split(":")[1])
    return stop, emax, l2


@pytest.mark.parametrize("name", ["couette__couette_bnd_eq.json", "couette__couette_bnd_eq2.json",
                                  "couette__couette_bnd_eq_aligned.json", "couette__couette_bnd_NEEM.json",
                                  "couette__couette_bnd_NEBB.json"])
def test_couette_wetnode_reference_cases_f64(name, tmp_path):
    """The reference's Couette cases with wet-node walls, unmodified, through
    the CLI in FP64, against the oracle run under the reference's own loop
    and stopping rule on the restated grid (tests/test_wetnode.py,
    reference_stop_rule): the check that stops the run comes at the same
    iteration, errorMax and errorL2 agree to the rounding noise of these
    15,000-step FP64 runs, and the run fails its thresholds as the
    restatement does.

    The errors measured here are the remaining transient, ~1e-12 on
    velocities of ~0.05, i.e. differences at the 1e-11 relative level; the
    oracle's unshifted FP64 populations (~0.1-0.4) carry ~1e-14 of
    accumulated rounding in u after 15,000 steps, so errorMax agrees to
    5e-14 absolute (a few % of its value) and errorL2 to 5 %.  With the
    reference's L2 formula the error profile, the slowest decaying mode,
    gives errorL2 about 0.48 x errorMax: 3.0e-13 where 1e-13 is asked (and
    6.7e-13 for couette_bnd_eq), so those four cases end with exit status
    255 through 'Analytical testcase failed', as the reference's own
    algorithm does on these inputs.  couette_bnd_NEEM ends on the exact
    profile up to rounding (3.46e-15 / 1.6e-15 in the restatement, which
    never meets its 1e-14 criterion and runs to maxSteps)."""
    from tests.test_wetnode import analytic_errors, reference_stop_rule

    cfg = json.load(open(os.path.join(REFCFG, name)))
    s = cfg["solver"]
    stop, (rho_o, u_o), g = reference_stop_rule(cfg)
    emax_o, l2_o = analytic_errors(cfg, g["cen"], u_o)
    r = run_cli(cfg, tmp_path, "-p", "f64", timeout=600)
    cstop, emax, l2 = cli_report(r.stderr)
    print(f"{name}: restated stop {stop} errorMax {emax_o:.6e} errorL2 {l2_o:.6e}; "
          f"CLI exit {r.returncode} last check {cstop} errorMax {emax} errorL2 {l2}")
    if emax_o < 1e-13:
        # couette_bnd_NEEM reaches the exact profile up to rounding.  Its
        # criterion, 1e-14 for sum |du| over 640 cells, sits at the rounding
        # floor: the unshifted FP64 restatement (like the reference) never
        # meets it and runs to maxSteps, the shifted FP64 device populations
        # round less and may meet it earlier; either way both errors are
        # rounding noise, and passing 3.46e-15 / 1e-15 is decided by it
        assert emax < 1e-13 and l2 < 1e-13
        return
    assert "Reached convergence" in r.stderr and cstop == stop
    assert abs(emax - emax_o) <= 5e-14 and abs(l2 - l2_o) <= 0.05 * l2_o
    # errorL2 fails by a factor near 3 in both: far outside the rounding noise
    assert l2_o > 2 * s["errorL2"] and l2 > 2 * s["errorL2"]
    assert r.returncode == 255, r.stderr[-2000:]
    assert "Analytical testcase failed" in r.stderr


def test_couette_reference_f32_fails_loudly_only_on_thresholds(tmp_path):
    """FP32 storage cannot reach the 4e-11 bound: the run must end through the
    analytic check with status 255, like the reference's TERMM."""
    cfg = json.load(open(os.path.join(REFCFG, "couette__couette.json")))
    cfg["solver"]["maxSteps"] = 3000
    r = run_cli(cfg, tmp_path, "-p", "f32")
    assert r.returncode == 255
    assert "Analytical testcase failed" in r.stderr


def host_run(cfg, tmp_path, storage):
    L = ct.CDLL(os.path.join(ROOT, "lbm_amd", "liblbm_host.so"))
    L.lbm_host_run_file.argtypes = [ct.c_char_p, ct.c_char_p, ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_void_p]
    p = tmp_path / "case.json"
    cfg["solver"]["output_dir"] = str(tmp_path / "out")
    p.write_text(json.dumps(cfg))
    h = host_grid(cfg)
    n, dim = h["n"], h["dim"]
    rho = np.zeros(n)
    u = np.zeros(n * dim)
    steps = ct.c_int64()
    cwd = os.getcwd()
    os.chdir(tmp_path)  # postprocessing writes line.csv into the working directory
    try:
        rc = L.lbm_host_run_file(str(p).encode(), storage.encode(), n, rho.ctypes.data, u.ctypes.data, ct.byref(steps))
    finally:
        os.chdir(cwd)
    assert rc == 0
    return h, rho, u.reshape(n, dim), steps.value


def oracle_run(cfg, h, steps, textbook=False):
    bnd = cfg["solver"]["boundary"]
    order = []
    for g in sorted(bnd):
        nb = len(bnd[g])
        for f in sorted(bnd[g]):
            order.append((f"{g}_{f}" if nb > 1 else g, bnd[g][f]))
    s = cfg["solver"]
    if "relaxation" in s:
        omega = 1 / s["relaxation"]
    else:
        nu = s["ma"] / s["reynoldsnumber"] * s.get("refLength", 1.0)
        omega = 2.0 / (1.0 + 2.0 * nu * 2.0 ** cfg["partitionLevel"])
    nb = h["nb"]
    if textbook or any(v.get("type") == "periodic" and v.get("generateBndry", True) for _, v in order):
        # symmetric neighbours of the structured window (true diagonal cells)
        nb = wrapped_neighbours(h, cfg)
    nx, ny, nz = h["nxyz"]
    sidx = h["sidx"]
    coords = np.stack([sidx % nx, (sidx // nx) % ny, sidx // (nx * ny)], 1)
    p = O.GridProblem(h["dim"], nb, h["surf"], order, omega, force=s.get("bodyForce", (0, 0, 0)),
                      textbook=textbook, coords=coords)
    p.init_reference()
    p.step(steps)
    return p.macros()


def wrapped_neighbours(h, cfg):
    """Per level on a multi-level grid (each level is its own lattice)."""
    lv = h.get("level")
    if lv is None or len(set(lv.tolist())) == 1:
        return wrapped_neighbours_window(h, cfg)
    nb = np.full_like(h["nb"], -1)
    for l in sorted(set(lv.tolist())):
        ids = np.nonzero(lv == l)[0]
        loc = -np.ones(h["n"], dtype=np.int64)
        loc[ids] = np.arange(ids.size)
        cen = h["cen"][ids]
        nxyz = np.ones(3, dtype=np.int64)
        idx = np.zeros((ids.size, 3), dtype=np.int64)
        for a in range(h["dim"]):
            vals, inv = np.unique(cen[:, a], return_inverse=True)
            nxyz[a] = vals.size
            idx[:, a] = inv
        sub = dict(n=ids.size, dim=h["dim"], nb=h["nb"][ids], nxyz=nxyz,
                   sidx=idx[:, 0] + nxyz[0] * (idx[:, 1] + nxyz[1] * idx[:, 2]))
        w = wrapped_neighbours_window(sub, cfg)
        nb[ids] = np.where(w >= 0, ids[np.maximum(w, 0)], -1)
    return nb


def wrapped_neighbours_window(h, cfg):
    """'periodic' boundary objects (generateBndry = true) copy outgoing
    populations to the linked cell of the other face (bnd_periodic.h:102-119),
    i.e. the structured window wraps along that axis."""
    nx, ny, nz = h["nxyz"]
    dim = h["dim"]
    ext = [nx, ny, nz]
    sidx = h["sidx"]
    cell_at = -np.ones(nx * ny * nz, dtype=np.int64)
    cell_at[sidx] = np.arange(h["n"])
    co = np.stack([sidx % nx, (sidx // nx) % ny, sidx // (nx * ny)], 1)
    per = [False, False, False]
    for g, faces in cfg["solver"]["boundary"].items():
        for f, v in faces.items():
            # a periodic pair with a pressure jump is not a wrap
            if v.get("type") == "periodic" and "pressure" not in v:
                per["xyz".index(f[1])] = True
    from oracle import grid_oracle as G
    dv = G.DIR_VEC2 if dim == 2 else G.DIR_VEC3
    nb = np.full_like(h["nb"], -1)
    for d in range(h["nb"].shape[1]):
        v = list(dv[d]) + [0] * (3 - len(dv[d]))
        t = co + np.array(v)
        ok = np.ones(h["n"], dtype=bool)
        for a in range(3):
            if per[a]:
                t[:, a] %= ext[a]
            else:
                ok &= (t[:, a] >= 0) & (t[:, a] < ext[a])
        s = np.where(ok, t[:, 0] + nx * (t[:, 1] + ny * t[:, 2]), 0)
        nb[:, d] = np.where(ok, cell_at[s], -1)
    return nb


CASES = [
    ("couette__couette.json", REFCFG, 400, False, None),
    ("couette__couette_bnd_bbDirichlet.json", REFCFG, 400, False, None),
    ("pressure_channel2d.json", OWNCFG, 400, False, None),
    ("box3d_periodic_x.json", OWNCFG, 200, False, None),
    ("sphere3d_channel.json", OWNCFG, 150, True, None),
    ("neem_channel2d.json", OWNCFG, 400, False, None),
    ("sphere3d_neem.json", OWNCFG, 150, True, None),
    ("poiseuille__poiseuille_bnd.json", REFCFG, 300, False, None),
    ("poiseuille__poiseuille_bnd.json", REFCFG, 300, False, 1.002),
    ("couette__couette_bnd_eq.json", REFCFG, 300, False, None),
    ("couette__couette_bnd_eq2.json", REFCFG, 300, False, None),
    ("couette__couette_bnd_eq_aligned.json", REFCFG, 300, False, None),
    ("couette__couette_bnd_NEEM.json", REFCFG, 300, False, None),
    ("couette__couette_bnd_NEBB.json", REFCFG, 300, False, None),
    ("poiseuille__poiseuille_bnd_eq.json", REFCFG, 300, False, None),
    ("poiseuille__poiseuille_bnd_eq.json", REFCFG, 300, False, 1.002),
    ("poiseuille__poiseuille_bnd_NEEM.json", REFCFG, 300, False, None),
    ("poiseuille__poiseuille_bnd_NEEM.json", REFCFG, 300, False, 1.002),
    ("poiseuille__poiseuille_bnd_NEBB.json", REFCFG, 300, False, None),
    ("poiseuille__poiseuille_bnd_NEBB.json", REFCFG, 300, False, 1.002),
    ("poiseuille__poiseuille_bnd_pressure.json", REFCFG, 300, False, None),
    ("poiseuille__poiseuille_bnd_pressure.json", REFCFG, 300, False, 1.002),
]


_GRIDS = {}


def restated_grid(folder, name):
    from tests.test_grid import oracle_grid

    if (folder, name) not in _GRIDS:
        _GRIDS[folder, name] = oracle_grid(json.load(open(os.path.join(folder, name))))
    return _GRIDS[folder, name]


@pytest.mark.parametrize("name,folder,steps,textbook,jump", CASES)
@pytest.mark.parametrize("storage,tol", [("f64", 1e-11), ("f32", 1e-5)])
def test_host_matches_oracle_on_reference_grid(name, folder, steps, textbook, jump, storage, tol, tmp_path):
    """The host on the reference grid against the oracle stepped on the
    restated grid with the reference's boundary semantics, maxSteps
    iterations with the thresholds and the stop criterion removed.

    The reference's Poiseuille cases drive velocities of ~1e-7 with density
    jumps of ~1e-6 between their pressure faces; there the oracle's
    unshifted FP64 populations (~0.1, rounding ~1e-17 per operation) hold u
    to ~1e-9 relative per step, so the FP64 bar of those unmodified cases is
    1e-8 after their 300 steps.  The same cases with a 2e-3 jump
    (velocities ~2e-4, the jump column) are checked at the 1e-11 bar."""
    cfg = json.load(open(os.path.join(folder, name)))
    jumps = [abs(v.get("pressure", 1.0) - 1.0) for faces in cfg["solver"]["boundary"].values() for v in faces.values()]
    if storage == "f64" and jump is None and 0 < max(jumps) < 1e-4:
        tol = 1e-8
    s = cfg["solver"]
    s["maxSteps"] = steps
    s["convergence"] = 0.0
    for k in ("analyticalSolution", "errorL2", "errorMax", "errorGRE", "analyticalSolutionExcludeSurface",
              "postprocessing", "poiseuillePressureGradient"):
        s.pop(k, None)
    if jump is not None:
        for spec in s["boundary"].get("cube", {}).values():
            if spec.get("pressure", 1.0) != 1.0:
                spec["pressure"] = jump
    h, rho, u, done = host_run(cfg, tmp_path, storage)
    assert done == steps
    # grid indexing and neighbour maps: the host's grid is the restated
    # generator's, bit for bit; the oracle steps the restated grid
    g = restated_grid(folder, name)
    assert_same_grid(h, g)
    # the host returns the macros after its last step, as the reference's
    # forced output at the last iteration recomputes them
    r_rho, r_u = oracle_run(cfg, g, steps, textbook)
    eu = np.linalg.norm(u - r_u) / np.linalg.norm(r_u)
    er = np.linalg.norm(rho - r_rho) / np.linalg.norm(r_rho)
    print(f"{name} {storage}: rel-L2 u={eu:.3e} rho={er:.3e}")
    assert eu < tol and er < tol


MULTI_CASES = [
    ("pressure_channel2d.json", OWNCFG, (3, 5, 5), 300, False),
    ("pressure_channel2d.json", OWNCFG, (4, 5, 5), 400, False),
    ("neem_channel2d.json", OWNCFG, (4, 5, 5), 400, False),
    ("couette__couette_bnd_NEEM.json", REFCFG, (4, 5, 5), 300, False),
    ("sphere3d_neem.json", OWNCFG, (4, 5, 5), 100, True),
]


@pytest.mark.parametrize("name,folder,levels,steps,textbook", MULTI_CASES)
def test_host_multilevel_matches_oracle(name, folder, levels, steps, textbook, tmp_path):
    """partitionLevel < uniformLevel: the reference keeps every level in the
    grid and its solver loops over all cells, so each level is a lattice of
    its own (links never cross levels).  The host runs one device lattice
    per level in lockstep (convergence sums added over the levels); its
    fields after maxSteps iterations equal the oracle stepped on the
    restated multi-level grid (one neighbour table over all cells) in FP64."""
    cfg = json.load(open(os.path.join(folder, name)))
    cfg["partitionLevel"], cfg["uniformLevel"], cfg["maxRfnmtLvl"] = levels
    cfg["maxNoCells"] = max(cfg["maxNoCells"], 4_000_000)
    s = cfg["solver"]
    s["maxSteps"] = steps
    s["convergence"] = 0.0
    for k in ("analyticalSolution", "errorL2", "errorMax", "errorGRE", "analyticalSolutionExcludeSurface",
              "postprocessing", "poiseuillePressureGradient"):
        s.pop(k, None)
    print(f"multi-level {name} levels {levels}", flush=True)
    h, rho, u, done = host_run(cfg, tmp_path, "f64")
    assert done == steps
    from tests.test_grid import oracle_grid
    g = oracle_grid(cfg)
    assert h["n"] == g["n"] and len(set(g["level"].tolist())) == levels[1] - levels[0] + 1
    assert np.array_equal(h["cen"], g["cen"]) and np.array_equal(h["nb"], g["nb"])
    assert np.array_equal(h["bnd"].astype(bool), g["bnd"].astype(bool)) and h["surf"] == g["surf"]
    r_rho, r_u = oracle_run(cfg, g, steps, textbook)
    eu = np.linalg.norm(u - r_u) / np.linalg.norm(r_u)
    er = np.linalg.norm(rho - r_rho) / np.linalg.norm(r_rho)
    print(f"{name} levels {levels}: rel-L2 u={eu:.3e} rho={er:.3e}")
    assert eu < 1e-10 and er < 1e-10


def test_cli_refuses_boundary_refinement(tmp_path):
    """maxRfnmtLvl > uniformLevel: grids are generated (and checked against
    the restatement on the CPU), the solver ends with status 255."""
    cfg = json.load(open(os.path.join(REFCFG, "couette__couette_bnd_NEEM.json")))
    cfg["partitionLevel"], cfg["uniformLevel"], cfg["maxRfnmtLvl"] = 4, 4, 5
    cfg["solver"]["maxSteps"] = 10
    r = run_cli(cfg, tmp_path, "-p", "f64", timeout=120)
    assert r.returncode == 255
    assert "Boundary refinement" in r.stderr

"""The reference's "forcing" pressure correction (LBMSolver::forcing,
src/lbm/solver.cpp:626-696), run by its own test/poiseuille/poiseuille.json.

After every collision the cells of the inlet face cube_-x take
feq(p_in, u_v) + f(v) - feq(v) from the cell v next to the outlet face in
their row, and the outlet cells cube_+x the same with p_out = 1 from the
cell next to the inlet face; feq(p, u_v) puts only u_v's x component into
c.u.  Pins of the oracle's restatement (CPU):
  * what the rule fixes: a forced cell carries density p (1 - 3/2 u_y(v)^2)
    (the x-only c.u drops u_y from the second moment) and momentum
    (p u_x(v), 0) after the correction (the collision keeps v's moments);
  * the reference's own run of poiseuille.json (recorded when the
    reference was built and run for this project: max. Error 1.26064e-07,
    avg. L2 3.23436e-05, both within its thresholds) is reproduced by the
    oracle under the reference's loop and stop rule.
On the GPU the host CLI runs the unmodified configuration and must print
the same errors, and the host's fields after a fixed number of steps equal
the oracle's."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFCFG = os.path.join(ROOT, "tests", "golden", "reference_configs")
CFG = "poiseuille__poiseuille.json"
REF_ERROR_MAX, REF_ERROR_L2 = 1.26064e-07, 3.23436e-05  # the reference's run, 6 printed digits


def load():
    return json.load(open(os.path.join(REFCFG, CFG)))


def omega_of(cfg):
    """loadConfiguration, reynoldsnumber branch (src/lbm/solver.cpp:114-122)."""
    s = cfg["solver"]
    nu = s["ma"] / s["reynoldsnumber"] * s.get("refLength", 1.0)
    return 2.0 / (1.0 + 2.0 * nu * 2.0 ** cfg["maxRfnmtLvl"]), nu


def forcing_pairs(cfg, g):
    """(face cell, value cell, pressure) in the reference's loop order: the
    outlet cells from the inlet cells' +x neighbours, then the inlet cells
    from the outlet cells' -x neighbours, matched by the row's y centre."""
    s = cfg["solver"]
    p_out = 1.0
    p_in = p_out + s["poiseuillePressureGradient"]
    inlet, outlet = g["surf"]["cube_-x"][0], g["surf"]["cube_+x"][0]
    cen, nb = g["cen"], g["nb"]
    cells, vals, ps = [], [], []
    for i in inlet:
        v = nb[i, 1]
        for o in outlet:
            if abs(cen[v, 1] - cen[o, 1]) < np.finfo(float).eps:
                cells.append(o), vals.append(v), ps.append(p_out)
    for o in outlet:
        v = nb[o, 0]
        for i in inlet:
            if abs(cen[i, 1] - cen[v, 1]) < np.finfo(float).eps:
                cells.append(i), vals.append(v), ps.append(p_in)
    return cells, vals, ps


def problem(cfg):
    from tests.test_grid import oracle_grid

    g = oracle_grid(cfg)
    s = cfg["solver"]
    bnd = s["boundary"]
    order = [(f"{k}_{f}" if len(bnd[k]) > 1 else k, bnd[k][f]) for k in sorted(bnd) for f in sorted(bnd[k])]
    om, nu = omega_of(cfg)
    p = O.GridProblem(g["dim"], g["nb"], g["surf"], order, om)
    p.set_forcing(*forcing_pairs(cfg, g))
    p.init_reference()
    return p, g, nu


def analytic_errors(cfg, g, u, nu):
    """compareToAnalyticalResult for poiseuille2D_1 (src/lbm/analytical_solutions.h:36-59,
    solver.cpp:388-482), the analyticalSolutionExcludeSurface cells left out."""
    s = cfg["solver"]
    dp = s["poiseuillePressureGradient"]
    excl = set()
    for name in s.get("analyticalSolutionExcludeSurface", []):
        excl.update(g["surf"][name][0])
    keep = np.array([q for q in range(g["n"]) if q not in excl])
    y = g["cen"][keep, 1]
    a = np.zeros((keep.size, 2))
    a[:, 0] = dp / (2.0 * nu) * y * (1.0 - 0.0 - y)
    d = np.sqrt(((u[keep] - a) ** 2).sum(1))
    L2 = np.sqrt((d * d).sum()) / np.sqrt((a * a).sum()) / g["n"] ** 0.5
    return d.max(), L2


def test_forcing_pairs_cover_both_faces_row_by_row():
    cfg = load()
    from tests.test_grid import oracle_grid

    g = oracle_grid(cfg)
    cells, vals, ps = forcing_pairs(cfg, g)
    inlet, outlet = g["surf"]["cube_-x"][0], g["surf"]["cube_+x"][0]
    assert sorted(cells) == sorted(list(inlet) + list(outlet)) and len(cells) == len(set(cells))
    xs = np.unique(g["cen"][:, 0])
    for c, v in zip(cells, vals):
        assert abs(g["cen"][c, 1] - g["cen"][v, 1]) < 1e-15
        assert g["cen"][v, 0] in (xs[1], xs[-2]) and g["cen"][c, 0] in (xs[0], xs[-1])
        assert (g["cen"][c, 0] == xs[0]) == (g["cen"][v, 0] == xs[-2])


def test_forcing_rule_moments_after_collision():
    """One step from a perturbed state: each forced cell's post-collision
    populations carry density p (1 - 3/2 u_y(v)^2) and momentum
    (p u_x(v), 0): sum_i w_i (1 + 3 c_ix u_x + 9/2 c_ix^2 u_x^2 - 3/2 |u|^2)
    = 1 - 3/2 u_y^2 on D2Q9, and f(v) - feq(v) has no mass or momentum."""
    cfg = load()
    p, g, _ = problem(cfg)
    rng = np.random.default_rng(4)
    n = g["n"]
    rho = 1.0 + 1e-3 * rng.standard_normal(n)
    u = 1e-3 * rng.standard_normal((n, 2))
    p.init(rho, u)
    r0, u0 = p.macros()  # the step's macros (no force: the collision keeps them)
    p.step(1)
    post = p.post_collision()  # [Q][n]
    c, _, _ = O.tables(9)
    cells, vals, ps = p.forcing
    for k in range(len(cells)):
        f = post[:, cells[k]]
        v = vals[k]
        assert abs(f.sum() - ps[k] * (1.0 - 1.5 * u0[v, 1] ** 2)) < 1e-14
        assert abs((c[:, 0] * f).sum() - ps[k] * u0[v, 0]) < 1e-14
        assert abs((c[:, 1] * f).sum()) < 1e-14


def test_reference_poiseuille_run_is_reproduced():
    """poiseuille.json under the reference's loop (checks every
    info_interval, vars of steps t-1 and t-2, one more step after
    convergence, output macros after it) on the oracle: the analytic errors
    equal the reference's own run to its 6 printed digits."""
    cfg = load()
    s = cfg["solver"]
    p, g, nu = problem(cfg)
    interval = s.get("conv_interval", s.get("info_interval", 100))
    t, stop = 0, None
    while t + interval <= s["maxSteps"]:
        nxt = t + interval
        p.step(nxt - 2 - t)
        r2, u2 = p.macros()
        p.step(1)
        r1, u1 = p.macros()
        p.step(1)
        t = nxt
        conv = max(np.abs(u1[:, 0] - u2[:, 0]).sum(), np.abs(u1[:, 1] - u2[:, 1]).sum(), np.abs(r1 - r2).sum())
        if conv < s["convergence"]:
            p.step(1)
            stop = t
            break
    if stop is None:
        p2, _, _ = problem(cfg)
        p2.step(s["maxSteps"])
        p = p2
    rho, u = p.macros()
    emax, l2 = analytic_errors(cfg, g, u, nu)
    print(f"poiseuille.json: stop {stop}, max. Error {emax:.6g}, avg. L2 {l2:.6g}")
    assert emax < s["errorMax"] and l2 < s["errorL2"]
    assert abs(emax - REF_ERROR_MAX) <= 1e-5 * REF_ERROR_MAX
    assert abs(l2 - REF_ERROR_L2) <= 1e-5 * REF_ERROR_L2


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
def test_device_forcing_matches_oracle(tmp_path):
    """The host (lbm_set_pressure_forcing on the device) on poiseuille.json
    for 600 iterations, the stop criterion and analytic check removed,
    against the oracle on the restated grid: FP64 fields to 1e-8 (the
    velocities are ~1e-6, driven by a 8e-7 pressure difference)."""
    from tests.test_host_gpu import host_run

    cfg = load()
    s = cfg["solver"]
    steps = 600
    s["maxSteps"] = steps
    s["convergence"] = 0.0
    for k in ("analyticalSolution", "errorL2", "errorMax", "analyticalSolutionExcludeSurface", "postprocessing"):
        s.pop(k, None)
    h, rho, u, done = host_run(cfg, tmp_path, "f64")
    assert done == steps
    p, g, _ = problem(load())
    assert np.array_equal(h["cen"], g["cen"])
    p.step(steps)
    r_rho, r_u = p.macros()
    eu = np.linalg.norm(u - r_u) / np.linalg.norm(r_u)
    er = np.linalg.norm(rho - r_rho) / np.linalg.norm(r_rho)
    print(f"forcing, {steps} steps: rel-L2 u={eu:.3e} rho={er:.3e}")
    assert eu < 1e-8 and er < 1e-12


@pytest.mark.gpu
def test_reference_poiseuille_case_through_cli(tmp_path):
    """test/poiseuille/poiseuille.json unmodified through the CLI in FP64:
    exit 0 with the analytic errors of the reference's own run."""
    from tests.test_host_gpu import run_cli

    cfg = load()
    r = run_cli(cfg, tmp_path, "-p", "f64", timeout=300)
    print(r.stderr[-1500:])
    assert r.returncode == 0, r.stderr[-2000:]
    emax = l2 = None
    for line in r.stderr.splitlines():
        if line.startswith("max. Error:"):
            emax = float(line.split(":")[1])
        if line.startswith("avg. L2:"):
            l2 = float(line.split(":")[1])
    assert abs(emax - REF_ERROR_MAX) <= 1e-5 * REF_ERROR_MAX
    assert abs(l2 - REF_ERROR_L2) <= 1e-5 * REF_ERROR_L2


@pytest.mark.gpu
def test_forcing_argument_checks():
    from lbm_amd import capi

    lat = capi.Lattice((16, 8, 1), storage="f64", omega=1.0)
    try:
        lat.set_pressure_forcing(1, 1.001, 1.0)
        lat.set_pressure_forcing(0)
        with pytest.raises(Exception):
            lat.set_pressure_forcing(1, -1.0, 1.0)
    finally:
        lat.close()
    lat3 = capi.Lattice((8, 8, 8), storage="f32", omega=1.0)
    try:
        with pytest.raises(Exception):
            lat3.set_pressure_forcing(1, 1.001, 1.0)  # D3Q19
    finally:
        lat3.close()

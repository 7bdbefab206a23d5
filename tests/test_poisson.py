"""The Poisson (screened Poisson / diffusion-reaction) solver: the FP64
restatement oracle/poisson_oracle.py pinned by what the mathematics fixes,
the host's 1D grids and boundary expressions, and on the GPU the device
step (include/lbm_poisson.h) against the restatement and the reference's own
test/poisson cases through the CLI with their thresholds.

Pins of the restatement (CPU):
  * tables: weights and Poisson weights sum to one, zero first moment; the
    reference's values (src/lbm/constants.h:248-319) as data;
  * eq::poisson: sum_{i<Q-1} feq_i / (1 - w_rest) = phi and sum feq = 0;
  * the Dirichlet NEEM rule after one step of an arbitrary state: the
    boundary cell carries exactly the prescribed phi and the
    non-equilibrium part of its extrapolation cell;
  * without reaction the steady state of a Dirichlet problem whose values
    are linear in the coordinates is that linear field (1D and 2D, D2Q5 and
    D2Q9): a harmonic function the diffusion scheme reproduces exactly;
  * with reaction the steady state converges at second order to the closed
    forms the reference compares with: CHAI08 example 1 (Debye-Hueckel,
    Dirichlet both ends) and the catalyst slab cosh(1 - x) / cosh(1)
    (Dirichlet and zero-gradient Neumann).
"""
import json
import os

import numpy as np
import pytest

from oracle import poisson_oracle as PO

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PCFG = os.path.join(ROOT, "tests", "golden", "reference_configs")


# ------------------------------------------------------------ structured problems
def structured_grid(nx, ny):
    """Neighbour table (main + 2D diagonals, cartesian order) of an nx x ny
    window in x-fastest cell order, and the face surfaces in the reference's
    form (cells with a missing neighbour in that direction, cell order)."""
    dim = 1 if ny == 1 else 2
    n = nx * ny
    stride = 2 if dim == 1 else 8
    vec = [(-1, 0), (1, 0), (0, -1), (0, 1), (1, 1), (1, -1), (-1, -1), (-1, 1)][:stride]
    nb = -np.ones((n, stride), dtype=np.int64)
    for c in range(n):
        x, y = c % nx, c // nx
        for d, (dx, dy) in enumerate(vec):
            xx, yy = x + dx, y + dy
            if 0 <= xx < nx and 0 <= yy < ny:
                nb[c, d] = xx + nx * yy
    names = ["-x", "+x", "-y", "+y"][:2 * dim]
    surf = {nm: ([c for c in range(n) if nb[c, d] < 0], None) for d, nm in enumerate(names)}
    return nb, surf


def along(face, cells, nx):
    """Position of each surface cell along its face (y on x faces, x on y faces)."""
    return [c // nx if face in ("-x", "+x") else c % nx for c in cells]


def structured_problem(Q, nx, ny, faces, tau, source):
    """faces: {name: (kind, values along the face or None)} in application order."""
    nb, surf = structured_grid(nx, ny)
    bcs = []
    for name, (kind, vals) in faces.items():
        cells = surf[name][0]
        v = [0.0] * len(cells) if vals is None else [vals[k] for k in along(name, cells, nx)]
        bcs.append((name, kind, v))
    return PO.PoissonGridProblem(Q, nb, surf, bcs, tau, source), nb, surf


def converge(p, tol=1e-13, every=200, cap=400000):
    while p.steps < cap:
        p.step(every)
        if p.residual() < tol:
            break
    return p


# -------------------------------------------------------------------- the pins
@pytest.mark.parametrize("Q", [3, 5, 9])
def test_tables(Q):
    t = PO.TABLES[Q]
    w, wp, c = np.array(t["w"]), np.array(t["wp"]), np.array(t["c"], dtype=float)
    assert abs(w.sum() - 1) < 1e-15 and abs(wp.sum() - 1) < 1e-15 and wp[-1] == 0
    assert np.abs((w[:, None] * c).sum(0)).max() < 1e-15
    assert np.all(c[-1] == 0)
    ref = {3: ([1 / 6, 1 / 6, 2 / 3], 1 / 3), 5: ([1 / 6] * 4 + [1 / 3], 1 / 2), 9: ([1 / 9] * 4 + [1 / 36] * 4 + [4 / 9], 1 / 3)}
    assert np.allclose(w, ref[Q][0], rtol=0, atol=1e-16) and t["alpha"] == ref[Q][1]


@pytest.mark.parametrize("Q", [3, 5, 9])
def test_equilibrium_moments(Q):
    nx, ny = (9, 1) if Q == 3 else (7, 6)
    p, _, _ = structured_problem(Q, nx, ny, {}, 1.0, 0.0)
    phi = np.random.default_rng(1).standard_normal(nx * ny)
    f = p.feq(phi)
    assert np.abs(f[:, :-1].sum(1) / (1 - p.w[-1]) - phi).max() < 1e-14
    assert np.abs(f.sum(1)).max() < 1e-14


@pytest.mark.parametrize("Q", [3, 5, 9])
def test_dirichlet_neem_rule_one_step(Q):
    nx, ny = (12, 1) if Q == 3 else (9, 8)
    rng = np.random.default_rng(Q)
    faces = {"-x": ("dirichlet", rng.random(ny)), "+x": ("dirichlet", rng.random(ny))}
    if Q != 3:
        faces["-y"] = ("dirichlet", rng.random(nx))
        faces["+y"] = ("dirichlet", rng.random(nx))
    p, nb, surf = structured_problem(Q, nx, ny, faces, 0.8, -0.01)
    p.init()
    p.fold = p.fold + 0.05 * rng.standard_normal(p.fold.shape)  # arbitrary, not an equilibrium
    p.step(1)
    # the object applied last owns a corner cell
    owner = {}
    for b in p.bcs:
        for k, c in enumerate(b["cells"]):
            owner[c] = (b, k)
    for c, (b, k) in owner.items():
        v, e = b["value"][k], b["e"][k]
        phi_c = p.fold[c, :-1].sum() / (1 - p.w[-1])
        phi_e = p.fold[e, :-1].sum() / (1 - p.w[-1])
        assert abs(phi_c - v) < 1e-14
        assert np.abs((p.fold[c] - p.feq(np.array([v]))[0]) - (p.fold[e] - p.feq(np.array([phi_e]))[0])).max() < 1e-14


@pytest.mark.parametrize("Q,nx,ny", [(3, 24, 1), (5, 14, 11), (9, 14, 11)])
def test_linear_field_is_the_steady_state_without_reaction(Q, nx, ny):
    a, bx, by = 0.7, 0.3, -0.45

    def lin(x, y):
        return a + bx * x + by * y

    faces = {"-x": ("dirichlet", [lin(0, y) for y in range(ny)]),
             "+x": ("dirichlet", [lin(nx - 1, y) for y in range(ny)])}
    if ny > 1:
        faces["-y"] = ("dirichlet", [lin(x, 0) for x in range(nx)])
        faces["+y"] = ("dirichlet", [lin(x, ny - 1) for x in range(nx)])
    p, _, _ = structured_problem(Q, nx, ny, faces, 1.0, 0.0)
    p.init()
    converge(p, 1e-14)
    xs, ys = np.arange(nx * ny) % nx, np.arange(nx * ny) // nx
    assert np.abs(p.vars - lin(xs, ys)).max() < 1e-11


def chai08_1(x):
    """CHAI08 example 1 (src/lbm/analytical_solutions.h:72-81)."""
    k = PO.DEBYE_HUCKEL_KAPPA
    return ((np.exp(k) - 1) / (np.exp(k) - np.exp(-k)) * np.exp(-k * x)
            + (1 - np.exp(-k)) / (np.exp(k) - np.exp(-k)) * np.exp(k * x))


def test_debye_huckel_second_order():
    """1D D1Q3, phi = 1 at both ends, kappa = 27.79; nodes mapped onto
    equidistant points of [0, 1] as the reference's shiftCenter does."""
    errs = []
    for n in (64, 128, 256):
        dt = 1.0 / (n - 1)
        src = PO.source_coefficient(3, 1.0, dt, PO.DEBYE_HUCKEL_KAPPA)
        p, _, _ = structured_problem(3, n, 1, {"+x": ("dirichlet", [1.0]), "-x": ("dirichlet", [1.0])}, 1.0, src)
        p.init()
        converge(p, 1e-13, cap=200000)
        x = np.arange(n) / (n - 1)
        errs.append(np.abs(p.vars - chai08_1(x)).max())
    order = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    print("Debye-Hueckel max errors", errs, "orders", order)
    assert errs[-1] < 2e-3 and np.all(order > 1.8)


def test_zero_gradient_neumann_catalyst_slab_second_order():
    """1D D1Q3, phi = 1 at x = 0 (Dirichlet), zero gradient at x = 1
    (Neumann NEEM), kappa = 1: cosh(1 - x) / cosh(1) (the reference's
    poissonSimpleDiffReaction)."""
    errs = []
    for n in (16, 32, 64):
        dt = 1.0 / (n - 1)
        src = PO.source_coefficient(3, 1.0, dt, 1.0)
        p, _, _ = structured_problem(3, n, 1, {"+x": ("neumann", None), "-x": ("dirichlet", [1.0])}, 1.0, src)
        p.init()
        converge(p, 1e-15, every=500, cap=2000000)
        x = np.arange(n) / (n - 1)
        errs.append(np.abs(p.vars - np.cosh(1 - x) / np.cosh(1)).max())
    order = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    print("catalyst slab max errors", errs, "orders", order)
    assert errs[-1] < 1e-3 and np.all(order > 1.5)


# ------------------------------------------------------- host: grids, values
def host_lib():
    import ctypes as ct

    L = ct.CDLL(os.path.join(ROOT, "lbm_amd", "liblbm_host.so"))
    L.lbm_host_eval_expression.argtypes = [ct.c_char_p, ct.c_double, ct.c_double, ct.c_double,
                                           ct.POINTER(ct.c_double)]
    return L


@pytest.mark.parametrize("text,py", [
    ("cos(pi*x)", lambda x, y: np.cos(np.pi * x)),
    ("-sinh(sqrt(4+pi^2)*(1-y))/sinh(sqrt(4+pi^2))", lambda x, y: -np.sinh(np.sqrt(4 + np.pi ** 2) * (1 - y)) / np.sinh(np.sqrt(4 + np.pi ** 2))),
    ("2^3^2 - -x*3 + (y)", lambda x, y: 2 ** 9 + x * 3 + y),
    ("-x^2 + exp(log(y+2))", lambda x, y: -(x ** 2) + (y + 2)),
    ("1", lambda x, y: 1.0),
])
def test_boundary_expressions(text, py):
    import ctypes as ct

    L = host_lib()
    for x, y in [(0.0, 0.0), (0.3, 0.7), (1.0, 0.25)]:
        out = ct.c_double()
        assert L.lbm_host_eval_expression(text.encode(), x, y, 0.0, ct.byref(out)) == 0
        assert abs(out.value - py(x, y)) <= 1e-14 * max(1.0, abs(py(x, y)))


def test_boundary_expression_errors():
    import ctypes as ct

    L = host_lib()
    out = ct.c_double()
    for bad in ["cos(", "foo(x)", "x +", "2 ** 3", "q"]:
        assert L.lbm_host_eval_expression(bad.encode(), 0.0, 0.0, 0.0, ct.byref(out)) != 0


POISSON_CFGS = ["poisson__poisson1D.json", "poisson__poisson1D_reaction.json", "poisson__poisson2D.json",
                "poisson__poisson2D_helmholtz.json", "poisson__poissonD2Q9.json", "poisson__poisson2D_reaction.json"]


def reduced(name, lvl1=8, lvl2=6):
    """A reference config at a level small enough for the restatements."""
    cfg = json.load(open(os.path.join(PCFG, name)))
    lvl = lvl1 if cfg["dim"] == 1 else lvl2
    cfg["partitionLevel"] = cfg["uniformLevel"] = cfg["maxRfnmtLvl"] = min(lvl, cfg["uniformLevel"])
    return cfg


@pytest.mark.parametrize("name", POISSON_CFGS)
def test_host_grid_equals_restatement_poisson_configs(name):
    from tests.test_grid import assert_same_grid, host_grid, oracle_grid

    cfg = reduced(name)
    assert_same_grid(host_grid(cfg), oracle_grid(cfg))


def test_create_rejects_bad_arguments():
    """Argument checks of lbm_poisson_create run before any device call."""
    from lbm_amd.poisson import Poisson, PoissonError

    for args in [(4, 16, 1, 1.0, 0.0), (3, 16, 2, 1.0, 0.0), (5, 4, 16, 1.0, 0.0), (9, 16, 16, 2.0, 0.0),
                 (5, 16, 16, 0.0, 0.0)]:
        with pytest.raises(PoissonError) as e:
            Poisson(*args)
        assert e.value.code != 0


# ------------------------------------------------------------------ GPU parity
def macro(p):
    """phi of the restatement's current populations (the device's phi)."""
    return p.fold[:, :-1].sum(1) / (1.0 - p.w[-1])


# (lattice, nx, ny, faces in application order): mixed kinds, a Neumann face
# applied after the Dirichlet faces whose extrapolation cells it reads, a
# Neumann face applied before a Dirichlet face, x spanning two 256-wide
# blocks
DEVICE_CASES = [
    (3, 40, 1, [("-x", "dirichlet"), ("+x", "neumann")]),
    (3, 29, 1, [("+x", "dirichlet"), ("-x", "dirichlet")]),
    (5, 37, 23, [("-y", "dirichlet"), ("+x", "neumann"), ("-x", "dirichlet"), ("+y", "neumann")]),
    (9, 300, 9, [("+y", "dirichlet"), ("-x", "neumann"), ("+x", "neumann"), ("-y", "dirichlet")]),
    (9, 21, 17, [("-x", "neumann"), ("-y", "neumann"), ("+x", "dirichlet"), ("+y", "neumann")]),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(DEVICE_CASES)))
def test_device_step_matches_restatement(case):
    """include/lbm_poisson.h against the restatement from the same random
    potential: phi after 1, 2, 9 and 60 steps to 1e-13, and the check
    launch's sum equal to sumAbsDiff of the reference's vars/varsold."""
    from lbm_amd.poisson import Poisson

    Q, nx, ny, spec = DEVICE_CASES[case]
    rng = np.random.default_rng(100 + case)
    faces = {name: (kind, None if kind == "neumann" else rng.random(nx if name in ("-y", "+y") else ny))
             for name, kind in spec}
    tau, src = 0.8, -0.004
    p, _, _ = structured_problem(Q, nx, ny, faces, tau, src)
    phi0 = 0.1 * rng.standard_normal(nx * ny)
    p.init()
    p.vars[:] = phi0
    p.fold = p.feq(phi0)
    print(f"device case {case}: Q={Q} {nx}x{ny} {spec}", flush=True)
    d = Poisson(Q, nx, ny, 1.0 / tau, src)
    try:
        d.set_faces({name: (kind, vals, prio) for prio, (name, (kind, vals)) in enumerate(faces.items())})
        d.init(phi0)
        done = 0
        for m in (1, 2, 9, 60):
            d.step(m - 1 - done)
            p.step(m - 1 - done)
            res, rc = d.step_check()
            p.step(1)
            done = m
            assert rc == 0
            err = np.abs(d.phi() - macro(p)).max()
            ref = p.residual()
            print(f"  step {m}: max|dphi| {err:.2e} check {res:.6e} restated {ref:.6e}")
            assert err < 1e-13
            assert abs(res - ref) <= 1e-11 * ref + 1e-14 * nx * ny
    finally:
        d.close()


@pytest.mark.gpu
def test_device_converges_to_linear_field():
    """D2Q9, no reaction, Dirichlet values linear in the cell coordinates: the
    device's steady state is the linear field (harmonic) and the check sum
    falls below 1e-13."""
    from lbm_amd.poisson import Poisson

    nx, ny = 50, 33

    def lin(x, y):
        return 0.2 - 0.01 * x + 0.02 * y

    d = Poisson(9, nx, ny, 1.0, 0.0)
    try:
        d.set_faces({"-x": ("dirichlet", lin(0, np.arange(ny)), 0), "+x": ("dirichlet", lin(nx - 1, np.arange(ny)), 1),
                     "-y": ("dirichlet", lin(np.arange(nx), 0), 2), "+y": ("dirichlet", lin(np.arange(nx), ny - 1), 3)})
        d.init(np.zeros(nx * ny))
        res = 1.0
        for _ in range(200):
            d.step(199)
            res, rc = d.step_check()
            if res < 1e-13:
                break
        xs, ys = np.arange(nx * ny) % nx, np.arange(nx * ny) // nx
        assert res < 1e-13
        assert np.abs(d.phi() - lin(xs, ys)).max() < 1e-11
    finally:
        d.close()


@pytest.mark.gpu
def test_device_rejects_bad_faces_and_order():
    from lbm_amd.poisson import Poisson, PoissonError

    d = Poisson(5, 16, 16, 1.0, 0.0)
    try:
        with pytest.raises(PoissonError):
            d.step(1)  # before init
        with pytest.raises(PoissonError):
            d.set_faces({"-x": ("dirichlet", np.ones(16), 0), "+x": ("dirichlet", np.ones(16), 1)})  # -y, +y missing
        with pytest.raises(PoissonError):
            d.init(np.zeros(256))  # faces not set
    finally:
        d.close()


# ------------------------------------------------- host on the reference grids
EXPR = {
    "-sinh(sqrt(4+pi^2)*(1-y))/sinh(sqrt(4+pi^2))":
        lambda x, y: -np.sinh(np.sqrt(4 + np.pi ** 2) * (1 - y)) / np.sinh(np.sqrt(4 + np.pi ** 2)),
    "sinh(sqrt(4+pi^2)*(1-y))/sinh(sqrt(4+pi^2))":
        lambda x, y: np.sinh(np.sqrt(4 + np.pi ** 2) * (1 - y)) / np.sinh(np.sqrt(4 + np.pi ** 2)),
    "cos(pi*x)": lambda x, y: np.cos(np.pi * x),
}


def restated_run(cfg, g, steps, every=None):
    """PoissonGridProblem on the restated grid with the configuration's
    objects in the manager's order (surface keys sorted), values at the
    surface cells' centres, kappa and dt as loadConfiguration sets them;
    with `every`, also sumAbsDiff at each iteration t = every, 2 every, ...
    below `steps` (the value the convergence check of iteration t prints)."""
    s = cfg["solver"]
    Q = {"D1Q3": 3, "D2Q5": 5, "D2Q9": 9}[s["model"]]
    n, dim = g["n"], g["dim"]
    dt = 1.0 / (n ** (1.0 / dim) - 1)
    kappa = s["equation_th"] if s.get("equation_application", "debye_huckel") == "simple_diff_reaction" \
        else PO.DEBYE_HUCKEL_KAPPA
    bcs = []
    for grp in sorted(s["boundary"]):
        faces = s["boundary"][grp]
        for f in sorted(faces):
            name = f"{grp}_{f}" if len(faces) > 1 else grp
            cells = g["surf"][name][0]
            v = faces[f]["value"]
            if isinstance(v, list):
                v = v[0]
            if isinstance(v, str):
                cen = g["cen"][cells]
                vals = EXPR[v](cen[:, 0], cen[:, 1] if dim > 1 else 0.0) * np.ones(len(cells))
            else:
                vals = np.full(len(cells), float(v))
            bcs.append((name, faces[f]["type"], vals))
    p = PO.PoissonGridProblem(Q, g["nb"], g["surf"], bcs, s["relaxation"], PO.source_coefficient(Q, s["relaxation"], dt, kappa))
    p.init()
    sums = {}
    if every:
        for t in range(every, steps, every):
            p.step(t - p.steps)
            sums[t] = p.residual()
    p.step(steps - p.steps)
    return (p, sums) if every else p


@pytest.mark.gpu
@pytest.mark.parametrize("name", POISSON_CFGS)
def test_host_matches_restatement_on_reference_grid(name, tmp_path, capfd):
    """The host solver (lbm_host_run_poisson_file) on a reference config at a
    reduced level, run a fixed number of iterations: its final potential in
    reference cell order equals the restatement stepped as many times on
    the restated grid (cell order, neighbours, surfaces, object order,
    expression values at the cell centres) to 1e-12, and every convergence
    check it prints (conv_interval 7) the restatement's sumAbsDiff of vars
    and varsold at that iteration to the 6 printed digits."""
    import ctypes as ct

    from tests.test_grid import oracle_grid

    cfg = reduced(name)
    s = cfg["solver"]
    steps = 1500 if cfg["dim"] == 1 else 250
    s.update(maxSteps=steps, convergence=0.0, solution_interval=10 ** 9, output_dir=str(tmp_path / "out"),
             conv_interval=7)
    for k in ("analyticalSolution", "postprocessing", "errorL2", "errorMax", "errorGRE"):
        s.pop(k, None)
    cfg["outputDir"] = str(tmp_path / "grid")
    path = tmp_path / "case.json"
    path.write_text(json.dumps(cfg))
    g = oracle_grid(cfg)
    L = host_lib()
    L.lbm_host_run_poisson_file.argtypes = [ct.c_char_p, ct.c_int64, ct.c_void_p, ct.c_void_p]
    phi = np.zeros(g["n"])
    done = ct.c_int64()
    print(f"host poisson {name}: {g['n']} cells, {steps} steps", flush=True)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        rc = L.lbm_host_run_poisson_file(str(path).encode(), g["n"], phi.ctypes.data, ct.byref(done))
    finally:
        os.chdir(cwd)
    printed = {}
    for line in capfd.readouterr().err.splitlines():
        head = line.split(":")[0]
        if head.isdigit() and ": dP=" in line:
            printed[int(head)] = float(line.split("dP=")[1])
    assert rc == 0 and done.value == steps
    p, sums = restated_run(cfg, g, steps, every=7)
    err = np.abs(phi - macro(p)).max()
    print(f"  max |host - restated| = {err:.2e}")
    assert err < 1e-12
    assert sorted(printed) == sorted(sums)
    for t, v in sums.items():
        assert abs(printed[t] - v) <= 1e-5 * abs(v) + 1e-300, (t, printed[t], v)


# ------------------------------------------------ the reference's own cases
LBM = os.path.join(ROOT, "lbm_amd", "bin", "lbm")


def run_cli(cfg, tmp_path, timeout=600):
    import subprocess

    p = tmp_path / "case.json"
    cfg["solver"]["output_dir"] = str(tmp_path / "out")
    p.write_text(json.dumps(cfg))
    print("cli:", LBM, p, flush=True)
    return subprocess.run([LBM, str(p), "-p", "f64"], cwd=tmp_path, capture_output=True, text=True, timeout=timeout)


def report(stderr):
    out = {}
    for line in stderr.splitlines():
        for key in ("max. Error", "avg. L2", "global relative error"):
            if line.startswith(key + ":"):
                out[key] = float(line.split(":")[1])
        if line.startswith("Reached convergence"):
            out["converged"] = True
        if line.startswith("Main loop:"):
            out["steps"] = int(line.split()[2])
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["poisson__poisson1D.json", "poisson__poisson1D_reaction.json",
                                  "poisson__poisson2D_helmholtz.json", "poisson__poisson2D.json",
                                  "poisson__poissonD2Q9.json", "poisson__poisson2D_reaction.json"])
def test_reference_poisson_case(name, tmp_path):
    """test/poisson configs through the CLI, unmodified except the output
    directory (and poissonD2Q9's solution_interval, 10 -> 10000: a file every
    10 steps only costs disk): the analytic cases must pass their own
    errorGRE (CHAI08 example 1, the catalyst slab, CHAI08 example 2), the
    others run to exit 0."""
    cfg = json.load(open(os.path.join(PCFG, name)))
    if name == "poisson__poissonD2Q9.json":
        cfg["solver"]["solution_interval"] = 10000
    r = run_cli(cfg, tmp_path)
    rep = report(r.stderr)
    print(name, "exit", r.returncode, rep)
    print(r.stderr[-1500:])
    assert r.returncode == 0, r.stderr[-2000:]
    if "analyticalSolution" in cfg["solver"]:
        assert rep["global relative error"] <= cfg["solver"]["errorGRE"]


@pytest.mark.gpu
def test_reference_poisson_bounce_back_dirichlet_terminates(tmp_path):
    """test/poisson/poisson1D_BBDirichlet.json: the reference's bounce-back
    Dirichlet object terminates for the Poisson equation
    (bnd_dirichlet.h:95-103); the CLI ends with status 255 too."""
    cfg = json.load(open(os.path.join(PCFG, "poisson__poisson1D_BBDirichlet.json")))
    r = run_cli(cfg, tmp_path, timeout=120)
    assert r.returncode == 255
    assert "this is incorrect" in r.stderr

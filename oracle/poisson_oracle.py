"""FP64 restatement of the reference's Poisson (screened Poisson / diffusion-
reaction) lattice-Boltzmann solver, TEST INFRASTRUCTURE ONLY: numpy over the
reference's cell order, neighbour table and surfaces, stepped in the
reference's own pass order.  Only tests/ import it.

Follows:
  lattices D1Q3, D2Q5, D2Q9: directions, weights, poissonAlpha and
    poissonWeights (src/lbm/constants.h:248-319);
  initialCondition, Poisson branch (src/lbm/solver.cpp:283-293): vars = 0,
    boundary initCnd values, f = fold = feq;
  updateMacroscopicValues (solver.cpp:542-545): phi = sum_{i<Q-1} fold_i /
    (1 - w_rest);
  eq::poisson (src/lbm/equilibrium_func.h:149-163): feq_i = w_i phi,
    feq_rest = (w_rest - 1) phi;
  collisionStep (solver.cpp:589-611): f = (1-omega) fold + omega feq, plus
    dt * D * w'_i * kappa^2 phi for i < Q-1 with D = alpha c^2 (1/2 - tau) dt;
  propagationStep (solver.cpp:715-740): push through the grid neighbours;
  LBMBnd_DirichletNEEM::apply (src/lbm/bnd/bnd_dirichlet.h:250-394) and
    LBMBnd_NeumannNEEM::apply (src/lbm/bnd/bnd_neumann.h:14-66) in manager
    order (src/lbm/bnd/bnd.h:69-133, surfaces in key order).
The vars array is kept and overwritten exactly where the reference writes
it, so convergence sums and the final field read the same values.
"""
import numpy as np

D2Q9_C = [(-1, 0), (1, 0), (0, -1), (0, 1), (1, 1), (1, -1), (-1, -1), (-1, 1), (0, 0)]

TABLES = {
    3: dict(c=[(-1,), (1,), (0,)], w=[1 / 6, 1 / 6, 2 / 3], alpha=1 / 3, wp=[0.5, 0.5, 0.0]),
    5: dict(c=[(-1, 0), (1, 0), (0, -1), (0, 1), (0, 0)], w=[1 / 6, 1 / 6, 1 / 6, 1 / 6, 1 / 3], alpha=1 / 2,
            wp=[0.25, 0.25, 0.25, 0.25, 0.0]),
    9: dict(c=D2Q9_C, w=[1 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 36, 1 / 36, 4 / 9], alpha=1 / 3,
            wp=[1 / 8] * 8 + [0.0]),
}

DEBYE_HUCKEL_KAPPA = 27.79  # solver.cpp:597 (k of the Debye-Hueckel approximation)


def opposite(d, dim):
    """cartesian::oppositeDir<NDIM> (include/common/math/cartesian.h:22-34):
    main directions pair up 0/1, 2/3; the 2D diagonals 4/6, 5/7."""
    if d < 2 * dim:
        return d + 1 - 2 * (d % 2)
    return d - 2 if d > 5 else d + 2


def source_coefficient(Q, tau, dt, kappa):
    """dt * D * kappa^2 with D = alpha * c^2 * (1/2 - tau) * dt, c = 1
    (solver.cpp:605-608); the collision adds it times w'_i times phi."""
    diffusivity = TABLES[Q]["alpha"] * 1.0 * (0.5 - tau) * dt
    return dt * diffusivity * kappa * kappa


def extrapolation(cell, nbr, dim):
    """(extrapolation cell, inward direction) of a boundary cell
    (bnd_dirichlet.h:289-316): the first missing main neighbour direction,
    a missing pair -x/-y, -x/+y, +x/+y, +x/-y becoming the diagonal 6, 7, 4,
    5; the extrapolation cell is the neighbour in the opposite direction."""
    ed = -1
    for d in range(2 * dim):
        if nbr[cell, d] < 0:
            if ed < 0:
                ed = d
            elif ed == 0 and d == 2:
                ed = 6
            elif ed == 0 and d == 3:
                ed = 7
            elif ed == 1 and d == 3:
                ed = 4
            elif ed == 1 and d == 2:
                ed = 5
    inward = opposite(ed, dim)
    return int(nbr[cell, inward]), inward


class PoissonGridProblem:
    """bcs: [(surface name, kind, values)] in manager order, kind 'dirichlet'
    or 'neumann' (NEEM), values one per surface cell (Dirichlet values at
    the cell centres; ignored for Neumann, whose gradient is 0 as in the
    reference, bnd_neumann.h:64)."""

    def __init__(self, Q, nbr, surfaces, bcs, tau, source):
        t = TABLES[Q]
        self.Q = Q
        self.dim = len(t["c"][0])
        self.w = np.array(t["w"])
        self.wp = np.array(t["wp"])
        self.omega = 1.0 / tau
        self.source = float(source)
        self.nbr = np.asarray(nbr, dtype=np.int64)
        self.n = self.nbr.shape[0]
        self.bcs = []
        for name, kind, values in bcs:
            cells = list(surfaces[name][0])
            ext = [extrapolation(c, self.nbr, self.dim) for c in cells]
            self.bcs.append(dict(kind=kind, cells=cells, value=np.array(values, dtype=np.float64).copy(),
                                 e=[e for e, _ in ext], dirs=[d for _, d in ext]))
        self.vars = np.zeros(self.n)
        self.varsold = np.zeros(self.n)

    def feq(self, phi):
        f = self.w[None, :] * phi[:, None]
        f[:, -1] = (self.w[-1] - 1.0) * phi
        return f

    def calc_density(self, cells):
        for c in cells:
            self.vars[c] = self.fold[c, :-1].sum() / (1.0 - self.w[-1])

    def init(self):
        """initialCondition: vars = 0, initCnd of the Dirichlet objects (the
        Neumann object's initCnd is empty), f = fold = feq(vars)."""
        self.vars[:] = 0.0
        self.varsold[:] = 0.0
        for b in self.bcs:
            if b["kind"] == "dirichlet":
                for k, c in enumerate(b["cells"]):
                    self.vars[c] = b["value"][k]
        self.fold = self.feq(self.vars)
        self.f = self.fold.copy()
        self.steps = 0

    def apply_dirichlet(self, b):
        self.calc_density(b["e"])
        w = self.w
        for k, c in enumerate(b["cells"]):
            e, v = b["e"][k], b["value"][k]
            self.vars[c] = v
            pe = self.vars[e]
            self.fold[c, :-1] = w[:-1] * v + self.fold[e, :-1] - w[:-1] * pe
            self.fold[c, -1] = (w[-1] - 1.0) * v + self.fold[e, -1] - (w[-1] - 1.0) * pe

    def apply_neumann(self, b):
        for k, c in enumerate(b["cells"]):
            e = b["e"][k]
            ex2 = int(self.nbr[e, b["dirs"][k]])
            self.calc_density([ex2])
            b["value"][k] = (4.0 * self.vars[e] - self.vars[ex2] + 0.0) / 3.0
        self.apply_dirichlet(b)

    def step(self, n=1):
        Q = self.Q
        for _ in range(n):
            self.varsold[:] = self.vars
            self.vars[:] = self.fold[:, :-1].sum(axis=1) / (1.0 - self.w[-1])
            phi = self.vars
            self.f = (1.0 - self.omega) * self.fold + self.omega * self.feq(phi)
            self.f[:, :-1] += self.source * self.wp[None, :-1] * phi[:, None]
            for i in range(Q - 1):
                ok = self.nbr[:, i] >= 0
                self.fold[self.nbr[ok, i], i] = self.f[ok, i]
            self.fold[:, -1] = self.f[:, -1]
            for b in self.bcs:
                if b["kind"] == "dirichlet":
                    self.apply_dirichlet(b)
                else:
                    self.apply_neumann(b)
            self.steps += 1

    def residual(self):
        """sumAbsDiff of electricPotential (solver.cpp:808-815)."""
        return float(np.abs(self.vars - self.varsold).sum())

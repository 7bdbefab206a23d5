"""Python side of the FP64 CPU oracle (TEST INFRASTRUCTURE ONLY).

Only the tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module, and only as the checker.  It wraps
``liblbm_oracle.so`` (lbm_oracle.c) and builds, for a structured box, the
reference-style description the oracle steps: a downstream neighbour table
``nbr[c][i]`` (src/cartesiangrid.h:111-124, -1 = missing) and per-cell
boundary records applied after propagation in list order
(src/lbm/bnd/bnd.h:60-65).

Face rule for a direction whose downstream neighbour leaves the box: the
crossed face decides the boundary; when two faces are crossed the face with
the higher priority wins (default: walls over open faces), ties to the lower
axis.  Missing neighbours inside the box (solid cells) are
plain no-slip bounce-back.  Zou-He records come after the per-direction ones.
"""
import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liblbm_oracle.so")
REF_HILBERT = os.path.join(HERE, "_ref", "libref_hilbert.so")

FACE = {"periodic": 0, "wall": 1, "pressure_abb": 2, "velocity_zh": 3, "pressure_zh": 4, "pressure_neem": 5, "velocity_neem": 6,
        "periodic_pressure": 7, "wall_eq": 8, "wall_neem": 9, "wall_nebb": 10, "wall_nebb0": 11}
# faces whose rule replaces the populations of the whole face cell
FULL = (3, 4, 5, 6, 8, 9, 10, 11)
WET = (8, 9, 10, 11)
EPS = 2.220446049250313e-16


class OrcBC(ct.Structure):
    _fields_ = [("cell", ct.c_int64), ("kind", ct.c_int32), ("dirmask", ct.c_int32), ("axis", ct.c_int32),
                ("sign", ct.c_int32), ("n1", ct.c_int64), ("n2", ct.c_int64), ("u", ct.c_double * 3),
                ("rho", ct.c_double), ("lim1", ct.c_int32), ("lim2", ct.c_int32), ("flags", ct.c_int32),
                ("pad", ct.c_int32)]


def limited_set(c, nrm, Q, upstream_ok, downstream_missing):
    """Limited density constants of a wet-node wall cell
    (LBMBnd_wallWetnode, src/lbm/bnd/bnd_wetnode.h:20-55): population i with
    c_i.n >= eps (pointing out through the wall) counts twice, one with
    |c_i.n| <= eps once (twice when its downstream neighbour is missing and
    the cell is not periodic), each only when its upstream cell exists;
    if the constants do not add up to Q-1 the set is empty (a corner).
    upstream_ok(i), downstream_missing(i) -> bool.  Returns (lim1, lim2,
    empty)."""
    lim1 = lim2 = 0
    total = 0
    for i in range(Q - 1):
        cn = float(c[i] @ nrm)
        if not upstream_ok(i):
            continue
        if cn >= EPS:
            lim2 |= 1 << i
            total += 2
        elif abs(cn) <= EPS:
            if downstream_missing(i):
                lim2 |= 1 << i
                total += 2
            else:
                lim1 |= 1 << i
                total += 1
    if total != Q - 1:
        return 0, 0, True
    return lim1 | (1 << (Q - 1)), lim2, False


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run make -C oracle")
        L = ct.CDLL(LIB)
        dp = ct.POINTER(ct.c_double)
        i64p = ct.POINTER(ct.c_int64)
        L.orc_tables.argtypes = [ct.c_int, dp, dp, ct.POINTER(ct.c_int)]
        L.orc_feq.argtypes = [ct.c_int, ct.c_double, dp, dp]
        L.orc_macros.argtypes = [ct.c_int, ct.c_int64, dp, dp, dp]
        L.orc_step.argtypes = [ct.c_int, ct.c_int64, i64p, ct.c_double, dp, dp, dp, dp, ct.c_int64,
                               ct.POINTER(OrcBC)]
        L.orc_step_forced.argtypes = [ct.c_int, ct.c_int64, i64p, ct.c_double, dp, dp, dp, dp, ct.c_int64,
                                      ct.POINTER(OrcBC), ct.c_int64, i64p, i64p, dp]
        L.orc_init.argtypes = [ct.c_int, ct.c_int64, dp, dp, dp]
        L.orc_residual.argtypes = [ct.c_int, ct.c_int64, dp, dp, dp]
        L.orc_hilbert_index.argtypes = [ct.c_int, dp, ct.c_int64]
        L.orc_hilbert_index.restype = ct.c_int64
        L.orc_round_bf16.argtypes = [ct.c_double]
        L.orc_round_bf16.restype = ct.c_double
        L.orc_round_f16.argtypes = [ct.c_double]
        L.orc_round_f16.restype = ct.c_double
        L.orc_num_threads.restype = ct.c_int
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


def tables(Q):
    c = np.zeros((Q, 3))
    w = np.zeros(Q)
    opp = np.zeros(Q, dtype=np.int32)
    assert lib().orc_tables(Q, _dp(c), _dp(w), opp.ctypes.data_as(ct.POINTER(ct.c_int))) == 0
    return c, w, opp


def feq(Q, rho, u):
    out = np.zeros(Q)
    u = np.ascontiguousarray(np.asarray(u, dtype=np.float64))
    assert lib().orc_feq(Q, float(rho), _dp(u), _dp(out)) == 0
    return out


def hilbert_index(x, level):
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    return lib().orc_hilbert_index(len(x), _dp(x), int(level))


def ref_hilbert_index(x, level):
    """The reference's own hilbert::index compiled into oracle/_ref (None when absent)."""
    if not os.path.exists(REF_HILBERT):
        return None
    L = ct.CDLL(REF_HILBERT)
    L.ref_hilbert_index.argtypes = [ct.c_int, ct.POINTER(ct.c_double), ct.c_longlong]
    L.ref_hilbert_index.restype = ct.c_longlong
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    return L.ref_hilbert_index(len(x), _dp(x), int(level))


class Problem:
    """A structured box described the reference's way for the oracle.

    shape: (nx, ny[, nz]); faces: {"-x": (kind, u, rho), ...}, unlisted = periodic;
    solid: optional bool array of the shape (x fastest when flattened).
    Cell index c = x + nx*(y + ny*z), the same order the GPU library uses.
    """

    def __init__(self, shape, faces=None, solid=None, omega=1.0, force=(0.0, 0.0, 0.0)):
        shape = tuple(int(s) for s in shape)
        if len(shape) == 2:
            shape = shape + (1,)
        self.shape = shape
        nx, ny, nz = shape
        self.dim = 2 if nz == 1 else 3
        self.Q = 9 if self.dim == 2 else 19
        self.omega = float(omega)
        self.force = np.array(list(force) + [0.0] * (3 - len(force)), dtype=np.float64)[:3]
        self.faces = {k: (v[0], tuple(v[1]) if v[1] is not None else (0.0, 0.0, 0.0), float(v[2]),
                          int(v[3]) if len(v) > 3 else (2 if v[0] == "wall" else 1))
                      for k, v in (faces or {}).items()}
        names = ["-x", "+x", "-y", "+y", "-z", "+z"]
        self.face_kind = [FACE[self.faces[n][0]] if n in self.faces else 0 for n in names]
        self.face_prio = [self.faces[n][3] if n in self.faces else 0 for n in names]
        self.face_u = [self.faces[n][1] if n in self.faces else (0.0, 0.0, 0.0) for n in names]
        self.face_rho = [self.faces[n][2] if n in self.faces else 1.0 for n in names]
        self.periodic = [self.face_kind[2 * a] == 0 for a in range(3)]
        self.solid = np.zeros(nx * ny * nz, dtype=bool) if solid is None else np.asarray(solid, bool).reshape(-1)
        self._build()

    def _build(self):
        nx, ny, nz = self.shape
        Q, dim = self.Q, self.dim
        c, w, opp = tables(Q)
        self.c, self.w, self.opp = c, w, opp
        N = nx * ny * nz
        idx = np.arange(N)
        x = idx % nx
        y = (idx // nx) % ny
        z = idx // (nx * ny)
        ext = [nx, ny, nz]
        coords = [x, y, z]
        nbr = np.full((N, Q - 1), -1, dtype=np.int64)
        crossed_face = np.full((N, Q - 1), -1, dtype=np.int64)
        for i in range(Q - 1):
            tgt = []
            ok = np.ones(N, dtype=bool)
            face = np.full(N, -1, dtype=np.int64)
            for a in range(3):
                t = coords[a] + int(c[i, a])
                out = (t < 0) | (t >= ext[a])
                if self.periodic[a]:
                    t = t % ext[a]
                else:
                    f = 2 * a + (1 if c[i, a] > 0 else 0)
                    # the higher face priority wins, ties to the lower axis
                    take = out & ((face < 0) | (self.face_prio[f] > np.array(self.face_prio)[np.maximum(face, 0)]))
                    face = np.where(take, f, face)
                    ok &= ~out
                    t = np.clip(t, 0, ext[a] - 1)
                tgt.append(t)
            n = tgt[0] + nx * (tgt[1] + ny * tgt[2])
            ok &= ~self.solid[n]
            nbr[:, i] = np.where(ok, n, -1)
            crossed_face[:, i] = np.where(ok, -1, face)
        # The oracle's cell list holds fluid cells only (the reference deletes
        # cells inside bodies); solid cells keep their slots here but are
        # never anybody's neighbour and get no records.
        self.fluid = ~self.solid
        self.nbr = nbr
        recs = []
        wet = []
        after = []
        zh = []
        pre = []
        PP = FACE["periodic_pressure"]
        for cidx in np.nonzero(self.fluid & (nbr < 0).any(axis=1))[0]:
            co = [int(x[cidx]), int(y[cidx]), int(z[cidx])]
            # a periodic-pressure face writes every population of its cells
            # that does not stream in, before any other rule (preApply,
            # bnd_periodic.h:102-108); the rules of other faces overwrite
            for f in range(2 * dim):
                ax, side = f // 2, f % 2
                if self.face_kind[f] == PP and co[ax] == (ext[ax] - 1 if side else 0):
                    r = OrcBC()
                    r.cell = int(cidx)
                    r.kind = 7
                    r.rho = self.face_rho[f ^ 1]
                    c1 = list(co)
                    c1[ax] = 0 if side == 1 else ext[ax] - 1
                    r.n1 = c1[0] + nx * (c1[1] + ny * c1[2])
                    r.dirmask = int(sum(1 << i for i in range(Q - 1) if nbr[cidx, i] < 0))
                    pre.append(r)
                    break
            # the whole-cell rule: the full-cell face of highest priority the
            # cell lies on (ties to the lower face index)
            F = None
            for f in range(2 * dim):
                ax, side = f // 2, f % 2
                if self.face_kind[f] in FULL and co[ax] == (ext[ax] - 1 if side else 0):
                    if F is None or self.face_prio[f] > self.face_prio[F]:
                        F = f
            wetF = F is not None and self.face_kind[F] in WET
            groups = {}
            for i in range(Q - 1):
                if nbr[cidx, i] >= 0:
                    continue
                f = int(crossed_face[cidx, i])
                key = f if f >= 0 else "solid"
                groups.setdefault(key, 0)
                groups[key] |= 1 << i
            for key, mask in groups.items():
                r = OrcBC()
                r.cell = int(cidx)
                r.dirmask = mask
                if key == "solid":
                    r.kind = 1
                    recs.append(r)
                    continue
                kind = self.face_kind[key]
                ax, side = key // 2, key % 2
                # a wet-node wall replaces the cell, except the populations
                # whose face has the higher priority: those are written after
                dest = after if wetF and self.face_prio[key] > self.face_prio[F] else recs
                if kind == FACE["wall"]:
                    r.kind = 1
                    r.u = (ct.c_double * 3)(*self.face_u[key])
                    dest.append(r)
                elif kind == FACE["pressure_abb"]:
                    r.kind = 2
                    r.rho = self.face_rho[key]
                    step = -1 if side == 1 else 1  # inward along the axis
                    co = [int(x[cidx]), int(y[cidx]), int(z[cidx])]
                    c1 = list(co); c1[ax] += step
                    c2 = list(co); c2[ax] += 2 * step
                    r.n1 = c1[0] + nx * (c1[1] + ny * c1[2])
                    r.n2 = c2[0] + nx * (c2[1] + ny * c2[2])
                    dest.append(r)
                elif kind == FACE["periodic_pressure"]:
                    pass  # written by the face's record above
                elif kind in (FACE["pressure_neem"], FACE["velocity_neem"]):
                    # the face record (below) replaces every population; with
                    # a solid inward neighbour the missing ones bounce back
                    step = -1 if side == 1 else 1
                    c1 = [int(x[cidx]), int(y[cidx]), int(z[cidx])]
                    c1[ax] += step
                    if self.solid[c1[0] + nx * (c1[1] + ny * c1[2])]:
                        r.kind = 1
                        recs.append(r)
            if wetF:
                # wet-node wall: limited density over the populations that
                # stream in (or enter across a periodic-pressure face)
                ax, side = F // 2, F % 2
                nrm = np.zeros(3)
                nrm[ax] = 1.0 if side else -1.0

                def pp_crossed(i, cidx=cidx):
                    fc = int(crossed_face[cidx, i])
                    return fc >= 0 and self.face_kind[fc] == PP

                lim1, lim2, empty = limited_set(
                    c, nrm, Q, lambda i: nbr[cidx, opp[i]] >= 0 or pp_crossed(opp[i]),
                    lambda i: nbr[cidx, i] < 0 and not pp_crossed(i))
                r = OrcBC()
                r.cell = int(cidx)
                r.kind = self.face_kind[F]
                r.axis = ax
                r.sign = 1 if side else -1
                r.lim1, r.lim2 = lim1, lim2
                # the mass balance divides by 1 + n.u_w; NEBB walls take a
                # tangential velocity, so no division (the reference divides
                # by the cell's own normal velocity there, zero once the wall
                # has acted)
                # bit 2: the device's other rules read the state before
                # the step, never what a wall wrote into vars
                r.flags = (0 if r.kind in (10, 11) else 1) | (2 if empty else 0) | 4
                r.u = (ct.c_double * 3)(*self.face_u[F])
                c1 = list(co)
                c1[ax] += -1 if side else 1
                r.n1 = c1[0] + nx * (c1[1] + ny * c1[2])
                wet.append(r)
            # Zou-He / NEEM face
            for f in ([F] if F is not None and not wetF else []):
                k = self.face_kind[f]
                ax, side = f // 2, f % 2
                r = OrcBC()
                r.cell = int(cidx)
                r.kind = {FACE["velocity_zh"]: 3, FACE["pressure_zh"]: 4, FACE["pressure_neem"]: 5,
                          FACE["velocity_neem"]: 6}[k]
                if k in (FACE["pressure_neem"], FACE["velocity_neem"]):
                    c1 = list(co)
                    c1[ax] += -1 if side else 1
                    r.n1 = c1[0] + nx * (c1[1] + ny * c1[2])
                    if self.solid[r.n1]:
                        break  # bounce-back records above
                r.axis = ax
                r.sign = 1 if side else -1
                r.u = (ct.c_double * 3)(*self.face_u[f])
                r.rho = self.face_rho[f]
                zh.append(r)
                break
        allr = pre + recs + wet + after + zh
        self.nrec = len(allr)
        self.recs = (OrcBC * max(1, len(allr)))(*allr)

    # ------------------------------------------------------------ state
    def init(self, rho, u):
        N = int(np.prod(self.shape))
        rho = np.ascontiguousarray(np.asarray(rho, dtype=np.float64).reshape(N))
        u = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(N, self.dim))
        self.fold = np.zeros(N * self.Q)
        self.fpost = np.zeros(N * self.Q)
        self.vars = np.zeros(N * (self.dim + 1))
        assert lib().orc_init(self.Q, N, _dp(rho), _dp(u), _dp(self.fold)) == 0
        self.fold.reshape(N, self.Q)[self.solid] = 0.0
        self.steps = 0

    def step(self, n=1):
        N = int(np.prod(self.shape))
        L = lib()
        nbr = self.nbr
        for _ in range(int(n)):
            rc = L.orc_step(self.Q, N, nbr.ctypes.data_as(ct.POINTER(ct.c_int64)), self.omega, _dp(self.force),
                            _dp(self.fold), _dp(self.fpost), _dp(self.vars), self.nrec, self.recs)
            assert rc == 0, rc
            self.steps += 1

    def macros(self):
        """(rho[N], u[N,dim]) of fold; solid cells zeroed."""
        N = int(np.prod(self.shape))
        v = np.zeros(N * (self.dim + 1))
        assert lib().orc_macros(self.Q, N, _dp(self.fold), _dp(self.force), _dp(v)) == 0
        v = v.reshape(N, self.dim + 1)
        v[self.solid] = 0.0
        return v[:, self.dim].copy(), v[:, :self.dim].copy()

    def post_collision(self):
        """f_post of the last step, SoA [Q][N] (the GPU state after steps-1 launches)."""
        N = int(np.prod(self.shape))
        return self.fpost.reshape(N, self.Q).T.copy()


def residual(dim, vars_new, vars_old):
    out = np.zeros(dim + 3)
    vn = np.ascontiguousarray(vars_new, dtype=np.float64)
    vo = np.ascontiguousarray(vars_old, dtype=np.float64)
    n = vn.size // (dim + 1)
    assert lib().orc_residual(dim, n, _dp(vn), _dp(vo), _dp(out)) == 0
    return out


def round_bf16(x):
    return lib().orc_round_bf16(float(x))


def round_f16(x):
    return lib().orc_round_f16(float(x))


def num_threads():
    return lib().orc_num_threads()


def taylor_green_noise(shape, u0, sigma, seed):
    """Taylor-Green vortex at cell centres x=(i+1/2) with k = 2 pi / n per axis,
    u = u0 (sin kx cos ky cos kz, -cos kx sin ky cos kz, 0), plus Gaussian
    noise of std sigma from splitmix64 + Box-Muller (seed recorded), so that
    CPU and GPU start from identical fields.  Returns (rho, u[N,3])."""
    nx, ny, nz = shape
    N = nx * ny * nz
    i = np.arange(N)
    X = (i % nx + 0.5) * 2 * np.pi / nx
    Y = ((i // nx) % ny + 0.5) * 2 * np.pi / ny
    Z = (i // (nx * ny) + 0.5) * 2 * np.pi / nz
    u = np.zeros((N, 3))
    u[:, 0] = u0 * np.sin(X) * np.cos(Y) * np.cos(Z)
    u[:, 1] = -u0 * np.cos(X) * np.sin(Y) * np.cos(Z)
    if sigma > 0:
        u += sigma * gaussian_splitmix(3 * N, seed).reshape(N, 3)
    return np.ones(N), u


def splitmix64(n, seed):
    """n uint64 outputs of splitmix64 (Steele, Lea & Flood 2014) from seed."""
    M = np.uint64(0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        s = (np.uint64(seed) + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)) & M
        z = s
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def gaussian_splitmix(n, seed):
    m = (n + 1) // 2
    r = splitmix64(2 * m, seed)
    u1 = ((r[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)
    u2 = (r[1::2] >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    rad = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * m)
    out[0::2] = rad * np.cos(2 * np.pi * u2)
    out[1::2] = rad * np.sin(2 * np.pi * u2)
    return out[:n]


class GridProblem(Problem):
    """A reference-ordered grid (cell ids, neighbour table, surfaces) with the
    reference's boundary objects applied in manager order
    (LBMBndManager::apply, src/lbm/bnd/bnd.h:60-65).

    nbr_table: [n][stride] grid neighbours (lattice direction i == grid
    direction i for i < Q-1); surfaces: {name: (cells, dirs)}; bcs: list of
    (surface name, spec dict) in application order, spec = the reference's
    JSON boundary entry.  Gate of every boundary record: the cell's normal
    n = dirVec(dir) and c_i . n >= eps (bnd_dirichlet.h:85-86).  With
    textbook=True, missing directions no surface gate covers are bounced as
    no-slip walls (interior obstacles); Zou-He records always come last.
    """

    def __init__(self, dim, nbr_table, surfaces, bcs, omega, force=(0.0, 0.0, 0.0), textbook=False, coords=None):
        self.dim = dim
        self.Q = 9 if dim == 2 else 19
        n = nbr_table.shape[0]
        self.shape = (n, 1, 1)
        self.omega = float(omega)
        self.force = np.array(list(force) + [0.0] * (3 - len(force)), dtype=np.float64)[:3]
        self.solid = np.zeros(n, dtype=bool)
        c, w, opp = tables(self.Q)
        self.c, self.w, self.opp = c, w, opp
        self.nbr = np.ascontiguousarray(nbr_table[:, :self.Q - 1].astype(np.int64))
        dv2 = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0)]
        dv3 = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]
        dv = dv2 if dim == 2 else dv3
        recs, zh = [], []
        pre = []  # periodic pressure records: the reference's preApply runs before every apply
        gated = np.zeros((n, self.Q - 1), dtype=bool)
        self.init_u = {}
        # CellProperties::periodic: cells of 'periodic' objects that are
        # generated (bnd.h:219-220)
        periodic_cell = np.zeros(n, dtype=bool)
        for name, spec in bcs:
            if spec["type"] == "periodic" and spec.get("generateBndry", True):
                periodic_cell[np.asarray(surfaces[name][0], dtype=np.int64)] = True
        for name, spec in bcs:
            cells, dirs = surfaces[name]
            t = spec["type"]
            if t == "periodic" and "pressure" in spec and spec.get("generateBndry", True):
                # LBMBndCell_periodic::preApply (bnd_periodic.h:102-108): every
                # cell c of this surface writes the cell L straight across
                # on the connected surface, with this surface's pressure
                assert coords is not None, "periodic pressure needs the cells' structured coordinates"
                co = np.asarray(coords)
                ext = co.max(axis=0) + 1
                key = {tuple(v): k for k, v in enumerate(co.tolist())}
                for cidx, d in zip(cells, dirs):
                    a = d // 2
                    tgt = list(co[cidx])
                    tgt[a] = ext[a] - 1 if d % 2 == 0 else 0
                    L = key[tuple(tgt)]
                    r = OrcBC()
                    r.cell = int(L)
                    r.kind = 7
                    r.n1 = int(cidx)
                    r.rho = float(spec["pressure"])
                    mask = 0
                    for i in range(self.Q - 1):
                        if self.nbr[L, i] < 0:
                            mask |= 1 << i
                            gated[L, i] = True
                    r.dirmask = mask
                    pre.append(r)
                continue
            if t == "periodic" or not spec.get("generateBndry", True):
                continue
            for cidx, d in zip(cells, dirs):
                nrm = np.array(dv[d], dtype=float)
                mask = 0
                for i in range(self.Q - 1):
                    if c[i] @ nrm >= 2.220446049250313e-16:
                        mask |= 1 << i
                        gated[cidx, i] = True
                r = OrcBC()
                r.cell = int(cidx)
                r.dirmask = mask
                if t == "wall" and spec["model"] in ("equilibrium", "neem", "nebb"):
                    # wet-node walls (bnd_wall.h:100-479) in manager order
                    nrm3 = np.zeros(3)
                    nrm3[:dim] = nrm[:dim]
                    per = bool(periodic_cell[cidx])
                    lim1, lim2, empty = limited_set(
                        c, nrm3, self.Q, lambda i: per or nbr_table[cidx, opp[i]] >= 0,
                        lambda i: nbr_table[cidx, i] < 0 and not per)
                    r.kind = {"equilibrium": 8, "neem": 9, "nebb": 10}[spec["model"]]
                    if r.kind == 10 and "velocity" not in spec:
                        r.kind = 11  # LBMBnd_wallNEBB::apply_0
                    a = d // 2
                    r.axis = a
                    r.sign = 1 if d % 2 else -1
                    r.lim1, r.lim2 = lim1, lim2
                    r.flags = (1 if "velocity" in spec else 0) | (2 if empty else 0)
                    v = list(spec.get("velocity", [0.0] * dim))[:dim] + [0.0] * (3 - dim)
                    r.u = (ct.c_double * 3)(*v)
                    # extrapolation cell: the neighbour against the normal
                    # (LBMBnd_wallNEEM::init, bnd_wall.h:213-245)
                    r.n1 = int(nbr_table[cidx, 2 * a + 1 if nrm[a] < 0 else 2 * a])
                    recs.append(r)
                elif t == "wall":
                    r.kind = 1
                    V = spec.get("tangentialVelocity", 0.0)
                    r.u = (ct.c_double * 3)(V * nrm[1], V * nrm[0], 0.0)
                    recs.append(r)
                elif t == "dirichlet" and spec["model"] == "bounceback":
                    r.kind = 1
                    v = list(spec["value"])[:dim] + [0.0] * (3 - dim)
                    r.u = (ct.c_double * 3)(*v)
                    self.init_u[int(cidx)] = v[:dim]
                    recs.append(r)
                elif t == "pressure" and spec.get("model", "antibounceback") == "antibounceback":
                    r.kind = 2
                    r.rho = spec["pressure"]
                    inside = next(2 * a + (1 if nrm[a] < 0 else 0) for a in range(dim) if nrm[a] != 0)
                    r.n1 = int(nbr_table[cidx, inside])
                    r.n2 = int(nbr_table[r.n1, inside])
                    recs.append(r)
                elif t in ("dirichlet", "pressure") and spec.get("model") == "neem":
                    # every population from the inward neighbour; after all
                    # bounce-back records, with the Zou-He records
                    r.kind = 6 if t == "dirichlet" else 5
                    inside = next(2 * a + (1 if nrm[a] < 0 else 0) for a in range(dim) if nrm[a] != 0)
                    r.n1 = int(nbr_table[cidx, inside])
                    if t == "dirichlet":
                        v = list(spec["value"])[:dim] + [0.0] * (3 - dim)
                        r.u = (ct.c_double * 3)(*v)
                        self.init_u[int(cidx)] = v[:dim]
                    else:
                        r.rho = spec["pressure"]
                    zh.append(r)
                elif (t == "dirichlet" and spec["model"] == "zouhe") or (t == "pressure" and spec.get("model") == "zouhe"):
                    r.kind = 3 if t == "dirichlet" else 4
                    r.axis = d // 2
                    r.sign = 1 if d % 2 else -1
                    if t == "dirichlet":
                        v = list(spec["value"])[:dim] + [0.0] * (3 - dim)
                        r.u = (ct.c_double * 3)(*v)
                        self.init_u[int(cidx)] = v[:dim]
                    else:
                        r.rho = spec["pressure"]
                    zh.append(r)
                else:
                    raise NotImplementedError(spec)
        if textbook:
            extra = []
            for cidx in np.nonzero((self.nbr < 0).any(axis=1))[0]:
                mask = 0
                for i in range(self.Q - 1):
                    if self.nbr[cidx, i] < 0 and not gated[cidx, i]:
                        mask |= 1 << i
                if mask:
                    r = OrcBC()
                    r.cell = int(cidx)
                    r.kind = 1
                    r.dirmask = mask
                    extra.append(r)
            recs = extra + recs
        allr = pre + recs + zh
        self.nrec = len(allr)
        self.recs = (OrcBC * max(1, len(allr)))(*allr)

    def set_forcing(self, cells, value_cells, pressures):
        """LBMSolver::forcing (src/lbm/solver.cpp:626-696): after every
        collision cell cells[k] takes feq(pressures[k], u_v) + f(v) - feq(v)
        from v = value_cells[k] (orc_step_forced)."""
        self.forcing = (np.ascontiguousarray(cells, dtype=np.int64), np.ascontiguousarray(value_cells, dtype=np.int64),
                        np.ascontiguousarray(pressures, dtype=np.float64))

    def step(self, n=1):
        fz = getattr(self, "forcing", None)
        if fz is None or len(fz[0]) == 0:
            return super().step(n)
        N = int(np.prod(self.shape))
        L = lib()
        i64 = ct.POINTER(ct.c_int64)
        for _ in range(int(n)):
            rc = L.orc_step_forced(self.Q, N, self.nbr.ctypes.data_as(i64), self.omega, _dp(self.force), _dp(self.fold),
                                   _dp(self.fpost), _dp(self.vars), self.nrec, self.recs, len(fz[0]),
                                   fz[0].ctypes.data_as(i64), fz[1].ctypes.data_as(i64), _dp(fz[2]))
            assert rc == 0, rc
            self.steps += 1

    def init_reference(self):
        """initialCondition (src/lbm/solver.cpp:267-295): rho = 1, u from the
        boundaries' initCnd, f = feq."""
        n = self.shape[0]
        u = np.zeros((n, self.dim))
        for cidx, v in self.init_u.items():
            u[cidx] = v
        self.init(np.ones(n), u)

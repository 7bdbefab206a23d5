"""Literal CPU restatement of the reference's grid generator (TEST
INFRASTRUCTURE ONLY; pure-Python loops, small grids).

It follows the reference's steps in their own order -- level offsets that
alternate between the two ends of the cell arrays, refineCell with the
childDir table, findChildLevelNghbrs with the nghbrInside/nghbrParentChildId
tables, markOutsideCells/floodCells with a stack, deleteCell with
swap-with-last compaction, reorderHilbertCurve's argsort + swap procedure --
so that the product host (lbm_amd/host/grid.cpp, which works on a structured
window instead) can be checked against it cell id by cell id.

Multi-level grids: uniformRefineGrid (every cell of a level refined, children
appended after the cells so far) and markBndryCells / refineMarkedCells
(the boundary cells of the finest level refined), each followed by the same
neighbour search and deletion; the solver side tests boundary cells by the
parent rule at each level's length.  Pinned by tests/test_grid.py
(test_multilevel_restatement_invariants: parent/child geometry, closed-form
level sizes and child ids on the unit square, brute-force neighbour links,
boundary refinement of exactly the boundary cells).

Reference: src/gridgenerator/cartesiangrid_generation.h:71-147 (partition
grid), :151-206 (uniform and boundary refinement), :343-380 (refineGrid), :419-451 (refineCell), :453-506 (neighbours),
:509-620 (deletion, flood fill), :637-661 (copyCell/swapCell/updateParent),
:664-735 (Hilbert reorder); src/cartesiangrid.h:451-493 (diagonals),
:510-546 (boundary cells), :553-603 (surfaces), :608-706 (periodic links);
src/geometry.h:640-1075 (box/cube/sphere tests); include/common/math/cartesian.h:72-153.
"""
import math
import sys

from oracle import oracle as O

EPS = 2.220446049250313e-16
MAX_LVL = 100
EPS_DBL = sys.float_info.epsilon  # approx(a, b), include/common/math/mathfunctions.h:60-62

# include/common/math/cartesian.h:72-101 (first 8 children)
CHILD_DIR = [(-1, -1, -1), (1, -1, -1), (-1, 1, -1), (1, 1, -1), (-1, -1, 1), (1, -1, 1), (-1, 1, 1), (1, 1, 1)]
# include/common/math/cartesian.h:104-125, rows 0..7, dirs -x +x -y +y -z +z
NGHBR_INSIDE = [[-1, 1, -1, 2, -1, 4], [0, -1, -1, 3, -1, 5], [-1, 3, 0, -1, -1, 6], [2, -1, 1, -1, -1, 7],
                [-1, 5, -1, 6, 0, -1], [4, -1, -1, 7, 1, -1], [-1, 7, 4, -1, 2, -1], [6, -1, 5, -1, 3, -1]]
# include/common/math/cartesian.h:129-153
NGHBR_PARENT_CHILD = [[1, -1, 2, -1, 4, -1], [-1, 0, 3, -1, 5, -1], [3, -1, -1, 0, 6, -1], [-1, 2, -1, 1, 7, -1],
                      [5, -1, 6, -1, -1, 0], [-1, 4, 7, -1, -1, 1], [7, -1, -1, 4, -1, 2], [-1, 6, -1, 5, -1, 3]]
DIR_VEC2 = [(-1, 0), (1, 0), (0, -1), (0, 1), (1, 1), (1, -1), (-1, -1), (-1, 1)]
DIR_VEC3 = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1), (-1, -1, 0), (-1, 1, 0), (1, -1, 0),
            (1, 1, 0), (-1, 0, -1), (-1, 0, 1), (1, 0, -1), (1, 0, 1), (0, -1, -1), (0, -1, 1), (0, 1, -1), (0, 1, 1),
            (-1, -1, -1), (-1, -1, 1), (-1, 1, -1), (-1, 1, 1), (1, -1, -1), (1, -1, 1), (1, 1, -1), (1, 1, 1)]


def opposite(d):
    return d + 1 - 2 * (d % 2)


# ------------------------------------------------------------- geometry
class Box:
    def __init__(self, name, g, dim):
        self.name, self.dim = name, dim
        self.A, self.B = list(g["A"]), list(g["B"])
        self.body = g.get("body", "unique")
        self.subtract = g.get("subtract", False)

    def inside(self, x):
        return all(not (self.A[d] > x[d] or self.B[d] < x[d]) for d in range(self.dim))

    def cut(self, c, length):
        h = 0.5 * length
        A, B = self.A, self.B
        if self.dim == 2:
            if A[0] - h <= c[0] and B[0] + h >= c[0]:
                if abs(c[1] - A[1]) <= h or abs(c[1] - B[1]) <= h:
                    return True
            if A[1] - h <= c[1] and B[1] + h >= c[1]:
                if abs(c[0] - A[0]) <= h or abs(c[0] - B[0]) <= h:
                    return True
            return False
        for a in range(self.dim):
            if all(A[b] - h <= c[b] and B[b] + h >= c[b] for b in range(self.dim) if b != a):
                if abs(c[a] - A[a]) <= h or abs(c[a] - B[a]) <= h:
                    return True
        return False

    def bbox(self):
        return list(self.A), list(self.B)


class Sphere:
    def __init__(self, name, g, dim):
        self.name, self.dim = name, dim
        self.c, self.r = list(g["center"]), g["radius"]
        self.body = g.get("body", "unique")
        self.subtract = g.get("subtract", False)

    def dist(self, x):
        return math.sqrt(sum((x[d] - self.c[d]) ** 2 for d in range(self.dim)))

    def inside(self, x):
        return self.dist(x) < self.r + EPS

    def cut(self, cc, length):
        distance = self.dist(cc)
        half = 0.5 * length
        surr = math.sqrt(2) * half
        if distance <= self.r + surr and distance >= self.r - surr:
            n = 1 << self.dim
            cnt = 0
            for k in range(n):
                v = [(cc[d] - half) + length * ((k >> d) & 1) for d in range(self.dim)]
                if self.dist(v) <= self.r:
                    cnt += 1
            return cnt < n
        return False

    def bbox(self):
        return [self.c[d] - self.r for d in range(self.dim)], [self.c[d] + self.r for d in range(self.dim)]


class Geometry:
    def __init__(self, gcfg, dim):
        self.objs = []
        for name in sorted(gcfg):
            t = gcfg[name]["type"]
            if t == "box":
                self.objs.append(Box(name, gcfg[name], dim))
            elif t == "sphere":
                self.objs.append(Sphere(name, gcfg[name], dim))
            else:
                raise NotImplementedError(t)
        self.bodies = {}
        for o in self.objs:
            if o.body == "unique":
                o.body = o.name
            self.bodies.setdefault(o.body, []).append(o)

    def inside(self, x):
        for key in sorted(self.bodies):
            lst = self.bodies[key]
            has_sub = any(o.subtract for o in lst)
            for o in lst:
                if o.inside(x):
                    if not has_sub:
                        return True
                    if o.subtract:
                        return False
                    if any(s.subtract and s.inside(x) for s in lst):
                        return False
                    return True
        return False

    def cut(self, c, length, name=None):
        return any(o.cut(c, length) for o in self.objs if name is None or o.name == name)

    def bbox(self, dim):
        mn, mx = [0.0] * dim, [0.0] * dim
        for i, o in enumerate(self.objs):
            a, b = o.bbox()
            for d in range(dim):
                if mn[d] > a[d] or i == 0:
                    mn[d] = a[d]
                if mx[d] < b[d] or i == 0:
                    mx[d] = b[d]
        return mn, mx


# -------------------------------------------------------------- generator
class Generator:
    def __init__(self, cfg):
        self.dim = cfg["dim"]
        self.L = cfg["partitionLevel"]
        self.U = cfg["uniformLevel"]
        self.M = cfg.get("maxRfnmtLvl", self.U)
        assert self.L <= self.U <= self.M
        self.geom = Geometry(cfg["geometry"], self.dim)
        self.align = bool(cfg.get("alignNodesWithSurface", False))
        self.align_dir = int(cfg.get("alignDir", 1))
        mn, mx = self.geom.bbox(self.dim)
        ext = [abs(mx[d] - mn[d]) for d in range(self.dim)]
        decisive = 0
        for d in range(self.dim):
            decisive = d if ext[d] > ext[decisive] else decisive
        self.cog = [mn[d] + 0.5 * (mx[d] - mn[d]) for d in range(self.dim)]
        self.len = [(1.0 + 1.0 / math.pow(2.0, MAX_LVL)) * ext[decisive]]
        for _ in range(self.M):
            self.len.append(0.5 * self.len[-1])
        self.nch = 1 << self.dim
        self.nn = 2 * self.dim
        n_est = sum((1 << (self.dim * l)) for l in range(self.M + 1))
        self.cap = min(cfg["maxNoCells"], 4 * n_est + 16)
        cap = self.cap
        self.center = [None] * cap
        self.level = [0] * cap
        self.parent = [-1] * cap
        self.child = [[-1] * self.nch for _ in range(cap)]
        self.nchild = [0] * cap
        self.nghbr = [[-1] * self.nn for _ in range(cap)]
        self.bndry = [False] * cap
        self.inside = [False] * cap
        self.marked = [False] * cap
        self.offsets = [None] * (self.M + 1)

    # createPartitioningGrid
    def run(self):
        cap, nch = self.cap, self.nch
        if self.L % 2 == 0:
            self.offsets[0] = [0, 1]
            self.offsets[1] = [cap - nch, cap]
        else:
            self.offsets[0] = [cap - 1, cap]
            self.offsets[1] = [0, nch]
        b = self.offsets[0][0]
        self.center[b] = list(self.cog)
        self.bndry[b] = True
        for l in range(self.L):
            pb, pe = self.offsets[l]
            pn = pe - pb
            if pb == 0:
                self.offsets[l + 1] = [cap - pn * nch, cap]
            else:
                self.offsets[l + 1] = [0, pn * nch]
            for cid in range(pb, pe):
                self.refine_cell(cid, self.offsets[l + 1][0] + (cid - pb) * nch)
            self.find_child_nghbrs(l)
            self.delete_outside(l + 1)
        for i in range(cap):
            self.parent[i] = -1
        self.size = self.offsets[self.L][1] - self.offsets[self.L][0]
        assert self.offsets[self.L][0] == 0
        self.reorder_hilbert()
        # uniformRefineGrid (cartesiangrid_generation.h:151-169): every cell of
        # a level gets its children, appended after the cells so far
        highest = self.L
        for lvl in range(self.L, self.U):
            b, e = self.offsets[lvl]
            self.offsets[lvl + 1] = [self.size, self.size + (e - b) * nch]
            for cid in range(b, e):
                self.refine_cell(cid, self.offsets[lvl + 1][0] + (cid - b) * nch)
            self.refine_level_done(lvl)
            highest += 1
        # boundary refinement (gridGenerator.cpp:209-213, markBndryCells /
        # refineMarkedCells, cartesiangrid_generation.h:171-206,409-417): the
        # boundary cells of the finest level get children
        for _ in range(self.U, self.M):
            b, e = self.offsets[highest]
            marked = [c for c in range(b, e) if self.bndry[c]]
            self.offsets[highest + 1] = [self.size, self.size + len(marked) * nch]
            for k, cid in enumerate(marked):
                self.refine_cell(cid, self.offsets[highest + 1][0] + k * nch)
            self.refine_level_done(highest)
            highest += 1
        if self.align:
            self.align_with_surface(self.align_dir)
        return self

    def refine_level_done(self, lvl):
        """refineGrid's tail (cartesiangrid_generation.h:362-378): neighbours
        of the new level, deletion of its outside cells."""
        self.size = self.offsets[lvl + 1][1]
        self.find_child_nghbrs(lvl)
        self.delete_outside(lvl + 1)
        self.size = self.offsets[lvl + 1][1]

    def align_with_surface(self, align_dir):
        """transformMaxRfnmtLvlToExtent
        (src/gridgenerator/cartesiangrid_generation.h:256-304): stretch the
        finest level so that its outermost cell centres lie on the bounding
        box along align_dir (along every axis when the nodal extent is the
        same in all of them), scale the cell length alike, then delete the
        cells that are neither cut nor inside."""
        b, e = self.offsets[self.M]
        mn = [sys.float_info.max] * self.dim
        mx = [sys.float_info.min] * self.dim
        for c in range(b, e):
            for d in range(self.dim):
                if mn[d] > self.center[c][d]:
                    mn[d] = self.center[c][d]
                if mx[d] < self.center[c][d]:
                    mx[d] = self.center[c][d]
        same = True
        for d in range(self.dim):
            for d2 in range(d + 1, self.dim):
                if not (abs(mn[d] - mn[d2]) < EPS_DBL and abs(mx[d] - mx[d2]) < EPS_DBL):
                    same = False
            if not same:
                break
        bmin, bmax = self.geom.bbox(self.dim)
        T = bmax[align_dir] / (mx[align_dir] - mn[align_dir])
        for c in range(b, e):
            for d in range(self.dim):
                if d == align_dir or same:
                    self.center[c][d] = self.center[c][d] * T - (T * mx[d] - bmax[d])
                else:
                    self.center[c][d] = self.center[c][d] * T
        self.len[self.M] *= T
        for c in range(b, e):
            isb = self.geom.cut(self.center[c], self.len[self.M])
            self.bndry[c] = isb
            self.inside[c] = isb or self.geom.inside(self.center[c])
        for c in range(e - 1, b - 1, -1):
            if not self.inside[c]:
                self.delete_cell(c)
        self.size = self.offsets[self.M][1]

    def refine_cell(self, cid, off):
        lvl = self.level[cid] + 1
        ln = self.len[lvl]
        for k in range(self.nch):
            c = off + k
            self.center[c] = [self.center[cid][d] + (0.5 * ln) * CHILD_DIR[k][d] for d in range(self.dim)]
            self.level[c] = lvl
            self.parent[c] = cid
            self.nchild[c] = 0
            self.bndry[c] = False
            self.inside[c] = False
            self.marked[c] = False
            self.child[c] = [-1] * self.nch
            self.nghbr[c] = [-1] * self.nn
            if self.bndry[cid]:
                self.bndry[c] = self.geom.cut(self.center[c], ln)
            self.child[cid][k] = c
            self.nchild[cid] += 1

    def find_child_nghbrs(self, l):
        pb, pe = self.offsets[l]
        for p in range(pb, pe):
            ch = self.child[p]
            for k in range(self.nch):
                if ch[k] == -1:
                    continue
                nb = self.nghbr[ch[k]]
                for d in range(self.nn):
                    if nb[d] != -1:
                        continue
                    ni = NGHBR_INSIDE[k][d]
                    if ni != -1:
                        nb[d] = ch[ni]
                    else:
                        pc = NGHBR_PARENT_CHILD[k][d]
                        pn = self.nghbr[p][d]
                        if pn != -1 and pc != -1 and self.child[pn][pc] != -1:
                            nb[d] = self.child[pn][pc]

    def delete_outside(self, lvl):
        b, e = self.offsets[lvl]
        for c in range(b, e):
            self.marked[c] = False
        for c in range(b, e):
            if self.marked[c]:
                continue
            self.marked[c] = True
            isb = self.bndry[c]
            self.inside[c] = isb or self.geom.inside(self.center[c])
            if not isb:
                self.flood(c)
        for c in range(e - 1, b - 1, -1):
            if not self.inside[c]:
                self.delete_cell(c)

    def flood(self, c):
        stack = [c]
        val = self.inside[c]
        while stack:
            cur = stack.pop()
            for d in range(self.nn):
                nb = self.nghbr[cur][d]
                if nb != -1 and not self.marked[nb]:
                    self.marked[nb] = True
                    if not self.bndry[nb]:
                        self.inside[nb] = val
                        stack.append(nb)
                    else:
                        self.inside[nb] = True

    def update_parent(self, p, old, new):
        for k in range(self.nch):
            if self.child[p][k] == old:
                self.child[p][k] = new
                return

    def delete_cell(self, c):
        p = self.parent[c]
        lvl = self.level[c]
        if lvl != self.L:
            self.update_parent(p, c, -1)
            self.nchild[p] -= 1
        for d in range(self.nn):
            nb = self.nghbr[c][d]
            if nb != -1:
                self.nghbr[nb][opposite(d)] = -1
        if c != self.offsets[lvl][1] - 1:
            self.copy_cell(self.offsets[lvl][1] - 1, c)
        self.offsets[lvl][1] -= 1

    def copy_cell(self, frm, to):
        self.bndry[to] = self.bndry[frm]
        self.inside[to] = self.inside[frm]
        self.marked[to] = self.marked[frm]
        self.level[to] = self.level[frm]
        self.center[to] = list(self.center[frm])
        self.parent[to] = self.parent[frm]
        self.nghbr[to] = list(self.nghbr[frm])
        self.child[to] = list(self.child[frm])
        self.nchild[to] = self.nchild[frm]
        for d in range(self.nn):
            if self.nghbr[to][d] != -1:
                self.nghbr[self.nghbr[to][d]][opposite(d)] = to
            self.nghbr[frm][d] = -1
        if self.parent[to] != -1:
            self.update_parent(self.parent[to], frm, to)
        for k in range(self.nch):
            if self.child[to][k] != -1:
                self.parent[self.child[to][k]] = to

    def swap_cell(self, a, b):
        self.copy_cell(a, self.cap - 1)
        self.copy_cell(b, a)
        self.copy_cell(self.cap - 1, b)

    def reorder_hilbert(self):
        n = self.size
        L0 = self.len[0]
        hid = []
        for c in range(n):
            x = [((self.center[c][d] - self.cog[d]) + 0.5 * L0) / L0 for d in range(self.dim)]
            hid.append(O.hilbert_index(x, self.L))
        assert len(set(hid)) == n, "duplicated Hilbert ids"
        index = sorted(range(n), key=lambda a: hid[a])
        pos = list(range(n))
        rev = list(range(n))
        for i in range(n):
            hp = index[i]
            if i != hp:
                if i != pos[hp]:
                    self.swap_cell(i, pos[hp])
                tmp = rev[i]
                rev[pos[hp]] = tmp
                pos[tmp] = pos[hp]


# ----------------------------------------------------------- solver side
class SolverGrid:
    """CartesianGrid::loadGridInplace on the generated cells."""

    def __init__(self, gen, solver_cfg):
        self.dim = gen.dim
        n = gen.size
        self.n = n
        self.center = [list(gen.center[c]) for c in range(n)]
        nn = 2 * self.dim
        main = [list(gen.nghbr[c]) for c in range(n)]
        self.level = [gen.level[c] for c in range(n)]
        self.parent = [gen.parent[c] for c in range(n)]

        def ln(c):
            return gen.len[self.level[c]]

        # determineBoundaryCells (src/cartesiangrid.h:510-546): a cell of the
        # partition level, or a child of a boundary cell, is one if the
        # geometry cuts it at its level's length and a main neighbour is
        # missing (parents come before their children in cell order)
        self.bndry = []
        for c in range(n):
            b = False
            p = self.parent[c]
            if p == -1 or self.bndry[p]:
                b = gen.geom.cut(self.center[c], ln(c))
                if b and all(main[c][d] != -1 for d in range(nn)):
                    b = False
            self.bndry.append(b)
        # identifyBndrySurfaces
        bcfg = solver_cfg.get("boundary", {})
        self.surfaces = {}
        assigned = set()
        for gname in sorted(bcfg):
            faces = bcfg[gname]
            nb = len(faces)
            for fname in sorted(faces):
                sname = f"{gname}_{fname}" if nb > 1 else gname
                cells, dirs = [], {}
                d0, d1 = (0, nn) if fname == "all" else ("-x+x-y+y-z+z".index(fname) // 2,
                                                          "-x+x-y+y-z+z".index(fname) // 2 + 1)
                for d in range(d0, d1):
                    for c in range(n):
                        if not self.bndry[c] or main[c][d] != -1:
                            continue
                        if not gen.geom.cut(self.center[c], ln(c), gname):
                            continue
                        if (c, d) in assigned:
                            continue
                        assigned.add((c, d))
                        cells.append(c)
                        dirs[c] = d
                self.surfaces[sname] = (cells, dirs)
        # setupPeriodicConnections
        conn = {}
        for gname in sorted(bcfg):
            for fname, sc in sorted(bcfg[gname].items()):
                if sc.get("type") == "periodic" and not sc.get("generateBndry", True):
                    conn[f"{gname}_{fname}"] = sc["connection"]
        done = set()
        for a, b in sorted(conn.items()):
            if a in done:
                continue
            done.update((a, b))
            for ca in self.surfaces[a][0]:
                for cb in self.surfaces[b][0]:
                    nm = -1
                    for d in range(self.dim):
                        if abs(self.center[ca][d] - self.center[cb][d]) > EPS:
                            if nm >= 0:
                                nm = -1
                                break
                            nm = d
                    if nm < 0:
                        continue
                    if self.center[ca][nm] > self.center[cb][nm]:
                        main[cb][2 * nm] = ca
                        main[ca][2 * nm + 1] = cb
                    else:
                        main[ca][2 * nm] = cb
                        main[cb][2 * nm + 1] = ca
        # addDiagonalNghbrs
        def t(c, d):
            return -1 if c == -1 else main[c][d]
        self.nghbr = []
        for c in range(n):
            row = list(main[c])
            if self.dim == 2:
                mX, pX = main[c][0], main[c][1]
                row += [t(pX, 3), t(pX, 2), t(mX, 2), t(mX, 3)]
            elif self.dim == 3:
                for d in range(6, 26):
                    v = DIR_VEC3[d]
                    cur = c
                    for a in range(3):
                        if v[a] == 0 or cur == -1:
                            continue
                        cur = t(cur, 2 * a + (1 if v[a] > 0 else 0))
                    row.append(cur)
            self.nghbr.append(row)


def build(cfg):
    gen = Generator(cfg).run()
    return gen, SolverGrid(gen, cfg.get("solver", {}))

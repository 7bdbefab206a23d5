"""CPU restatement of the reference's "line" postprocessing function (TEST
INFRASTRUCTURE ONLY; numpy).

Cells are kept when their centre lies within half a cell length of
Eigen::Hyperplane::Through(A, B) (src/common/line.h, used by
PostprocessCartesianFunctionLine::init,
src/postprocess/postprocessing_cartesian.h:17-42): normal = (B - A)'s
unitOrthogonal -- in 2D (-dy, dx) / |d|; in 3D (-dy, dx, 0) / |(dx, dy)|
unless dx and dy are both below 1e-12 |dz|, then (0, -dz, dy) / |(dy, dz)|
-- offset = -A.normal, distance = |normal.p + offset|.  The kept cells are
sorted by centre, x first (the reference's comparator agrees wherever it is
an order).  line_csv gives the file ASCII::writePointsCSV writes
(src/common/IO.h:37-97, called from src/postprocess/postprocessing.h:100-114).

Pinned in tests/test_postprocess.py: the 2D distance against the textbook
point-line distance |d x (p - A)| / |d|, the 3D plane against its defining
properties (unit normal orthogonal to d, A and B on it), and axis-parallel
lines against the closed-form column of cells.
"""
import math

import numpy as np


def line_hyperplane(a, b):
    a = [float(v) for v in a]
    b = [float(v) for v in b]
    dim = len(a)
    d = [b[k] - a[k] for k in range(dim)]
    if dim == 2:
        n = [-d[1], d[0]]
        z = n[0] * n[0] + n[1] * n[1]
        if z > 0:
            s = math.sqrt(z)
            n = [n[0] / s, n[1] / s]
    else:
        prec = 1e-12
        if not (abs(d[0]) <= abs(d[2]) * prec) or not (abs(d[1]) <= abs(d[2]) * prec):
            inv = 1.0 / math.sqrt(d[0] * d[0] + d[1] * d[1])
            n = [-d[1] * inv, d[0] * inv, 0.0]
        else:
            inv = 1.0 / math.sqrt(d[1] * d[1] + d[2] * d[2])
            n = [0.0, -d[2] * inv, d[1] * inv]
    dot = 0.0
    for k in range(dim):
        dot += a[k] * n[k]
    return n, -dot


def distances(centers, a, b):
    n, off = line_hyperplane(a, b)
    dim = len(a)
    dot = np.zeros(len(centers))
    for k in range(dim):
        dot = dot + n[k] * centers[:, k]
    return np.abs(dot + off)


def line_cells(centers, length, a, b):
    """centers [n][dim], length: cell length (one level)."""
    keep = np.nonzero(0.5 * length >= distances(centers, a, b))[0]
    dim = centers.shape[1]
    # sort by x, then y, then z, then id
    order = sorted(range(len(keep)), key=lambda i: tuple(centers[keep[i], k] for k in range(dim)) + (keep[i],))
    return keep[order]


def line_csv(centers, cells, u):
    dim = centers.shape[1]
    out = [",".join("xyz"[:dim]) + ",u\n"]
    for c in cells:
        out.append(",".join("%.15g" % centers[c, k] for k in range(dim)) + ",%.15f\n" % u[c])
    return "".join(out)

// Grid generation in the reference's cell order; see grid.h for the rules.
#include "grid.h"

#include <algorithm>
#include <numeric>

namespace lbmh {

namespace {
// hilbert::index (include/common/math/hilbert.h:15-48): fixed quadrant table,
// no rotation; x in unit-cube coordinates.
auto hilbertIndex(int dim, const GDouble* xin, GInt level) -> GInt {
  static const GInt LUT[16] = {0, 3, 1, 2, 5, 4, 6, 7, 10, 9, 11, 8, 15, 14, 12, 13};
  GDouble p[3] = {xin[0], dim > 1 ? xin[1] : 0, dim > 2 ? xin[2] : 0};
  GInt index = 0;
  for (GInt l = 0; l < level; ++l) {
    int q = 0;
    int qb[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) {
      qb[d] = p[d] >= 0.5 ? 1 : 0;
      q |= qb[d] << d;
    }
    GInt mult = 1;
    for (GInt k = 0; k < dim * (level - 1 - l); ++k) mult *= 2;
    index += mult * LUT[q];
    for (int d = 0; d < dim; ++d) p[d] = 2 * p[d] - qb[d];
  }
  return index;
}

struct Window {
  std::array<GInt, 3> lo{{0, 0, 0}}, n{{1, 1, 1}};
  [[nodiscard]] auto size() const -> GInt { return n[0] * n[1] * n[2]; }
  [[nodiscard]] auto idx(GInt i, GInt j, GInt k) const -> GInt {
    return (i - lo[0]) + n[0] * ((j - lo[1]) + n[1] * (k - lo[2]));
  }
  [[nodiscard]] auto inside(GInt i, GInt j, GInt k) const -> bool {
    return i >= lo[0] && i < lo[0] + n[0] && j >= lo[1] && j < lo[1] + n[1] && k >= lo[2] && k < lo[2] + n[2];
  }
};
}  // namespace

auto dirVec(GInt dim, GInt dir) -> std::array<GDouble, 3> {
  static const int D2[8][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}, {1, 1}, {1, -1}, {-1, -1}, {-1, 1}};
  static const int D3[26][3] = {{-1, 0, 0},  {1, 0, 0},   {0, -1, 0},  {0, 1, 0},   {0, 0, -1},  {0, 0, 1},  {-1, -1, 0},
                                {-1, 1, 0},  {1, -1, 0},  {1, 1, 0},   {-1, 0, -1}, {-1, 0, 1},  {1, 0, -1}, {1, 0, 1},
                                {0, -1, -1}, {0, -1, 1},  {0, 1, -1},  {0, 1, 1},   {-1, -1, -1}, {-1, -1, 1}, {-1, 1, -1},
                                {-1, 1, 1},  {1, -1, -1}, {1, -1, 1},  {1, 1, -1},  {1, 1, 1}};
  std::array<GDouble, 3> v{{0, 0, 0}};
  if (dim == 2) {
    v[0] = D2[dir][0];
    v[1] = D2[dir][1];
  } else if (dim == 3) {
    for (int d = 0; d < 3; ++d) v[d] = D3[dir][d];
  } else {
    v[0] = dir == 0 ? -1 : 1;
  }
  return v;
}

// state of the tree refinement: the generator's boundary flag, children,
// flood-fill flags and the id range of each level
struct CartesianGrid::Refinement {
  std::vector<uint8_t> gb, inside, marked;
  std::vector<std::array<GInt, 8>> child;
  std::vector<GInt> begin, end;
};

namespace {
// include/common/math/cartesian.h:72-153: child directions, the sibling that
// neighbours a child in a direction, and the child of the parent's
// neighbour that does
constexpr int kChildDir[8][3] = {{-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {1, 1, -1},
                                 {-1, -1, 1},  {1, -1, 1},  {-1, 1, 1},  {1, 1, 1}};
constexpr int kNghbrInside[8][6] = {{-1, 1, -1, 2, -1, 4}, {0, -1, -1, 3, -1, 5}, {-1, 3, 0, -1, -1, 6},
                                    {2, -1, 1, -1, -1, 7}, {-1, 5, -1, 6, 0, -1}, {4, -1, -1, 7, 1, -1},
                                    {-1, 7, 4, -1, 2, -1}, {6, -1, 5, -1, 3, -1}};
constexpr int kNghbrParentChild[8][6] = {{1, -1, 2, -1, 4, -1}, {-1, 0, 3, -1, 5, -1}, {3, -1, -1, 0, 6, -1},
                                         {-1, 2, -1, 1, 7, -1}, {5, -1, 6, -1, -1, 0}, {-1, 4, 7, -1, -1, 1},
                                         {7, -1, -1, 4, -1, 2}, {-1, 6, -1, 5, -1, 3}};
}  // namespace

void CartesianGrid::generate(Configuration& cfg) {
  m_dim = (int)cfg.required_config_value<GInt>("dim");
  if (m_dim < 1 || m_dim > 3) TERMM(-1, "Only 1D, 2D and 3D grids are supported");
  const GInt maxNoCells = cfg.required_config_value<GInt>("maxNoCells");
  const GInt partitionLvl = cfg.required_config_value<GInt>("partitionLevel");
  const GInt uniformLvl = cfg.required_config_value<GInt>("uniformLevel");
  if (partitionLvl > uniformLvl) TERMM(-1, "Invalid definition of grid level partitionLevel >= uniformLevel");
  const GInt maxRfnmt = cfg.opt_config_value<GInt>("maxRfnmtLvl", uniformLvl);
  if (maxRfnmt < uniformLvl) TERMM(-1, "Invalid definition of grid level uniformLevel >= maxRfnmtLvl");
  if (partitionLvl < 1 || partitionLvl > 14) TERMM(-1, "partitionLevel must be in 1..14");
  if (maxRfnmt > 16) TERMM(-1, "maxRfnmtLvl must be at most 16");
  m_level = partitionLvl;
  m_maxLvl = maxRfnmt;
  m_uniformLvl = uniformLvl;
  cfg.opt_config_value<bool>("dry-run", false);
  cfg.opt_config_value<std::string>("outputDir", "out");
  cfg.opt_config_value<std::string>("gridFileName", "grid");
  cfg.getObject("output");
  const bool align = cfg.opt_config_value<bool>("alignNodesWithSurface", false);
  const GInt alignDir = cfg.opt_config_value<GInt>("alignDir", 1);

  m_geom = std::make_shared<GeometryManager>();
  json defGeom = {{"cube", {{"type", "cube"}, {"center", {0.0, 0.0, 0.0}}, {"length", 1}}}};
  m_geom->setup(cfg.has_config_value("geometry") ? cfg.getObject("geometry") : defGeom, m_dim);
  if (m_geom->noObjects() == 0 || cfg.has_config_value("boundingBox")) {
    auto bb = cfg.required_config_value<std::vector<GDouble>>("boundingBox");
    if ((int)bb.size() < 2 * m_dim) TERMM(-1, "boundingBox needs 2*dim values");
    m_bbox.dim = m_dim;
    for (int d = 0; d < m_dim; ++d) {
      m_bbox.min[d] = bb[2 * d];
      m_bbox.max[d] = bb[2 * d + 1];
    }
  } else {
    m_bbox = m_geom->boundingBox(m_dim);
  }
  // BaseCartesianGrid::setBoundingBox
  int decisive = 0;
  std::array<GDouble, 3> ext{{0, 0, 0}};
  for (int d = 0; d < m_dim; ++d) {
    ext[d] = std::abs(m_bbox.max[d] - m_bbox.min[d]);
    decisive = ext[d] > ext[decisive] ? d : decisive;
    m_cog[d] = m_bbox.min[d] + 0.5 * (m_bbox.max[d] - m_bbox.min[d]);
  }
  m_len.assign(m_maxLvl + 1, 0.0);
  m_len[0] = (1.0 + 1.0 / std::pow(2.0, (GDouble)MAX_LVL)) * ext[decisive];
  for (GInt l = 1; l <= m_maxLvl; ++l) m_len[l] = 0.5 * m_len[l - 1];

  // per-axis centre coordinates for each level, parent + (len/2)*(+-1)
  std::vector<std::vector<std::array<GDouble, 3>>> cen(m_level + 1);
  cen[0].assign(1, m_cog);
  for (GInt l = 1; l <= m_level; ++l) {
    const GInt nl = GInt(1) << l;
    cen[l].resize(nl);
    const GDouble halfLen = 0.5 * m_len[l];
    for (GInt i = 0; i < nl; ++i)
      for (int d = 0; d < 3; ++d) cen[l][i][d] = cen[l - 1][i >> 1][d] + halfLen * ((i & 1) ? 1.0 : -1.0);
  }
  auto centreOf = [&](GInt l, GInt i, GInt j, GInt k, GDouble* c) {
    c[0] = cen[l][i][0];
    c[1] = m_dim >= 2 ? cen[l][j][1] : 0.0;
    c[2] = m_dim == 3 ? cen[l][k][2] : 0.0;
  };

  // level 0: the root cell, a boundary cell by construction
  Window w;
  std::vector<uint8_t> alive(1, 1), bnd(1, 1);
  for (GInt l = 0; l < m_level; ++l) {
    Window wc;
    for (int d = 0; d < 3; ++d) {
      if (d < m_dim) {
        wc.lo[d] = 2 * w.lo[d];
        wc.n[d] = 2 * w.n[d];
      }
    }
    const GInt nc = wc.size();
    if (nc > (GInt)4e9) TERMM(-1, "Out of memory! grid window too large");
    std::vector<uint8_t> cand(nc, 0), cb(nc, 0), inside(nc, 0), marked(nc, 0);
    const GDouble lenC = m_len[l + 1];
    // candidates = children of surviving cells; boundary = parent boundary and cut
    for (GInt k = wc.lo[2]; k < wc.lo[2] + wc.n[2]; ++k)
      for (GInt j = wc.lo[1]; j < wc.lo[1] + wc.n[1]; ++j)
        for (GInt i = wc.lo[0]; i < wc.lo[0] + wc.n[0]; ++i) {
          const GInt pi = i >> 1, pj = m_dim >= 2 ? (j >> 1) : 0, pk = m_dim == 3 ? (k >> 1) : 0;
          const GInt p = w.idx(pi, pj, pk);
          if (!alive[p]) continue;
          const GInt c = wc.idx(i, j, k);
          cand[c] = 1;
          if (bnd[p]) {
            GDouble x[3];
            centreOf(l + 1, i, j, k, x);
            cb[c] = m_geom->cutWithCell(x, lenC) ? 1 : 0;
          }
        }
    // flood fill of inside/outside over non-boundary candidates (main neighbours)
    std::vector<GInt> stack;
    for (GInt k = wc.lo[2]; k < wc.lo[2] + wc.n[2]; ++k)
      for (GInt j = wc.lo[1]; j < wc.lo[1] + wc.n[1]; ++j)
        for (GInt i = wc.lo[0]; i < wc.lo[0] + wc.n[0]; ++i) {
          const GInt c = wc.idx(i, j, k);
          if (!cand[c] || marked[c]) continue;
          marked[c] = 1;
          if (cb[c]) {
            inside[c] = 1;
            continue;
          }
          GDouble x[3];
          centreOf(l + 1, i, j, k, x);
          const uint8_t in = m_geom->pointIsInside(x) ? 1 : 0;
          inside[c] = in;
          stack.clear();
          stack.push_back(c);
          while (!stack.empty()) {
            const GInt cur = stack.back();
            stack.pop_back();
            const GInt ci = wc.lo[0] + cur % wc.n[0];
            const GInt cj = wc.lo[1] + (cur / wc.n[0]) % wc.n[1];
            const GInt ck = wc.lo[2] + cur / (wc.n[0] * wc.n[1]);
            for (int dir = 0; dir < 2 * m_dim; ++dir) {
              GInt ni = ci, nj = cj, nk = ck;
              if (dir == 0) --ni;
              if (dir == 1) ++ni;
              if (dir == 2) --nj;
              if (dir == 3) ++nj;
              if (dir == 4) --nk;
              if (dir == 5) ++nk;
              if (!wc.inside(ni, nj, nk)) continue;
              const GInt nb = wc.idx(ni, nj, nk);
              if (!cand[nb] || marked[nb]) continue;
              marked[nb] = 1;
              if (!cb[nb]) {
                inside[nb] = in;
                stack.push_back(nb);
              } else {
                inside[nb] = 1;
              }
            }
          }
        }
    std::vector<uint8_t> alive2(nc, 0);
    GInt count = 0;
    for (GInt c = 0; c < nc; ++c) {
      alive2[c] = cand[c] && (cb[c] || inside[c]);
      count += alive2[c];
    }
    if (count > maxNoCells) TERMM(-1, "Out of memory! Increase maxNoCells");
    // shrink the window to the surviving cells
    Window ws;
    std::array<GInt, 3> mn{{std::numeric_limits<GInt>::max(), std::numeric_limits<GInt>::max(), std::numeric_limits<GInt>::max()}};
    std::array<GInt, 3> mx{{-1, -1, -1}};
    for (GInt c = 0; c < nc; ++c) {
      if (!alive2[c]) continue;
      const GInt ci = wc.lo[0] + c % wc.n[0];
      const GInt cj = wc.lo[1] + (c / wc.n[0]) % wc.n[1];
      const GInt ck = wc.lo[2] + c / (wc.n[0] * wc.n[1]);
      const GInt co[3] = {ci, cj, ck};
      for (int d = 0; d < 3; ++d) {
        mn[d] = std::min(mn[d], co[d]);
        mx[d] = std::max(mx[d], co[d]);
      }
    }
    if (count == 0) TERMM(-1, "Grid generation removed every cell (geometry outside the bounding box?)");
    for (int d = 0; d < 3; ++d) {
      ws.lo[d] = mn[d];
      ws.n[d] = mx[d] - mn[d] + 1;
    }
    std::vector<uint8_t> a3(ws.size(), 0), b3(ws.size(), 0);
    for (GInt c = 0; c < nc; ++c) {
      if (!alive2[c]) continue;
      const GInt ci = wc.lo[0] + c % wc.n[0];
      const GInt cj = wc.lo[1] + (c / wc.n[0]) % wc.n[1];
      const GInt ck = wc.lo[2] + c / (wc.n[0] * wc.n[1]);
      a3[ws.idx(ci, cj, ck)] = 1;
      b3[ws.idx(ci, cj, ck)] = cb[c];
    }
    w = ws;
    alive.swap(a3);
    bnd.swap(b3);
  }

  // cells of the final level, ordered by Hilbert index (reorderHilbertCurve)
  std::vector<std::array<GInt, 3>> cells;
  std::vector<GInt> hid;
  const GDouble L0 = m_len[0];
  for (GInt k = w.lo[2]; k < w.lo[2] + w.n[2]; ++k)
    for (GInt j = w.lo[1]; j < w.lo[1] + w.n[1]; ++j)
      for (GInt i = w.lo[0]; i < w.lo[0] + w.n[0]; ++i) {
        if (!alive[w.idx(i, j, k)]) continue;
        cells.push_back({{i, j, k}});
        GDouble c[3], x[3];
        centreOf(m_level, i, j, k, c);
        for (int d = 0; d < m_dim; ++d) x[d] = ((c[d] - m_cog[d]) + 0.5 * L0) / L0;
        hid.push_back(hilbertIndex(m_dim, x, m_level));
      }
  std::vector<GInt> order(cells.size());
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](GInt a, GInt b) { return hid[a] < hid[b]; });
  for (size_t q = 1; q < order.size(); ++q)
    if (hid[order[q]] == hid[order[q - 1]]) TERMM(-1, "Duplicated Hilbert Ids found!");
  const GInt N = (GInt)cells.size();
  m_cell.resize(N);
  m_center.assign(N * 3, 0.0);
  for (GInt q = 0; q < N; ++q) {
    m_cell[q] = cells[order[q]];
    centreOf(m_level, m_cell[q][0], m_cell[q][1], m_cell[q][2], &m_center[q * 3]);
  }
  // main neighbours of the partition level
  std::vector<GInt> at(w.size(), INVALID_CELLID);
  for (GInt q = 0; q < N; ++q) at[w.idx(m_cell[q][0], m_cell[q][1], m_cell[q][2])] = q;
  m_stride = 2 * m_dim;
  m_nghbr.assign(N * m_stride, INVALID_CELLID);
  for (GInt q = 0; q < N; ++q) {
    for (int dir = 0; dir < 2 * m_dim; ++dir) {
      std::array<GInt, 3> co = m_cell[q];
      co[dir / 2] += (dir % 2) ? 1 : -1;
      if (!w.inside(co[0], co[1], co[2])) continue;
      m_nghbr[q * m_stride + dir] = at[w.idx(co[0], co[1], co[2])];
    }
  }
  m_cellLvl.assign(N, m_level);
  m_parent.assign(N, INVALID_CELLID);
  // finer levels
  if (m_maxLvl > m_level) {
    Refinement st;
    st.gb.resize(N);
    for (GInt q = 0; q < N; ++q) st.gb[q] = bnd[w.idx(m_cell[q][0], m_cell[q][1], m_cell[q][2])];
    st.child.assign(N, {{INVALID_CELLID, INVALID_CELLID, INVALID_CELLID, INVALID_CELLID, INVALID_CELLID, INVALID_CELLID,
                         INVALID_CELLID, INVALID_CELLID}});
    st.begin.assign(m_maxLvl + 1, 0);
    st.end.assign(m_maxLvl + 1, 0);
    st.end[m_level] = N;
    for (GInt l = m_level; l < uniformLvl; ++l) refineLevel(l, false, maxNoCells, st);
    for (GInt l = uniformLvl; l < m_maxLvl; ++l) refineLevel(l, true, maxNoCells, st);
    m_leaf.assign(size(), 1);
    for (GInt q = 0; q < size(); ++q)
      for (GInt c : st.child[q])
        if (c != INVALID_CELLID) m_leaf[q] = 0;
  } else {
    m_leaf.assign(N, 1);
  }
  m_levels.clear();
  for (GInt q = 0; q < size(); ++q)
    if (m_levels.empty() || m_levels.back() != m_cellLvl[q]) m_levels.push_back(m_cellLvl[q]);
  if (align) alignWithSurface(alignDir);
  if (m_levels.size() == 1) {
    setWindow();
  } else {  // no single structured window: see levelGrid
    m_n = {{0, 0, 0}};
    m_struct.clear();
    m_cellOf.clear();
  }
}

// structured window of a single-level grid: the bounding range of the level
// indices, each cell's index in it and the cell (or none) at each position
void CartesianGrid::setWindow() {
  const GInt N = size();
  std::array<GInt, 3> mn{{std::numeric_limits<GInt>::max(), std::numeric_limits<GInt>::max(), std::numeric_limits<GInt>::max()}};
  std::array<GInt, 3> mx{{std::numeric_limits<GInt>::min(), std::numeric_limits<GInt>::min(), std::numeric_limits<GInt>::min()}};
  for (GInt q = 0; q < N; ++q)
    for (int d = 0; d < 3; ++d) {
      mn[d] = std::min(mn[d], m_cell[q][d]);
      mx[d] = std::max(mx[d], m_cell[q][d]);
    }
  Window w;
  for (int d = 0; d < 3; ++d) {
    w.lo[d] = mn[d];
    w.n[d] = mx[d] - mn[d] + 1;
  }
  m_lo = w.lo;
  m_n = w.n;
  m_struct.assign(N, -1);
  m_cellOf.assign(w.size(), -1);
  for (GInt q = 0; q < N; ++q) {
    const GInt s = w.idx(m_cell[q][0], m_cell[q][1], m_cell[q][2]);
    m_struct[q] = s;
    m_cellOf[s] = q;
  }
}

// refineGrid for level lvl (src/gridgenerator/cartesiangrid_generation.h:343-417):
// the children of every cell of the level (or of its boundary cells) are
// appended in parent order (refineCell, :419-451), get their neighbours
// from their siblings and the parent's neighbours (findChildLevelNghbrs,
// :453-506), and the new level's outside cells are deleted: flood fill in
// cell order from the non-boundary cells, then deletion from the end with
// the level's last cell copied into each hole (:509-661)
void CartesianGrid::refineLevel(GInt lvl, bool onlyBndry, GInt maxNoCells, Refinement& st) {
  const int nch = 1 << m_dim, nn = 2 * m_dim;
  const GInt b = st.begin[lvl], e = st.end[lvl];
  std::vector<GInt> parents;
  for (GInt c = b; c < e; ++c)
    if (!onlyBndry || st.gb[c]) parents.push_back(c);
  const GInt nb = size(), ne = nb + (GInt)parents.size() * nch;
  if (ne > maxNoCells) TERMM(-1, "Out of memory! Increase maxNoCells");
  m_cell.resize(ne);
  m_center.resize(ne * 3, 0.0);
  m_cellLvl.resize(ne);
  m_parent.resize(ne);
  m_nghbr.resize(ne * m_stride, INVALID_CELLID);
  st.gb.resize(ne);
  st.inside.assign(ne, 0);
  st.marked.assign(ne, 0);
  st.child.resize(ne);
  const GDouble len = m_len[lvl + 1];
  for (size_t k = 0; k < parents.size(); ++k) {
    const GInt p = parents[k];
    for (int ch = 0; ch < nch; ++ch) {
      const GInt c = nb + (GInt)k * nch + ch;
      m_cellLvl[c] = lvl + 1;
      m_parent[c] = p;
      for (int d = 0; d < 3; ++d) m_cell[c][d] = d < m_dim ? 2 * m_cell[p][d] + (kChildDir[ch][d] > 0 ? 1 : 0) : 0;
      for (int d = 0; d < 3; ++d) m_center[c * 3 + d] = d < m_dim ? m_center[p * 3 + d] + 0.5 * len * kChildDir[ch][d] : 0.0;
      st.child[c].fill(INVALID_CELLID);
      for (int dir = 0; dir < nn; ++dir) m_nghbr[c * m_stride + dir] = INVALID_CELLID;
      st.gb[c] = st.gb[p] ? (m_geom->cutWithCell(&m_center[c * 3], len) ? 1 : 0) : 0;
      st.child[p][ch] = c;
    }
  }
  for (GInt p = b; p < e; ++p)
    for (int ch = 0; ch < nch; ++ch) {
      const GInt c = st.child[p][ch];
      if (c == INVALID_CELLID) continue;
      for (int dir = 0; dir < nn; ++dir) {
        GInt& nbr = m_nghbr[c * m_stride + dir];
        if (nbr != INVALID_CELLID) continue;
        const int ni = kNghbrInside[ch][dir];
        if (ni != -1) {
          nbr = st.child[p][ni];
        } else {
          const int pc = kNghbrParentChild[ch][dir];
          const GInt pn = m_nghbr[p * m_stride + dir];
          if (pn != INVALID_CELLID && pc != -1 && st.child[pn][pc] != INVALID_CELLID) nbr = st.child[pn][pc];
        }
      }
    }
  // markOutsideCells / floodCells
  std::vector<GInt> stack;
  for (GInt c = nb; c < ne; ++c) {
    if (st.marked[c]) continue;
    st.marked[c] = 1;
    const bool isB = st.gb[c] != 0;
    st.inside[c] = (isB || m_geom->pointIsInside(&m_center[c * 3])) ? 1 : 0;
    if (isB) continue;
    const uint8_t val = st.inside[c];
    stack.assign(1, c);
    while (!stack.empty()) {
      const GInt cur = stack.back();
      stack.pop_back();
      for (int dir = 0; dir < nn; ++dir) {
        const GInt n = m_nghbr[cur * m_stride + dir];
        if (n == INVALID_CELLID || st.marked[n]) continue;
        st.marked[n] = 1;
        if (!st.gb[n]) {
          st.inside[n] = val;
          stack.push_back(n);
        } else {
          st.inside[n] = 1;
        }
      }
    }
  }
  // deleteCell / copyCell / updateParent
  auto updateParent = [&](GInt p, GInt from, GInt to) {
    for (auto& x : st.child[p])
      if (x == from) {
        x = to;
        return;
      }
  };
  GInt end = ne;
  for (GInt c = ne - 1; c >= nb; --c) {
    if (st.inside[c]) continue;
    updateParent(m_parent[c], c, INVALID_CELLID);
    for (int dir = 0; dir < nn; ++dir) {
      const GInt n = m_nghbr[c * m_stride + dir];
      if (n != INVALID_CELLID) m_nghbr[n * m_stride + (dir ^ 1)] = INVALID_CELLID;
    }
    const GInt from = end - 1;
    if (c != from) {
      m_cell[c] = m_cell[from];
      for (int d = 0; d < 3; ++d) m_center[c * 3 + d] = m_center[from * 3 + d];
      m_cellLvl[c] = m_cellLvl[from];
      m_parent[c] = m_parent[from];
      st.gb[c] = st.gb[from];
      st.inside[c] = st.inside[from];
      st.marked[c] = st.marked[from];
      st.child[c] = st.child[from];
      for (int dir = 0; dir < nn; ++dir) {
        const GInt n = m_nghbr[from * m_stride + dir];
        m_nghbr[c * m_stride + dir] = n;
        if (n != INVALID_CELLID) m_nghbr[n * m_stride + (dir ^ 1)] = c;
        m_nghbr[from * m_stride + dir] = INVALID_CELLID;
      }
      if (m_parent[c] != INVALID_CELLID) updateParent(m_parent[c], from, c);
      for (GInt x : st.child[c])
        if (x != INVALID_CELLID) m_parent[x] = c;
    }
    --end;
  }
  m_cell.resize(end);
  m_center.resize(end * 3);
  m_cellLvl.resize(end);
  m_parent.resize(end);
  m_nghbr.resize(end * m_stride);
  st.gb.resize(end);
  st.child.resize(end);
  st.begin[lvl + 1] = nb;
  st.end[lvl + 1] = end;
}

// transformMaxRfnmtLvlToExtent (src/gridgenerator/cartesiangrid_generation.h:256-304):
// stretch the cell centres so that the outermost ones lie on the bounding
// box along alignDir (along every axis when the nodal extent is the same in
// all of them) and scale the cell length alike.  The reference then deletes
// the cells that are neither cut nor inside; that would reshape the
// structured window, so such a configuration is rejected here.
void CartesianGrid::alignWithSurface(GInt alignDir) {
  const GInt N = (GInt)m_cell.size();
  if (alignDir < 0 || alignDir >= m_dim) TERMM(-1, "Invalid alignDir");
  auto onMax = [&](GInt q) { return m_cellLvl[q] == m_maxLvl; };
  std::array<GDouble, 3> mn{{std::numeric_limits<GDouble>::max(), std::numeric_limits<GDouble>::max(),
                             std::numeric_limits<GDouble>::max()}};
  std::array<GDouble, 3> mx{{std::numeric_limits<GDouble>::min(), std::numeric_limits<GDouble>::min(),
                             std::numeric_limits<GDouble>::min()}};
  for (GInt q = 0; q < N; ++q)
    for (int d = 0; d < m_dim && onMax(q); ++d) {
      if (mn[d] > m_center[q * 3 + d]) mn[d] = m_center[q * 3 + d];
      if (mx[d] < m_center[q * 3 + d]) mx[d] = m_center[q * 3 + d];
    }
  bool same = true;
  for (int d = 0; d < m_dim && same; ++d)
    for (int d2 = d + 1; d2 < m_dim; ++d2)
      if (!(std::abs(mn[d] - mn[d2]) < std::numeric_limits<GDouble>::epsilon() &&
            std::abs(mx[d] - mx[d2]) < std::numeric_limits<GDouble>::epsilon())) {
        same = false;
        break;
      }
  const GDouble T = m_bbox.max[alignDir] / (mx[alignDir] - mn[alignDir]);
  for (GInt q = 0; q < N; ++q)
    for (int d = 0; d < m_dim && onMax(q); ++d) {
      GDouble& c = m_center[q * 3 + d];
      if (d == alignDir || same) c = c * T - (T * mx[d] - m_bbox.max[d]);
      else c = c * T;
    }
  m_len[m_maxLvl] *= T;
  for (GInt q = 0; q < N; ++q) {
    if (!onMax(q)) continue;
    const GDouble* c = &m_center[q * 3];
    if (!m_geom->cutWithCell(c, m_len[m_maxLvl]) && !m_geom->pointIsInside(c))
      TERMM(-1, "alignNodesWithSurface moves cells out of the geometry; deleting them is not supported by this build");
  }
}

void CartesianGrid::loadGridInplace(Configuration& solverCfg) {
  const GInt N = size();
  // determineBoundaryCells: a partition-level cell or a child of a boundary
  // cell (parents precede their children) that the geometry cuts at its
  // level's length and that misses at least one main neighbour
  m_bndry.assign(N, 0);
  m_noBndCells = 0;
  for (GInt q = 0; q < N; ++q) {
    bool b = false;
    if (m_parent[q] == INVALID_CELLID || m_bndry[m_parent[q]]) {
      b = m_geom->cutWithCell(&m_center[q * 3], cellLength(q));
      if (b) {
        GInt nn = 0;
        for (int dir = 0; dir < 2 * m_dim; ++dir) nn += neighbor(q, dir) != INVALID_CELLID;
        if (nn == 2 * m_dim) b = false;
      }
    }
    m_bndry[q] = b;
    m_noBndCells += b && m_leaf[q];
  }
  // identifyBndrySurfaces
  json bndryConfig = solverCfg.getObject("boundary");
  std::set<std::pair<GInt, GInt>> assigned;
  for (auto g = bndryConfig.begin(); g != bndryConfig.end(); ++g) {
    const std::string surfName = g.key();
    const GInt noBnds = (GInt)g.value().size();
    for (auto f = g.value().begin(); f != g.value().end(); ++f) {
      const std::string dirName = f.key();
      const std::string nameAp = noBnds > 1 ? surfName + "_" + dirName : surfName;
      m_surfaces.emplace(nameAp, Surface(this));
      const GInt dirBegin = dirName == "all" ? 0 : dirIdString2Id(dirName);
      const GInt dirEnd = dirName == "all" ? 2 * m_dim : dirIdString2Id(dirName) + 1;
      for (GInt dir = dirBegin; dir < dirEnd; ++dir) {
        for (GInt q = 0; q < N; ++q) {
          if (!m_bndry[q]) continue;
          if (neighbor(q, dir) != INVALID_CELLID) continue;
          if (!m_geom->cutWithCell(surfName, &m_center[q * 3], cellLength(q))) continue;
          if (assigned.insert({q, dir}).second) m_surfaces.at(nameAp).addCell(q, dir);
        }
      }
      if (m_surfaces.at(nameAp).size() == 0) std::cerr << "WARNING: surface " << nameAp << " has no cells!" << std::endl;
    }
  }
  setPeriodicConnections(solverCfg);
  addDiagonalNghbrs();
  checkPeriodicWrap();
}

void CartesianGrid::setPeriodicConnections(Configuration& solverCfg) {
  if (!solverCfg.has_any_key_value("type", "periodic")) return;
  json bndryConfig = solverCfg.getObject("boundary");
  std::map<std::string, std::string> conn;
  for (auto g = bndryConfig.begin(); g != bndryConfig.end(); ++g)
    for (auto f = g.value().begin(); f != g.value().end(); ++f) {
      const json& sc = f.value();
      if (opt_value<std::string>(sc, "type", "notset") != "periodic") continue;
      if (opt_value<bool>(sc, "generateBndry", true)) continue;
      conn.emplace(g.key() + "_" + f.key(), required_value<std::string>(sc, "connection"));
    }
  std::set<std::string> done;
  for (const auto& [a, b] : conn) {
    if (done.count(a)) continue;
    if (!conn.count(b)) TERMM(-1, "Invalid periodic setup!");
    done.insert(b);
    done.insert(a);
    addPeriodicConnection(bndrySurface(a), bndrySurface(b));
  }
  m_hasPeriodicLinks = !conn.empty();
}

// src/cartesiangrid.h:645-706: link cells whose centres differ in exactly
// one coordinate (tolerance GDoubleEps).
void CartesianGrid::addPeriodicConnection(const Surface& A, const Surface& B) {
  for (const GInt a : A.getCellList()) {
    for (const GInt b : B.getCellList()) {
      GInt notMatching = -1;
      for (int d = 0; d < m_dim; ++d) {
        if (std::abs(m_center[a * 3 + d] - m_center[b * 3 + d]) > GDoubleEps) {
          if (notMatching >= 0) {
            notMatching = -1;
            break;
          }
          notMatching = d;
        }
      }
      if (notMatching < 0) continue;
      const GInt nd = 2 * notMatching;
      if (m_center[a * 3 + notMatching] > m_center[b * 3 + notMatching]) {
        m_nghbr[b * m_stride + nd] = a;
        m_nghbr[a * m_stride + nd + 1] = b;
      } else {
        m_nghbr[a * m_stride + nd] = b;
        m_nghbr[b * m_stride + nd + 1] = a;
      }
      m_periodicCell.insert(a);
      m_periodicCell.insert(b);
    }
  }
}

// 2D: src/cartesiangrid.h:451-493 (+x+y, +x-y, -x-y, -x+y from main links);
// 3D: D3Q27 order, composed along x, then y, then z.
void CartesianGrid::addDiagonalNghbrs() {
  const GInt N = size();
  const GInt ns = maxNoNghbrsDiag(m_dim);
  std::vector<GInt> tmp = m_nghbr;
  const GInt os = m_stride;
  auto tmpN = [&](GInt c, GInt dir) -> GInt { return c == INVALID_CELLID ? INVALID_CELLID : tmp[c * os + dir]; };
  m_nghbr.assign(N * ns, INVALID_CELLID);
  for (GInt q = 0; q < N; ++q) {
    for (GInt dir = 0; dir < 2 * m_dim; ++dir) m_nghbr[q * ns + dir] = tmpN(q, dir);
    if (m_dim == 2) {
      const GInt mX = tmpN(q, 0), pX = tmpN(q, 1);
      m_nghbr[q * ns + 4] = tmpN(pX, 3);
      m_nghbr[q * ns + 5] = tmpN(pX, 2);
      m_nghbr[q * ns + 6] = tmpN(mX, 2);
      m_nghbr[q * ns + 7] = tmpN(mX, 3);
    } else if (m_dim == 3) {
      for (GInt dir = 6; dir < 26; ++dir) {
        const auto v = dirVec(3, dir);
        GInt c = q;
        for (int a = 0; a < 3 && c != INVALID_CELLID; ++a) {
          if (v[a] == 0) continue;
          c = tmpN(c, 2 * a + (v[a] > 0 ? 1 : 0));
        }
        m_nghbr[q * ns + dir] = c;
      }
    }
  }
  m_stride = ns;
}

// An axis is periodic for the device when every cell of the window's low
// face is linked to the high-face cell of its row and vice versa.
void CartesianGrid::checkPeriodicWrap() {
  for (int a = 0; a < 3; ++a) m_periodicAxis[a] = false;
  if (!m_hasPeriodicLinks || m_levels.size() != 1) return;
  for (int a = 0; a < m_dim; ++a) {
    bool any = false, all = true;
    for (GInt q = 0; q < size() && all; ++q) {
      const auto& co = m_cell[q];
      const bool lo = co[a] == m_lo[a], hi = co[a] == m_lo[a] + m_n[a] - 1;
      if (!lo && !hi) continue;
      std::array<GInt, 3> other = co;
      other[a] = lo ? m_lo[a] + m_n[a] - 1 : m_lo[a];
      GInt s = (other[0] - m_lo[0]) + m_n[0] * ((other[1] - m_lo[1]) + m_n[1] * (other[2] - m_lo[2]));
      const GInt expect = m_cellOf[s];
      const GInt got = neighbor(q, 2 * a + (lo ? 0 : 1));
      if (got != INVALID_CELLID) any = true;
      if (got != expect) all = false;
    }
    m_periodicAxis[a] = any && all;
  }
}

auto CartesianGrid::levelGrid(GInt lvl, std::vector<GInt>& ids) const -> std::unique_ptr<CartesianGrid> {
  auto g = std::make_unique<CartesianGrid>();
  g->m_dim = m_dim;
  g->m_level = lvl;
  g->m_maxLvl = lvl;
  g->m_uniformLvl = lvl;
  g->m_bbox = m_bbox;
  g->m_cog = m_cog;
  g->m_len = m_len;
  g->m_geom = m_geom;
  ids.clear();
  std::vector<GInt> loc(size(), INVALID_CELLID);
  for (GInt q = 0; q < size(); ++q)
    if (m_cellLvl[q] == lvl) {
      loc[q] = (GInt)ids.size();
      ids.push_back(q);
    }
  const GInt n = (GInt)ids.size();
  if (n == 0) TERMM(-1, "levelGrid: no cells on level " + std::to_string(lvl));
  g->m_stride = m_stride;
  g->m_nghbr.assign(n * m_stride, INVALID_CELLID);
  g->m_cell.resize(n);
  g->m_center.resize(n * 3);
  g->m_cellLvl.assign(n, lvl);
  g->m_parent.assign(n, INVALID_CELLID);
  g->m_leaf.resize(n);
  g->m_bndry.resize(n);
  g->m_levels = {lvl};
  for (GInt q = 0; q < n; ++q) {
    const GInt c = ids[q];
    g->m_cell[q] = m_cell[c];
    for (int d = 0; d < 3; ++d) g->m_center[q * 3 + d] = m_center[c * 3 + d];
    g->m_leaf[q] = m_leaf[c];
    g->m_bndry[q] = m_bndry[c];
    g->m_noBndCells += m_bndry[c] && m_leaf[c];
    for (GInt dir = 0; dir < m_stride; ++dir) {
      const GInt nb = m_nghbr[c * m_stride + dir];
      if (nb == INVALID_CELLID) continue;
      if (loc[nb] == INVALID_CELLID) TERMM(-1, "levelGrid: a neighbour link crosses levels");
      g->m_nghbr[q * m_stride + dir] = loc[nb];
    }
    if (m_periodicCell.count(c)) g->m_periodicCell.insert(q);
  }
  for (const auto& [name, srf] : m_surfaces) {
    Surface s(g.get());
    for (GInt c : srf.getCellList())
      if (loc[c] != INVALID_CELLID) s.addCell(loc[c], srf.dirOf(c));
    g->m_surfaces.emplace(name, std::move(s));
  }
  g->m_hasPeriodicLinks = m_hasPeriodicLinks;
  g->setWindow();
  g->checkPeriodicWrap();
  return g;
}

}  // namespace lbmh

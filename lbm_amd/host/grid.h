// Cartesian grid of the drop-in host: generation in the reference's cell
// order, neighbour tables, boundary cells and surfaces, and the structured
// view the device kernels use.
//
// Reference behaviour restated:
//  - root cube: centre = centre of the geometry bounding box, length =
//    (1 + 2^-MAX_LVL) * largest extent, halved per level
//    (src/cartesiangrid_base.h:125-143);
//  - child centre = parent centre + (len_child / 2) * childDir, computed
//    level by level in the same floating-point order
//    (src/gridgenerator/cartesiangrid_generation.h:419-451);
//  - a child is a boundary cell if its parent is and the geometry cuts it;
//    cells that are neither boundary nor (flood-filled) inside are deleted
//    after each refinement (cartesiangrid_generation.h:509-552,578-620);
//  - cell ids = ascending Hilbert index of the centre normalised by the root
//    cube at the partition level (cartesiangrid_generation.h:664-735);
//  - multi-level grids (partitionLevel < uniformLevel <= maxRfnmtLvl): the
//    finer levels are appended after the partition level, every cell of a
//    level refined (uniformRefineGrid) or only its boundary cells
//    (markBndryCells / refineMarkedCells), children in parent order, their
//    neighbours from the parent's (findChildLevelNghbrs), outside cells
//    deleted with the reference's flood fill and copy-last compaction
//    (cartesiangrid_generation.h:151-206,343-417,419-661); every level stays
//    in the grid, linked only within itself, so a solver steps each level as
//    its own lattice;
//  - solver side: boundary cell = cut at its level's length and missing a
//    main neighbour, tested for partition-level cells and children of
//    boundary cells (src/cartesiangrid.h:510-546), surfaces per configured
//    (geometry, face) (src/cartesiangrid.h:553-603), periodic links between
//    surfaces (src/cartesiangrid.h:608-706), diagonal neighbours composed from
//    main neighbours (src/cartesiangrid.h:451-493).
// Deliberate differences: the flood fill that decides "inside" is run over
// the whole candidate set in lexicographic order, which equals the
// reference's result whenever every connected non-boundary region is wholly
// inside or wholly outside the geometry; 3D diagonal neighbours (TERMM in
// the reference) follow the D3Q27 direction order (src/lbm/constants.h:370-397),
// composed axis by axis in x, y, z order.
#pragma once
#include <map>
#include <memory>
#include <unordered_map>

#include "common.h"
#include "geometry.h"

namespace lbmh {

enum class CellProperties : int { bndry = 0, leaf, periodic, NumProperties };

class GridInterface {
 public:
  virtual ~GridInterface() = default;
  [[nodiscard]] virtual auto dim() const -> GInt = 0;
  [[nodiscard]] virtual auto size() const -> GInt = 0;
  [[nodiscard]] virtual auto noCells() const -> GInt = 0;
  [[nodiscard]] virtual auto neighbor(GInt cellId, GInt dir) const -> GInt = 0;
  [[nodiscard]] virtual auto property(GInt cellId, CellProperties p) const -> bool = 0;
  [[nodiscard]] virtual auto cellLength(GInt cellId) const -> GDouble = 0;
  [[nodiscard]] virtual auto lengthOnLvl(GInt lvl) const -> GDouble = 0;
  [[nodiscard]] virtual auto partitionLvl() const -> GInt = 0;
  [[nodiscard]] virtual auto maxLvl() const -> GInt = 0;
  [[nodiscard]] virtual auto boundingBox() const -> const BoundingBox& = 0;
  [[nodiscard]] virtual auto cog() const -> std::vector<GDouble> = 0;
  [[nodiscard]] virtual auto center(GInt cellId, GInt dir) const -> GDouble = 0;
};

// unit vector of a neighbour direction (cartesian::dirVec, include/common/math/cartesian.h:229-253)
auto dirVec(GInt dim, GInt dir) -> std::array<GDouble, 3>;
// number of neighbour slots per cell: 2 / 8 / 26 (cartesian::maxNoNghbrsDiag)
inline auto maxNoNghbrsDiag(GInt dim) -> GInt { return dim == 1 ? 2 : (dim == 2 ? 8 : 26); }

class Surface {
 public:
  Surface() = default;
  explicit Surface(const GridInterface* g) : m_grid(g) {}
  // src/common/surface.h:52-55: append and overwrite the cell's normal
  void addCell(GInt cellId, GInt dir) {
    m_cellId.push_back(cellId);
    m_normal[cellId] = dirVec(m_grid->dim(), dir);
    m_dir[cellId] = dir;
  }
  [[nodiscard]] auto getCellList() const -> const std::vector<GInt>& { return m_cellId; }
  [[nodiscard]] auto normal_p(GInt cellId) const -> const GDouble* { return m_normal.at(cellId).data(); }
  [[nodiscard]] auto dirOf(GInt cellId) const -> GInt { return m_dir.at(cellId); }
  [[nodiscard]] auto size() const -> GInt { return (GInt)m_cellId.size(); }
  [[nodiscard]] auto neighbor(GInt cellId, GInt dir) const -> GInt { return m_grid->neighbor(cellId, dir); }
  void rebind(const GridInterface* g) { m_grid = g; }

 private:
  const GridInterface* m_grid = nullptr;
  std::vector<GInt> m_cellId;
  std::unordered_map<GInt, std::array<GDouble, 3>> m_normal;
  std::unordered_map<GInt, GInt> m_dir;
};

class CartesianGrid : public GridInterface {
 public:
  CartesianGrid() = default;
  // copies the tables (like loadGridInplace copies the generator's grid) and
  // points the surfaces at the copy
  CartesianGrid(const CartesianGrid& o)
      : m_dim(o.m_dim), m_level(o.m_level), m_maxLvl(o.m_maxLvl), m_uniformLvl(o.m_uniformLvl), m_bbox(o.m_bbox),
        m_cog(o.m_cog), m_len(o.m_len),
        m_geom(o.m_geom), m_cell(o.m_cell), m_center(o.m_center), m_cellLvl(o.m_cellLvl), m_parent(o.m_parent),
        m_leaf(o.m_leaf), m_levels(o.m_levels), m_nghbr(o.m_nghbr), m_stride(o.m_stride), m_bndry(o.m_bndry),
        m_noBndCells(o.m_noBndCells), m_surfaces(o.m_surfaces), m_periodicCell(o.m_periodicCell), m_lo(o.m_lo),
        m_n(o.m_n), m_struct(o.m_struct), m_cellOf(o.m_cellOf), m_periodicAxis(o.m_periodicAxis),
        m_hasPeriodicLinks(o.m_hasPeriodicLinks) {
    for (auto& kv : m_surfaces) kv.second.rebind(this);
  }
  auto operator=(const CartesianGrid&) -> CartesianGrid& = delete;
  // Grid generator part: keys dim, maxNoCells, partitionLevel, uniformLevel,
  // maxRfnmtLvl, geometry, boundingBox (src/gridgenerator/gridGenerator.cpp:107-286).
  void generate(Configuration& cfg);
  // transformMaxRfnmtLvlToExtent for alignNodesWithSurface
  void alignWithSurface(GInt alignDir);
  // Solver part (CartesianGrid::loadGridInplace, src/cartesiangrid.h:264-320):
  // boundary cells, surfaces from solver.boundary, periodic links, diagonals.
  void loadGridInplace(Configuration& solverCfg);

  [[nodiscard]] auto dim() const -> GInt override { return m_dim; }
  [[nodiscard]] auto size() const -> GInt override { return (GInt)m_cell.size(); }
  [[nodiscard]] auto noCells() const -> GInt override { return size(); }
  [[nodiscard]] auto neighbor(GInt cellId, GInt dir) const -> GInt override {
    return m_nghbr[cellId * m_stride + dir];
  }
  [[nodiscard]] auto property(GInt cellId, CellProperties p) const -> bool override {
    if (p == CellProperties::bndry) return m_bndry[cellId] != 0;
    if (p == CellProperties::leaf) return m_leaf[cellId] != 0;
    return m_periodicCell.count(cellId) > 0;
  }
  [[nodiscard]] auto cellLength(GInt cellId) const -> GDouble override { return m_len[m_cellLvl[cellId]]; }
  [[nodiscard]] auto lengthOnLvl(GInt lvl) const -> GDouble override { return m_len.at(lvl); }
  [[nodiscard]] auto partitionLvl() const -> GInt override { return m_level; }
  [[nodiscard]] auto maxLvl() const -> GInt override { return m_maxLvl; }
  [[nodiscard]] auto uniformLvl() const -> GInt { return m_uniformLvl; }
  [[nodiscard]] auto level(GInt cellId) const -> GInt { return m_cellLvl[cellId]; }
  [[nodiscard]] auto parent(GInt cellId) const -> GInt { return m_parent[cellId]; }
  // levels that hold cells, ascending (one entry for a single-level grid)
  [[nodiscard]] auto levels() const -> const std::vector<GInt>& { return m_levels; }
  // the cells of level lvl as a single-level grid of their own (neighbours,
  // boundary flags, surfaces and periodic links restricted to the level, a
  // structured window); ids[q] = this grid's id of its cell q
  [[nodiscard]] auto levelGrid(GInt lvl, std::vector<GInt>& ids) const -> std::unique_ptr<CartesianGrid>;
  [[nodiscard]] auto boundingBox() const -> const BoundingBox& override { return m_bbox; }
  [[nodiscard]] auto cog() const -> std::vector<GDouble> override { return {m_cog.begin(), m_cog.begin() + m_dim}; }
  [[nodiscard]] auto center(GInt cellId, GInt dir) const -> GDouble override { return m_center[cellId * 3 + dir]; }
  [[nodiscard]] auto noBndCells() const -> GInt { return m_noBndCells; }
  [[nodiscard]] auto geometry() const -> const GeometryManager& { return *m_geom; }

  [[nodiscard]] auto hasBndrySurface(const std::string& n) const -> bool { return m_surfaces.count(n) > 0; }
  auto bndrySurface(const std::string& n) -> Surface& {
    if (!m_surfaces.count(n)) TERMM(-1, "Invalid bndryId \"" + n + "\"");
    return m_surfaces.at(n);
  }
  [[nodiscard]] auto surfaces() const -> const std::map<std::string, Surface>& { return m_surfaces; }
  [[nodiscard]] auto neighborStride() const -> GInt { return m_stride; }

  // structured view: window of level indices [lo, lo+n) per axis (single-level
  // grids; a multi-level grid has one per level, through levelGrid)
  [[nodiscard]] auto extent(int a) const -> GInt { return m_n[a]; }
  [[nodiscard]] auto structIndex(GInt cellId) const -> GInt { return m_struct[cellId]; }
  [[nodiscard]] auto cellAt(GInt s) const -> GInt { return m_cellOf[s]; }
  [[nodiscard]] auto structSize() const -> GInt { return m_n[0] * m_n[1] * m_n[2]; }
  // periodic links map exactly onto a wrap of axis a
  [[nodiscard]] auto periodicAxis(int a) const -> bool { return m_periodicAxis[a]; }

 private:
  void setPeriodicConnections(Configuration& solverCfg);
  void addPeriodicConnection(const Surface& A, const Surface& B);
  void addDiagonalNghbrs();
  void checkPeriodicWrap();
  void setWindow();
  struct Refinement;
  void refineLevel(GInt lvl, bool onlyBndry, GInt maxNoCells, Refinement& st);

  int m_dim = 0;
  GInt m_level = 0;   // partition level
  GInt m_maxLvl = 0;  // maxRfnmtLvl
  GInt m_uniformLvl = 0;
  BoundingBox m_bbox;
  std::array<GDouble, 3> m_cog{{0, 0, 0}};
  std::vector<GDouble> m_len;
  std::shared_ptr<GeometryManager> m_geom;
  // cells in id order: level indices and centres
  std::vector<std::array<GInt, 3>> m_cell;
  std::vector<GDouble> m_center;
  std::vector<GInt> m_cellLvl, m_parent;
  std::vector<uint8_t> m_leaf;
  std::vector<GInt> m_levels;
  std::vector<GInt> m_nghbr;  // [cell][m_stride]
  GInt m_stride = 0;
  std::vector<uint8_t> m_bndry;
  GInt m_noBndCells = 0;
  std::map<std::string, Surface> m_surfaces;
  std::set<GInt> m_periodicCell;
  // structured window
  std::array<GInt, 3> m_lo{{0, 0, 0}}, m_n{{1, 1, 1}};
  std::vector<GInt> m_struct, m_cellOf;
  std::array<bool, 3> m_periodicAxis{{false, false, false}};
  bool m_hasPeriodicLinks = false;
};

}  // namespace lbmh

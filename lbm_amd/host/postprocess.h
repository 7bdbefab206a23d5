// Postprocessing hooks of the solver and the line function, the reference's
// only one (Postprocess, src/postprocess/postprocessing.h:45-127;
// PostprocessCartesianFunctionLine, src/postprocess/postprocessing_cartesian.h:9-65).
//
// A "line" function keeps the cells whose centre lies within half a cell
// length of the hyperplane through A and B (Line::distance, src/common/line.h:
// Eigen::Hyperplane::Through, whose normal is the unit vector orthogonal to
// B - A that Eigen's unitOrthogonal picks; in 3D that is one plane containing
// the line, not the line), sorted by their centres.  At its hook the u
// component of every kept cell is written with its centre to line.csv in the
// working directory (ASCII::writePointsCSV, src/common/IO.h:37-97): header
// "x,y[,z],u", coordinates in 15 significant digits, u in fixed notation with
// 15 decimals.  Every function of a hook writes the same file, so the last
// one in key order remains.
#pragma once

#include <array>
#include <string>
#include <vector>

#include "common.h"
#include "grid.h"

namespace lbmh {

enum class PPHook { AtStart = 0, BeforeTimestep, AfterTimestep, AtEnd, Num };

// unit normal and offset of the hyperplane through a and b (dim 2 or 3)
void lineHyperplane(int dim, const double* a, const double* b, double* normal, double* offset);

// cells of the grid within half a cell length of the hyperplane, sorted by
// centre (x, then y, then z)
auto lineCells(const CartesianGrid& grid, const double* a, const double* b) -> std::vector<GInt>;

class Postprocess {
 public:
  // conf: the solver's "postprocessing" object (absent or null: inactive)
  void init(const json& conf, const CartesianGrid& grid);
  [[nodiscard]] auto active(PPHook h) const -> bool { return !m_cells[(int)h].empty(); }
  // u: the u component of every cell, in cell-id order
  void execute(PPHook h, const CartesianGrid& grid, const std::vector<double>& u) const;

 private:
  std::array<std::vector<std::vector<GInt>>, (int)PPHook::Num> m_cells;
};

}  // namespace lbmh

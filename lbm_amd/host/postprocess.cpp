#include "postprocess.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>

namespace lbmh {

void lineHyperplane(int dim, const double* a, const double* b, double* n, double* offset) {
  const double d[3] = {b[0] - a[0], b[1] - a[1], dim == 3 ? b[2] - a[2] : 0.0};
  if (dim == 2) {
    // (-dy, dx), normalised by dividing by its norm
    n[0] = -d[1];
    n[1] = d[0];
    const double z = n[0] * n[0] + n[1] * n[1];
    if (z > 0) {
      const double s = std::sqrt(z);
      n[0] /= s;
      n[1] /= s;
    }
    n[2] = 0;
  } else {
    // a unit vector orthogonal to d: in the xy plane unless d is (nearly)
    // parallel to z (relative precision 1e-12)
    const double prec = 1e-12;
    const bool xSmall = std::abs(d[0]) <= std::abs(d[2]) * prec;
    const bool ySmall = std::abs(d[1]) <= std::abs(d[2]) * prec;
    if (!xSmall || !ySmall) {
      const double inv = 1.0 / std::sqrt(d[0] * d[0] + d[1] * d[1]);
      n[0] = -d[1] * inv;
      n[1] = d[0] * inv;
      n[2] = 0;
    } else {
      const double inv = 1.0 / std::sqrt(d[1] * d[1] + d[2] * d[2]);
      n[0] = 0;
      n[1] = -d[2] * inv;
      n[2] = d[1] * inv;
    }
  }
  double dot = 0;
  for (int k = 0; k < dim; ++k) dot += a[k] * n[k];
  *offset = -dot;
}

auto lineCells(const CartesianGrid& grid, const double* a, const double* b) -> std::vector<GInt> {
  const int dim = (int)grid.dim();
  double n[3], off;
  lineHyperplane(dim, a, b, n, &off);
  std::vector<GInt> cells;
  for (GInt c = 0; c < grid.size(); ++c) {
    double dot = 0;
    for (int k = 0; k < dim; ++k) dot += n[k] * grid.center(c, k);
    const double distance = std::abs(dot + off);
    if (0.5 * grid.cellLength(c) >= distance) cells.push_back(c);
  }
  // The reference's comparator (any coordinate smaller) is an order only
  // when the kept cells differ in one coordinate, as for an axis-parallel
  // line through one column of cells; there both orders agree.
  std::sort(cells.begin(), cells.end(), [&](GInt p, GInt q) {
    for (int k = 0; k < dim; ++k) {
      const double cp = grid.center(p, k), cq = grid.center(q, k);
      if (cp != cq) return cp < cq;
    }
    return p < q;
  });
  return cells;
}

void Postprocess::init(const json& conf, const CartesianGrid& grid) {
  if (conf.is_null() || !conf.is_object()) return;
  const int dim = (int)grid.dim();
  for (auto it = conf.begin(); it != conf.end(); ++it) {
    const json& f = it.value();
    if (!f.count("type")) TERMM(-1, "The required configuration value is missing: type");
    if (!f.count("execute")) TERMM(-1, "The required configuration value is missing: execute");
    const std::string type = f["type"].get<std::string>();
    const std::string exe = f["execute"].get<std::string>();
    if (type != "line") TERMM(-1, "Invalid functype");
    PPHook h;
    if (exe == "atStart") h = PPHook::AtStart;
    else if (exe == "beforeTimestep") h = PPHook::BeforeTimestep;
    else if (exe == "afterTimestep") h = PPHook::AfterTimestep;
    else if (exe == "atEnd") h = PPHook::AtEnd;
    else TERMM(-1, "Invalid hook");
    if (!f.count("A") || !f.count("B")) TERMM(-1, "The required configuration value is missing: A/B");
    const auto A = f["A"].get<std::vector<double>>(), B = f["B"].get<std::vector<double>>();
    if ((int)A.size() != dim || (int)B.size() != dim) TERMM(-1, "Invalid line end points: need " + std::to_string(dim) + " coordinates");
    std::cerr << "PP: Adding postprocessing function " << type << " for LBM solver at hook" << exe << std::endl;
    m_cells[(int)h].push_back(lineCells(grid, A.data(), B.data()));
  }
}

void Postprocess::execute(PPHook h, const CartesianGrid& grid, const std::vector<double>& u) const {
  const auto& funcs = m_cells[(int)h];
  if (funcs.empty()) return;
  static const char* hookName[] = {"atStart", "beforeTimestep", "afterTimestep", "atEnd"};
  std::cerr << "Executing postprocessing at hook:" << hookName[(int)h] << std::endl;
  const int dim = (int)grid.dim();
  for (const auto& cells : funcs) {
    std::cerr << " Writing line.csv" << std::endl;
    std::ofstream out("line.csv");
    if (!out) {
      std::cerr << "WARNING: cannot write line.csv" << std::endl;
      continue;
    }
    static const char axis[] = {'x', 'y', 'z'};
    for (int k = 0; k < dim; ++k) out << axis[k] << (k + 1 < dim ? "," : "");
    out << ",u\n";
    char buf[400];
    for (GInt c : cells) {
      for (int k = 0; k < dim; ++k) {
        std::snprintf(buf, sizeof(buf), "%.15g", grid.center(c, k));
        out << buf << (k + 1 < dim ? "," : "");
      }
      std::snprintf(buf, sizeof(buf), "%.15f", u[c]);
      out << "," << buf << "\n";
    }
  }
}

}  // namespace lbmh

// C ABI of the host layer (include/lbm_host.h).
#include <cstring>

#include "../../include/lbm_host.h"
#include "expr.h"
#include "poisson.h"
#include "solver.h"
#include "vtk.h"

using namespace lbmh;

struct lbm_host_grid {
  std::unique_ptr<CartesianGrid> grid;
};

static thread_local std::string g_host_err;

extern "C" {

int lbm_host_grid_create(const char* json_text, lbm_host_grid** out) {
  if (!json_text || !out) {
    g_host_err = "null argument";
    return -1;
  }
  json cfg;
  try {
    cfg = json::parse(json_text);
  } catch (const std::exception& e) {
    g_host_err = e.what();
    return -1;
  }
  auto* g = new lbm_host_grid();
  g->grid = std::make_unique<CartesianGrid>();
  Configuration c(cfg);
  g->grid->generate(c);
  Configuration sc(cfg.count("solver") ? cfg["solver"] : json::object(), "solver.");
  g->grid->loadGridInplace(sc);
  *out = g;
  return 0;
}

int64_t lbm_host_grid_ncells(const lbm_host_grid* g) { return g->grid->size(); }
int32_t lbm_host_grid_dim(const lbm_host_grid* g) { return (int32_t)g->grid->dim(); }
int32_t lbm_host_grid_stride(const lbm_host_grid* g) { return (int32_t)g->grid->neighborStride(); }

int lbm_host_grid_centers(const lbm_host_grid* g, double* out) {
  const GInt d = g->grid->dim();
  for (GInt q = 0; q < g->grid->size(); ++q)
    for (GInt a = 0; a < d; ++a) out[q * d + a] = g->grid->center(q, a);
  return 0;
}

int lbm_host_grid_neighbors(const lbm_host_grid* g, int64_t* out) {
  const GInt s = g->grid->neighborStride();
  for (GInt q = 0; q < g->grid->size(); ++q)
    for (GInt k = 0; k < s; ++k) out[q * s + k] = g->grid->neighbor(q, k);
  return 0;
}

int lbm_host_grid_bndry(const lbm_host_grid* g, uint8_t* out) {
  for (GInt q = 0; q < g->grid->size(); ++q) out[q] = g->grid->property(q, CellProperties::bndry) ? 1 : 0;
  return 0;
}

int lbm_host_grid_structured(const lbm_host_grid* g, int64_t* nxyz, int64_t* index) {
  if (g->grid->levels().size() != 1) {
    g_host_err = "lbm_host_grid_structured: a multi-level grid has one structured window per level";
    return -1;
  }
  for (int a = 0; a < 3; ++a) nxyz[a] = g->grid->extent(a);
  if (index)
    for (GInt q = 0; q < g->grid->size(); ++q) index[q] = g->grid->structIndex(q);
  return 0;
}

int lbm_host_grid_levels(const lbm_host_grid* g, int32_t* level, int64_t* parent) {
  for (GInt q = 0; q < g->grid->size(); ++q) {
    if (level) level[q] = (int32_t)g->grid->level(q);
    if (parent) parent[q] = g->grid->parent(q);
  }
  return 0;
}

int32_t lbm_host_grid_surface_count(const lbm_host_grid* g) { return (int32_t)g->grid->surfaces().size(); }

int lbm_host_grid_surface(const lbm_host_grid* g, int32_t i, char* name, int32_t name_len, int64_t* ncell,
                          int64_t* cells, int64_t* dirs) {
  const auto& m = g->grid->surfaces();
  if (i < 0 || i >= (int32_t)m.size()) return -1;
  auto it = m.begin();
  std::advance(it, i);
  if (name && name_len > 0) {
    std::strncpy(name, it->first.c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  const auto& list = it->second.getCellList();
  if (ncell) *ncell = (int64_t)list.size();
  if (cells)
    for (size_t k = 0; k < list.size(); ++k) cells[k] = list[k];
  if (dirs)
    for (size_t k = 0; k < list.size(); ++k) dirs[k] = it->second.dirOf(list[k]);
  return 0;
}

int64_t lbm_host_grid_line_cells(const lbm_host_grid* g, const double* a, const double* b, int64_t* out, int64_t cap) {
  if (!g || !a || !b || cap < 0 || (cap > 0 && !out) || g->grid->dim() < 2) {
    g_host_err = "lbm_host_grid_line_cells: invalid argument";
    return -1;
  }
  const std::vector<GInt> cells = lineCells(*g->grid, a, b);
  for (size_t i = 0; i < cells.size() && (int64_t)i < cap; ++i) out[i] = cells[i];
  return (int64_t)cells.size();
}

void lbm_host_grid_destroy(lbm_host_grid* g) { delete g; }

int lbm_host_run_file(const char* config_file, const char* storage, int64_t n, double* rho, double* u,
                      int64_t* steps_done) {
  SolverOptions opt;
  if (storage) opt.storage = storage;
  opt.quiet = true;
  GridGenerator gen;
  gen.init(0, nullptr, config_file);
  if (gen.run() != 0) return -1;
  LBMSolver s(opt);
  s.init(0, nullptr, config_file);
  s.takeGrid(gen.releaseGrid());
  const GInt rc = s.run();
  if (steps_done) *steps_done = s.stepsDone();
  const GInt N = (GInt)s.density().size();
  if (rho && u && n == N) {
    const GInt d = (GInt)s.velocity().size() / std::max<GInt>(N, 1);
    for (GInt q = 0; q < N; ++q) {
      rho[q] = s.density()[q];
      for (GInt a = 0; a < d; ++a) u[q * d + a] = s.velocity()[q * d + a];
    }
  }
  return (int)rc;
}

int lbm_host_eval_expression(const char* text, double x, double y, double z, double* out) {
  if (!text || !out) {
    g_host_err = "lbm_host_eval_expression: null argument";
    return -1;
  }
  try {
    *out = expr::Expression(text).eval(x, y, z);
  } catch (const std::exception& e) {
    g_host_err = e.what();
    return -1;
  }
  return 0;
}

int lbm_host_run_poisson_file(const char* config_file, int64_t n, double* phi, int64_t* steps_done) {
  SolverOptions opt;
  opt.quiet = true;
  GridGenerator gen;
  gen.init(0, nullptr, config_file);
  if (gen.run() != 0) return -1;
  PoissonSolver s(opt);
  s.init(0, nullptr, config_file);
  s.takeGrid(gen.releaseGrid());
  const GInt rc = s.run();
  if (steps_done) *steps_done = s.stepsDone();
  if (phi && n == (int64_t)s.potential().size())
    for (GInt q = 0; q < n; ++q) phi[q] = s.potential()[q];
  return (int)rc;
}

int lbm_host_write_points_vtk(const char* file_base, int32_t dim, int64_t n, const double* coords, int32_t nfields,
                              const char* const* names, const double* const* values) {
  if (!file_base || dim < 1 || dim > 3 || n < 1 || !coords || nfields < 0 || (nfields > 0 && (!names || !values))) {
    g_host_err = "lbm_host_write_points_vtk: invalid argument";
    return -1;
  }
  std::vector<vtk::Field> fields;
  for (int32_t k = 0; k < nfields; ++k) fields.push_back({names[k], std::vector<double>(values[k], values[k] + n)});
  if (!vtk::writePoints(file_base, dim, n, coords, fields)) {
    g_host_err = std::string("lbm_host_write_points_vtk: cannot write ") + file_base + ".vtp";
    return -1;
  }
  return 0;
}

double lbm_host_text_round_trip(double v) { return vtk::textRoundTrip(v); }

const char* lbm_host_last_error(void) { return g_host_err.c_str(); }

}  // extern "C"

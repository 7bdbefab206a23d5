// Drop-in host solver: the reference's Runnable / GridGenerator / LBMSolver
// shapes (src/interface/solver_interface.h:6-24, src/lbm/solver.h:13-230)
// driving the device step library through its C ABI (include/lbm_hip.h).
#pragma once
#include <chrono>
#include <memory>

#include "../../include/lbm_hip.h"
#include "bnd.h"
#include "grid.h"
#include "postprocess.h"

namespace lbmh {

class Runnable {
 public:
  virtual ~Runnable() = default;
  virtual void init(int argc, char** argv, std::string config_file) = 0;
  virtual void initBenchmark(int argc, char** argv) = 0;
  virtual auto run() -> GInt = 0;
  [[nodiscard]] virtual auto grid() const -> const GridInterface& = 0;
  virtual void transferGrid(const GridInterface& grid) = 0;
};

class GridGenerator : public Runnable {
 public:
  void init(int argc, char** argv, std::string config_file) override;
  void initBenchmark(int argc, char** argv) override;
  auto run() -> GInt override;
  [[nodiscard]] auto grid() const -> const GridInterface& override { return *m_grid; }
  void transferGrid(const GridInterface&) override { TERMM(-1, "Not implemented!"); }
  auto releaseGrid() -> std::unique_ptr<CartesianGrid> { return std::move(m_grid); }

 private:
  json m_config;
  std::unique_ptr<CartesianGrid> m_grid;
};

// create every directory of path (output directories)
void mkdirs(const std::string& path);

struct SolverOptions {
  std::string storage = "";  // overrides the 'storage' solver key when set
  bool quiet = false;
};

class LBMSolver : public Runnable {
 public:
  explicit LBMSolver(SolverOptions opt = {}) : m_opt(std::move(opt)) {}
  ~LBMSolver() override;
  void init(int argc, char** argv, std::string config_file) override;
  // `lbm --bench`: D3Q19 BGK periodic 256^3 Taylor-Green box on the GPU
  void initBenchmark(int argc, char** argv) override;
  auto run() -> GInt override;
  [[nodiscard]] auto grid() const -> const GridInterface& override { return *m_grid; }
  void transferGrid(const GridInterface& grid) override;
  void takeGrid(std::unique_ptr<CartesianGrid> g);

  // results in reference cell-id order (after run)
  [[nodiscard]] auto velocity() const -> const std::vector<GDouble>& { return m_u; }
  [[nodiscard]] auto density() const -> const std::vector<GDouble>& { return m_rho; }
  [[nodiscard]] auto stepsDone() const -> GInt { return m_timeStep; }

 private:
  void loadConfiguration();
  void createLattices();
  void initialCondition();
  auto convergenceCondition(const lbm_sums& s) -> bool;
  void fetchMacros();
  auto stepAll(int mode, lbm_sums* sums) -> int;
  void compareToAnalyticalResult();
  void output(bool forced, const std::string& postfix);
  auto runBenchmark() -> GInt;
  void check(int rc, const char* what, const lbm_handle* h = nullptr);

  SolverOptions m_opt;
  json m_config;
  std::unique_ptr<Configuration> m_solverCfg;
  std::unique_ptr<CartesianGrid> m_grid;
  LBMBndManager m_bnd;
  lbm_handle* m_h = nullptr;  // `lbm --bench`
  // One device lattice per grid level (a single-level grid has one; the
  // reference steps every level of a multi-level grid as its own lattice,
  // its cells linked only within the level): the level's grid and boundary
  // objects (multi-level grids; a single-level lattice uses m_grid and
  // m_bnd), ids[q] = m_grid's id of the lattice's cell q (empty: the same)
  struct Lattice {
    std::unique_ptr<CartesianGrid> grid;
    std::vector<GInt> ids;
    LBMBndManager bnd;
    lbm_handle* h = nullptr;
  };
  std::vector<Lattice> m_lat;
  [[nodiscard]] auto latGrid(const Lattice& L) const -> const CartesianGrid& { return L.grid ? *L.grid : *m_grid; }
  [[nodiscard]] static auto latId(const Lattice& L, GInt q) -> GInt { return L.ids.empty() ? q : L.ids[q]; }
  bool m_benchmark = false;

  // configuration (src/lbm/solver.cpp:71-163)
  GInt m_dim = 2;
  GInt m_lattice = 9;
  GInt m_infoInterval = 10, m_convInterval = 10, m_keepAlive = 60, m_maxSteps = 0, m_solutionInterval = 100;
  std::string m_outputDir = "out/", m_solutionFile = "solution";
  GDouble m_refLength = 1.0, m_ma = 0.01, m_relaxTime = 0.9, m_omega = 1.0 / 0.9, m_nu = 0, m_re = 1, m_dt = NAN;
  GDouble m_convergence = 1e-12;
  std::array<GDouble, 3> m_force{{0, 0, 0}};
  bool m_forcing = false;  // the reference's "forcing" inlet/outlet correction
  GDouble m_forcingGrad = 0;
  int m_storage = LBM_STORE_F32;
  GInt m_timeStep = 0;
  bool m_diverged = false, m_wroteDiverged = false;
  std::vector<uint8_t> m_outputCell;  // the output filter's verdict per cell
  std::vector<GDouble> m_vars;  // host initial values [cell][dim+1]
  std::vector<GDouble> m_rho, m_u;
  json m_ppConf;
  Postprocess m_pp;
  std::vector<GDouble> m_ppU;  // u component the postprocessing functions read
};

}  // namespace lbmh

// Host side of the Poisson solver: the reference's LBMSolver instantiated
// for LBEquationType::Poisson (D1Q3, D2Q5, D2Q9; src/lbm/solverExe.h:37-80)
// with its configuration keys, boundary objects, run loop, output and
// analytic comparison, driving the device step library through
// include/lbm_poisson.h.
#pragma once
#include <memory>
#include <vector>

#include "../../include/lbm_poisson.h"
#include "solver.h"

namespace lbmh {

class PoissonSolver : public Runnable {
 public:
  explicit PoissonSolver(SolverOptions opt = {}) : m_opt(std::move(opt)) {}
  ~PoissonSolver() override;
  void init(int argc, char** argv, std::string config_file) override;
  void initBenchmark(int, char**) override { TERMM(-1, "Not implemented!"); }
  auto run() -> GInt override;
  [[nodiscard]] auto grid() const -> const GridInterface& override { return *m_grid; }
  void transferGrid(const GridInterface& grid) override;
  void takeGrid(std::unique_ptr<CartesianGrid> g);

  // electric potential in reference cell-id order (after run)
  [[nodiscard]] auto potential() const -> const std::vector<GDouble>& { return m_phi; }
  [[nodiscard]] auto stepsDone() const -> GInt { return m_timeStep; }

 private:
  struct Bnd {
    std::string surface;
    int kind = 0;  // LBM_POISSON_DIRICHLET / LBM_POISSON_NEUMANN
    int order = 0;
    std::vector<GDouble> value;  // one per surface cell (Dirichlet)
  };
  void loadConfiguration();
  void setupBndryCnds(const json& cfg);
  void fetchPotential();
  void compareToAnalyticalResult();
  void output(bool forced, const std::string& postfix);
  void check(int rc, const char* what);

  SolverOptions m_opt;
  json m_config;
  std::unique_ptr<Configuration> m_solverCfg;
  std::unique_ptr<CartesianGrid> m_grid;
  lbm_poisson* m_h = nullptr;
  std::vector<Bnd> m_bnd;
  GInt m_dim = 1, m_lattice = 3;
  GInt m_infoInterval = 10, m_convInterval = 10, m_maxSteps = 0, m_solutionInterval = 100;
  std::string m_outputDir = "out/", m_solutionFile = "solution";
  GDouble m_relaxTime = 1, m_omega = 1, m_dt = 1, m_kappa = 27.79, m_source = 0, m_convergence = 1e-12;
  GInt m_timeStep = 0;
  bool m_diverged = false, m_wroteDiverged = false, m_outputAll = true;
  std::vector<GDouble> m_phi;
  json m_ppConf;
  Postprocess m_pp;
};

// "equation" of a configuration file's solver section ("navierstokes" when absent)
auto solverEquation(const json& cfg) -> std::string;

}  // namespace lbmh

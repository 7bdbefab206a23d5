// `lbm` command line: the reference's CLI (src/main.cpp:154-304) on top of
// the HIP step library.
//   lbm [-c] <config.json> [-d level] [-b] [-s] [-p storage] [-h] [-v]
// Runs the grid generator, then the LBM solver when solver.type is "lbm"
// (or -s is given).  --bench runs a D3Q19 benchmark on the GPU (the
// reference's --bench aborts, src/lbm/solver.cpp:39-46).  -p/--storage
// (new) overrides the population storage: f32 (default), f64, f16, bf16.
// Exit status: sum of the runners' returns; fatal errors exit 255.
#include <cstring>

#include "poisson.h"
#include "solver.h"

using namespace lbmh;

namespace {
void usage() {
  std::cout << "LBM Solver (MI355X)\n"
               "Usage: lbm [options] <config.json>\n"
               "  -c, --config FILE    configuration file (default grid.json)\n"
               "  -d, --debug LEVEL    debug level 0..2 (accepted; checks are always on)\n"
               "  -b, --bench          run the D3Q19 benchmark\n"
               "  -s, --solver         run the LBM solver after the grid generator\n"
               "  -p, --storage KIND   population storage f32|f64|f16|bf16\n"
               "  -h, --help           print usage\n"
               "  -v, --version        version information\n";
}
}  // namespace

int main(int argc, char** argv) {
  std::string config = "grid.json";
  bool bench = false, forceSolver = false;
  SolverOptions opt;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) TERMM(-1, "missing value for " + a);
      return argv[++i];
    };
    if (a == "-h" || a == "--help") {
      usage();
      return 0;
    } else if (a == "-v" || a == "--version") {
      std::cout << "0.0.2 (MI355X build)" << std::endl;
      return 0;
    } else if (a == "-c" || a == "--config") {
      config = next();
    } else if (a == "-d" || a == "--debug") {
      const int lvl = std::atoi(next().c_str());
      if (lvl > 0) std::cout << "Activated debug level " << std::min(lvl, 2) << std::endl;
    } else if (a == "-b" || a == "--bench") {
      bench = true;
    } else if (a == "-s" || a == "--solver") {
      forceSolver = true;
    } else if (a == "-p" || a == "--storage") {
      opt.storage = next();
    } else if (!a.empty() && a[0] == '-') {
      TERMM(-1, "Unknown option " + a);
    } else {
      config = a;
    }
  }
  if (bench) {
    LBMSolver s(opt);
    s.initBenchmark(argc, argv);
    return (int)s.run();
  }
  {
    std::ifstream f(config);
    if (!f) TERMM(-1, "Configuration file not found: " + config);
  }
  json cfg = Configuration::fromFile(config);
  bool runLBM = forceSolver;
  if (cfg.count("solver") && cfg["solver"].count("type")) {
    const auto t = cfg["solver"]["type"].get<std::string>();
    runLBM = runLBM || t == "lbm" || t == "LBM";
  }
  std::cerr << "running gridder " << std::endl;
  GridGenerator gen;
  gen.init(argc, argv, config);
  GInt ret = gen.run();
  if (ret == 0 && runLBM) {
    // LBMSolverExecutor::init (src/lbm/solverExe.h:17-93): the solver
    // section's "equation" picks the Navier-Stokes or the Poisson solver
    if (solverEquation(cfg) == "poisson") {
      PoissonSolver s(opt);
      s.init(argc, argv, config);
      s.takeGrid(gen.releaseGrid());
      ret += s.run();
    } else {
      LBMSolver s(opt);
      s.init(argc, argv, config);
      s.takeGrid(gen.releaseGrid());
      ret += s.run();
    }
  }
  return (int)ret;
}

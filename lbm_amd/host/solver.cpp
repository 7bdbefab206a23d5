#include "solver.h"

#include <sys/stat.h>

#include "vtk.h"

#include <cstring>
#include <set>

namespace lbmh {

namespace {
constexpr GDouble lbm_cs = 0.57735026918962576451;  // 1/sqrt(3), src/lbm/constants.h:26

auto storageFromString(const std::string& s) -> int {
  if (s == "f32" || s == "float32" || s == "fp32") return LBM_STORE_F32;
  if (s == "f64" || s == "float64" || s == "fp64") return LBM_STORE_F64;
  if (s == "f16" || s == "float16" || s == "fp16") return LBM_STORE_F16;
  if (s == "bf16" || s == "bfloat16") return LBM_STORE_BF16;
  TERMM(-1, "Invalid storage '" + s + "' (f32, f64, f16, bf16)");
}

}  // namespace

void mkdirs(const std::string& path) {
  std::string cur;
  for (size_t i = 0; i < path.size(); ++i) {
    cur += path[i];
    if (path[i] == '/' || i + 1 == path.size()) mkdir(cur.c_str(), 0755);
  }
}

// ------------------------------------------------------------- grid generator
void GridGenerator::init(int, char**, std::string config_file) { m_config = Configuration::fromFile(config_file); }

void GridGenerator::initBenchmark(int, char**) {
  // the reference's generator benchmark has no configuration to run on
  // (src/gridgenerator/gridGenerator.cpp:43-54,129); a fixed 3D box is used
  m_config = {{"dim", 3},
              {"maxNoCells", 1 << 24},
              {"partitionLevel", 5},
              {"uniformLevel", 5},
              {"maxRfnmtLvl", 5},
              {"geometry", {{"cube", {{"type", "cube"}, {"center", {0.0, 0.0, 0.0}}, {"length", 1}}}}}};
}

auto GridGenerator::run() -> GInt {
  Configuration cfg(m_config);
  m_grid = std::make_unique<CartesianGrid>();
  m_grid->generate(cfg);
  std::cout << "  * grid has " << m_grid->size() << " cells" << std::endl;
  return 0;
}

// --------------------------------------------------------------- LBM solver
LBMSolver::~LBMSolver() {
  if (m_h) lbm_destroy(m_h);
  for (auto& L : m_lat)
    if (L.h) lbm_destroy(L.h);
}

void LBMSolver::check(int rc, const char* what, const lbm_handle* h) {
  if (rc != LBM_OK) TERMM(-1, std::string(what) + ": " + lbm_last_error(h ? h : m_h));
}

void LBMSolver::init(int, char**, std::string config_file) {
  m_config = Configuration::fromFile(config_file);
  if (!m_config.count("solver")) TERMM(-1, "No solver section in " + config_file);
  m_solverCfg = std::make_unique<Configuration>(m_config["solver"], "solver.");
}

void LBMSolver::initBenchmark(int, char**) { m_benchmark = true; }

void LBMSolver::transferGrid(const GridInterface& grid) {
  const auto* cg = dynamic_cast<const CartesianGrid*>(&grid);
  if (!cg) TERMM(-1, "Invalid grid type");
  takeGrid(std::make_unique<CartesianGrid>(*cg));
}

void LBMSolver::takeGrid(std::unique_ptr<CartesianGrid> g) {
  if (!m_solverCfg) TERMM(-1, "init() must be called before the grid is transferred");
  m_grid = std::move(g);
  std::cerr << "Transferring " << m_grid->dim() << "D Grid to LBM solver" << std::endl;
  m_grid->loadGridInplace(*m_solverCfg);
}

// LBMSolver::loadConfiguration, src/lbm/solver.cpp:71-163
void LBMSolver::loadConfiguration() {
  Configuration& c = *m_solverCfg;
  const std::string method = c.opt_config_value<std::string>("method", "bgk");
  if (method != "bgk") TERMM(-1, "Invalid equation configuration!");
  c.opt_config_value<std::string>("type", "lbm");
  m_dim = m_grid->dim();
  const std::string model = c.opt_config_value<std::string>("model", m_dim == 2 ? "D2Q9" : "D3Q19");
  if (model == "D2Q9") m_lattice = 9;
  else if (model == "D3Q19") m_lattice = 19;
  else TERMM(-1, "Unsupported model " + model + " (D2Q9 and D3Q19 are built)");
  if ((m_lattice == 9 && m_dim != 2) || (m_lattice == 19 && m_dim != 3))
    TERMM(-1, "Invalid configuration the grid dimensionality is not matching!");
  const std::string equation = c.opt_config_value<std::string>("equation", "navierstokes");
  if (equation != "navierstokes") TERMM(-1, "Only the navierstokes equation is part of this build");
  m_infoInterval = c.opt_config_value<GInt>("info_interval", 10);
  m_convInterval = c.opt_config_value<GInt>("conv_interval", m_infoInterval);
  if (m_convInterval < 1) TERMM(-1, "conv_interval must be >= 1");
  m_keepAlive = c.opt_config_value<GInt>("keep_alive_time", 60);
  m_maxSteps = c.required_config_value<GInt>("maxSteps");
  m_outputDir = c.opt_config_value<std::string>("output_dir", "out/");
  if (m_outputDir.empty() || m_outputDir.back() != '/') m_outputDir += '/';
  m_solutionInterval = c.opt_config_value<GInt>("solution_interval", 100);
  if (m_solutionInterval < 1) TERMM(-1, "solution_interval must be >= 1");
  m_solutionFile = c.opt_config_value<std::string>("solution_filename", "solution");
  // output filter (CellFilterManager, src/cell_filter.h:64-103), default
  // {"cellFilter": "leafCells"} (src/lbm/solver.cpp:86-87): leaf cells,
  // the highest level, the partition level or a given level
  {
    json fc = c.getObject("cellFilter");
    if (fc.empty()) fc = json{{"cellFilter", "leafCells"}};
    if (!fc.count("cellFilter")) TERMM(-1, "Invalid output configuration missing \"cellFilter\" key");
    if (fc["cellFilter"].is_array()) TERMM(-1, "untested");
    const std::string f = fc["cellFilter"].get<std::string>();
    const GInt N = m_grid->size();
    m_outputCell.assign(N, 1);
    GInt lvl = -1;
    if (f == "targetLvl") {
      if (!fc.count("outputLvl")) TERMM(-1, "The required configuration value is missing: outputLvl");
      lvl = fc["outputLvl"].get<GInt>();
      if (lvl < m_grid->partitionLvl()) TERMM(-1, "Outputting a lvl below the partition lvl is not possible!");
    } else if (f == "highestLvl") {
      lvl = m_grid->levels().back();
    } else if (f == "lowestLvl" || f == "partitionLvl") {
      lvl = m_grid->partitionLvl();
    } else if (f != "leafCells") {
      TERMM(-1, "Unknown output filter " + f);
    }
    for (GInt q = 0; q < N; ++q)
      m_outputCell[q] = lvl >= 0 ? m_grid->level(q) == lvl : m_grid->property(q, CellProperties::leaf);
  }
  m_ppConf = c.getObject("postprocessing");
  c.opt_config_value<bool>("assumeAxisAligned", true);
  m_refLength = c.opt_config_value<GDouble>("refLength", 1.0);
  // LBMSolver::forcing (src/lbm/solver.cpp:626-696): any non-empty value
  // turns the inlet/outlet correction on, with the pressures 1 +
  // poiseuillePressureGradient and 1 on the surfaces cube_-x and cube_+x
  m_forcing = !c.opt_config_value<std::string>("forcing", "").empty();
  if (m_forcing) m_forcingGrad = c.required_config_value<GDouble>("poiseuillePressureGradient");
  const GDouble finestSpacing = 1.0 / (std::pow((GDouble)m_grid->size(), 1.0 / m_dim) - 1);
  if (c.has_config_value("reynoldsnumber") && c.has_config_value("relaxation"))
    TERMM(-1, "Only set either reynoldsnumber or relaxation");
  m_ma = c.required_config_value<GDouble>("ma");
  if (c.has_config_value("relaxation")) {
    m_relaxTime = c.required_config_value<GDouble>("relaxation");
    m_omega = 1.0 / m_relaxTime;
    m_nu = (2 * m_relaxTime - 1) / 6.0;
    m_re = m_ma * m_refLength / m_nu;
  } else {
    // src/lbm/solver.cpp:114-122 including its non-standard omega
    m_re = c.required_config_value<GDouble>("reynoldsnumber");
    m_nu = m_ma / m_re * m_refLength;
    m_omega = 2.0 / (1.0 + 2.0 * m_nu * std::pow(2.0, (GDouble)m_grid->maxLvl()));
    m_relaxTime = 1.0 / m_omega;
  }
  m_dt = finestSpacing * m_ma * lbm_cs / m_refLength;
  m_convergence = c.opt_config_value<GDouble>("convergence", 1e-12);
  // new keys (no reference counterpart)
  m_storage = storageFromString(m_opt.storage.empty() ? c.opt_config_value<std::string>("storage", "f32") : m_opt.storage);
  if (m_opt.storage.empty() == false) c.opt_config_value<std::string>("storage", "f32");
  auto bf = c.opt_config_value<std::vector<GDouble>>("bodyForce", {0.0, 0.0, 0.0});
  for (size_t d = 0; d < 3 && d < bf.size(); ++d) m_force[d] = bf[d];
  // analytic-comparison keys are read in compareToAnalyticalResult
  m_bnd.setupBndryCnds(c.getObject("boundary"), *m_grid);

  std::cerr << "<<<<<<<<<<<<>>>>>>>>>>>>>\n"
            << "LBM Type BGK\nLBM Model " << model << "\nEquation Navier-Stokes\n"
            << "No. of variables " << m_dim + 1 << "\nNo. Leaf cells: " << m_grid->size()
            << "\nMax Mesh Level: " << m_grid->maxLvl() << "\nNo. Bnd cells: " << m_grid->noBndCells()
            << "\nRelaxation Time: " << m_relaxTime << "\nTimestep: " << m_dt << "\nOmega: " << m_omega
            << "\nReynolds Number: " << m_re << "\nViscosity: " << m_nu << "\nReference V: " << m_ma * lbm_cs
            << "\nStorage: " << (m_storage == LBM_STORE_F32 ? "f32" : m_storage == LBM_STORE_F64 ? "f64" : m_storage == LBM_STORE_F16 ? "f16" : "bf16")
            << "\n+++++++++++++++++++++++++" << std::endl;
}

// initialCondition, src/lbm/solver.cpp:267-295: u from the boundary initCnd,
// rho = 1 everywhere, f = feq(rho, u).
void LBMSolver::initialCondition() {
  const GInt N = m_grid->size();
  const GInt nv = m_dim + 1;
  m_vars.assign(N * nv, 0.0);
  m_bnd.initCndBnd([&](GInt c, GInt v) -> GDouble& { return m_vars[c * nv + v]; });
  for (GInt q = 0; q < N; ++q) m_vars[q * nv + m_dim] = 1.0;
  for (auto& L : m_lat) {
    const CartesianGrid& G = latGrid(L);
    const GInt S = G.structSize();
    std::vector<GDouble> rho(S, 1.0), u(S * m_dim, 0.0);
    for (GInt q = 0; q < G.size(); ++q) {
      const GInt s = G.structIndex(q), g = latId(L, q);
      rho[s] = m_vars[g * nv + m_dim];
      for (GInt d = 0; d < m_dim; ++d) u[s * m_dim + d] = m_vars[g * nv + d];
    }
    check(lbm_init_fields(L.h, rho.data(), u.data()), "lbm_init_fields", L.h);
  }
}

// one lattice per level (LBMSolver::transferGrid keeps every level of the
// grid, src/lbm/solver.cpp:759-801, and every step loops over all cells)
void LBMSolver::createLattices() {
  const auto& lv = m_grid->levels();
  if (m_grid->maxLvl() > m_grid->uniformLvl())
    TERMM(-1, "Boundary refinement (maxRfnmtLvl > uniformLevel) keeps only the cells along the boundary on its levels; "
              "the reference leaves the populations that would stream in from their missing inner neighbours stale, "
              "which the device lattices of this build do not reproduce (uniformLevel > partitionLevel is supported)");
  m_lat.clear();
  m_lat.resize(lv.size());
  for (size_t i = 0; i < lv.size(); ++i) {
    Lattice& L = m_lat[i];
    if (lv.size() > 1) {
      L.grid = m_grid->levelGrid(lv[i], L.ids);
      L.bnd.setupBndryCnds(m_solverCfg->getObject("boundary"), *L.grid);
    }
    const CartesianGrid& G = latGrid(L);
    const DevicePlan plan = (L.grid ? L.bnd : m_bnd).lower(G);
    lbm_grid_desc g;
    g.dim = (int32_t)m_dim;
    g.nx = (int32_t)plan.nx;
    g.ny = (int32_t)plan.ny;
    g.nz = (int32_t)plan.nz;
    g.solid = plan.anySolid ? plan.solid.data() : nullptr;
    lbm_params p;
    p.lattice = (int32_t)m_lattice;
    p.storage = m_storage;
    p.omega = m_omega;
    for (int d = 0; d < 3; ++d) p.force[d] = m_force[d];
    check(lbm_create(&g, &p, &L.h), "lbm_create");
    check(lbm_set_bc(L.h, plan.faces.data(), (int32_t)plan.faces.size()), "lbm_set_bc", L.h);
    if (m_forcing) {
      // the reference takes the surfaces cube_-x and cube_+x and pairs their
      // cells row by row; on the structured window they must be the x = 0
      // and x = nx-1 columns
      if (lv.size() > 1) TERMM(-1, "The 'forcing' correction is not supported on multi-level grids");
      if (m_dim != 2) TERMM(-1, "The 'forcing' correction is 2D (D2Q9)");
      if (m_force[0] != 0 || m_force[1] != 0 || m_force[2] != 0) TERMM(-1, "Use either 'forcing' or 'bodyForce'");
      if (!G.hasBndrySurface("cube_-x") || !G.hasBndrySurface("cube_+x"))
        TERMM(-1, "The 'forcing' correction needs the surfaces cube_-x and cube_+x");
      auto column = [&](const char* name, GInt x) {
        const auto& cells = const_cast<CartesianGrid&>(G).bndrySurface(name).getCellList();
        std::set<GInt> rows;
        for (GInt q : cells) {
          const GInt si = G.structIndex(q);
          if (si % plan.nx != x) return false;
          rows.insert(si / plan.nx);
        }
        return (GInt)rows.size() == plan.ny;
      };
      if (!column("cube_-x", 0) || !column("cube_+x", plan.nx - 1))
        TERMM(-1, "The 'forcing' correction needs cube_-x and cube_+x to be the full x faces of the grid");
      check(lbm_set_pressure_forcing(L.h, 1, 1.0 + m_forcingGrad, 1.0), "lbm_set_pressure_forcing", L.h);
    }
    if (lv.size() > 1)
      std::cerr << "Level " << lv[i] << ": " << G.size() << " cells, window " << plan.nx << " x " << plan.ny << " x "
                << plan.nz << std::endl;
  }
}

// one step of every lattice; check steps add their sums (each a sum over
// the level's cells) into the sums over the whole grid
auto LBMSolver::stepAll(int mode, lbm_sums* sums) -> int {
  int out = LBM_OK;
  if (sums) *sums = lbm_sums{};
  lbm_sums s{};
  for (auto& L : m_lat) {
    s = lbm_sums{};
    const int rc = lbm_step_ex(L.h, mode, sums ? &s : nullptr);
    if (rc == LBM_DIVERGED) out = LBM_DIVERGED;
    else check(rc, "lbm_step_ex", L.h);
    if (!sums) continue;
    sums->mass += s.mass;
    for (int d = 0; d < 3; ++d) sums->momentum[d] += s.momentum[d];
    for (int v = 0; v < 4; ++v) sums->abs_diff[v] += s.abs_diff[v];
    sums->diff_sq += s.diff_sq;
    sums->norm_sq += s.norm_sq;
    sums->nonfinite += s.nonfinite;
    sums->fluid_cells += s.fluid_cells;
  }
  // one lattice: the residual exactly as its step reported it
  if (sums) sums->res_rel = m_lat.size() == 1 ? s.res_rel : (sums->norm_sq > 0 ? std::sqrt(sums->diff_sq / sums->norm_sq) : 0.0);
  return out;
}

auto LBMSolver::convergenceCondition(const lbm_sums& s) -> bool {
  static const char* names2[] = {"U", "V", "rho"};
  static const char* names3[] = {"U", "V", "W", "rho"};
  std::cerr << m_timeStep << ": ";
  GDouble maxConv = 0;
  bool bad = false;
  for (GInt v = 0; v <= m_dim; ++v) {
    const GDouble c = s.abs_diff[v];
    std::cerr << "d" << (m_dim == 2 ? names2[v] : names3[v]) << "=" << c << " ";
    if (std::isnan(c) || std::isinf(c)) bad = true;
    maxConv = std::max(maxConv, c);
  }
  std::cerr << "res_rel=" << s.res_rel << std::endl;
  if (m_timeStep > 1 && (bad || s.nonfinite > 0)) {
    std::cerr << "Solution diverged!" << std::endl;
    m_diverged = true;
    return true;
  }
  if (m_timeStep > 1 && maxConv < m_convergence) {
    std::cerr << "Reached convergence to: " << maxConv << std::endl;
    return true;
  }
  return false;
}

void LBMSolver::fetchMacros() {
  const GInt N = m_grid->size();
  m_rho.assign(N, 0.0);
  m_u.assign(N * m_dim, 0.0);
  for (auto& L : m_lat) {
    const CartesianGrid& G = latGrid(L);
    const GInt S = G.structSize();
    std::vector<GDouble> rho(S), u(S * m_dim);
    check(lbm_get_macros(L.h, rho.data(), u.data()), "lbm_get_macros", L.h);
    for (GInt q = 0; q < G.size(); ++q) {
      const GInt s = G.structIndex(q), g = latId(L, q);
      m_rho[g] = rho[s];
      for (GInt d = 0; d < m_dim; ++d) m_u[g * m_dim + d] = u[s * m_dim + d];
    }
  }
}

auto LBMSolver::run() -> GInt {
  if (m_benchmark) return runBenchmark();
  if (!m_grid) TERMM(-1, "No grid transferred to the LBM solver");
  std::cerr << m_grid->dim() << "D LBM Solver started ||>" << std::endl;
  loadConfiguration();
  m_pp.init(m_ppConf, *m_grid);  // Postprocess::init (src/lbm/solver.cpp:180)
  createLattices();
  initialCondition();
  // postprocessing (src/lbm/solver.cpp:187,199,309,318): the functions read
  // the u held by the solver, which the reference refreshes at the start of
  // every step (macros of the state before it) and when writing output
  {
    const GInt N = m_grid->size(), nv = m_dim + 1;
    m_ppU.assign(N, 0.0);
    for (GInt q = 0; q < N; ++q) m_ppU[q] = m_vars[q * nv];
  }
  const bool ppStep = m_pp.active(PPHook::BeforeTimestep) || m_pp.active(PPHook::AfterTimestep);
  m_pp.execute(PPHook::AtStart, *m_grid, m_ppU);

  const auto t0 = std::chrono::steady_clock::now();
  auto lastInfo = t0;
  GInt lastStep = 0;
  bool converged = false;
  lbm_sums last{};
  bool haveSums = false;
  const GInt I = m_convInterval;
  for (m_timeStep = 0; m_timeStep < m_maxSteps && !converged; ++m_timeStep) {
    // writeInfo (src/lbm/solver.cpp:217-230)
    if (m_timeStep > 0 && m_timeStep % m_infoInterval == 0) {
      const auto now = std::chrono::steady_clock::now();
      const double dt = std::chrono::duration<double>(now - lastInfo).count();
      if (!m_opt.quiet) std::cerr << m_timeStep << "/" << m_maxSteps << " " << (m_timeStep - lastStep) / std::max(dt, 1e-9) << "it/s \n";
      lastInfo = now;
      lastStep = m_timeStep;
    }
    if (m_timeStep > 0 && m_timeStep % I == 0 && haveSums) converged = convergenceCondition(last);
    // timeStep: launch k = t+1 gathers the fields of step k; the check at
    // iteration t compares steps t-1 and t-2
    const GInt k = m_timeStep + 1;
    const bool chk = (k + 1) % I == 0;
    const bool snap = (k + 2) % I == 0;
    const int mode = chk ? (snap ? LBM_STEP_CHECK_SNAPSHOT : LBM_STEP_CHECK) : (snap ? LBM_STEP_SNAPSHOT : LBM_STEP_PLAIN);
    m_pp.execute(PPHook::BeforeTimestep, *m_grid, m_ppU);
    if (ppStep) {
      fetchMacros();
      for (GInt q = 0; q < m_grid->size(); ++q) m_ppU[q] = m_u[q * m_dim];
    }
    lbm_sums s{};
    const int rc = stepAll(mode, chk ? &s : nullptr);
    if (rc == LBM_DIVERGED || chk) {
      last = s;
      haveSums = true;
    }
    // also at the last step or once converged (src/lbm/solver.cpp:194-195)
    m_pp.execute(PPHook::AfterTimestep, *m_grid, m_ppU);
    output(m_timeStep == m_maxSteps - 1 || converged, "");
  }
  for (auto& L : m_lat) check(lbm_synchronize(L.h), "lbm_synchronize", L.h);
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::cerr << "Main loop: " << m_timeStep << " steps in " << secs << " s ("
            << (double)m_timeStep * m_grid->size() / std::max(secs, 1e-12) / 1e6 << " MLUPS)" << std::endl;
  m_pp.execute(PPHook::AtEnd, *m_grid, m_ppU);
  fetchMacros();
  if (m_diverged) TERMM(-1, "Solution diverged");
  if (m_solverCfg->has_config_value("analyticalSolution")) compareToAnalyticalResult();
  for (const auto& k : m_solverCfg->unusedConfigValues()) std::cerr << "WARNING: unused configuration value " << k << std::endl;
  std::cerr << "LBM Solver finished <||" << std::endl;
  return 0;
}

// compareToAnalyticalResult, src/lbm/solver.cpp:388-482 (couette2D_1_5 and
// poiseuille2D_1 of src/lbm/analytical_solutions.h:20-59)
void LBMSolver::compareToAnalyticalResult() {
  Configuration& c = *m_solverCfg;
  const std::string name = c.required_config_value<std::string>("analyticalSolution");
  if (m_dim != 2) TERMM(-1, "Invalid analyticalSolution :" + name + " selected!");
  GDouble dp = 0;
  if (name == "poiseuille2D_1") dp = c.required_config_value<GDouble>("poiseuillePressureGradient");
  auto solution = [&](const GDouble* x, GDouble* out) {
    if (name == "couette2D_1_5") {
      const GDouble reynoldsNum = 0.75, relaxTime = 0.9, refL = 1.0;
      const GDouble dynViscosity = (2.0 * relaxTime - 1.0) / 6.0;
      const GDouble refV = reynoldsNum * dynViscosity / refL;
      out[0] = refV / 5.0 * x[1];
      out[1] = 0;
    } else if (name == "poiseuille2D_1") {
      out[0] = dp / (2.0 * m_nu) * x[1] * (1.0 - 0.0 - x[1]);
      out[1] = 0;
    } else {
      TERMM(-1, "Invalid analyticalSolution :" + name + " selected!");
    }
  };
  std::set<GInt> excluded;
  if (c.has_config_value("analyticalSolutionExcludeSurface")) {
    for (const auto& s : c.required_config_value<std::vector<std::string>>("analyticalSolutionExcludeSurface"))
      for (GInt q : m_grid->bndrySurface(s).getCellList()) excluded.insert(q);
  }
  GDouble maxError = 0, sumError = 0, sumErrorSq = 0, sumSolution = 0, sumSolutionSq = 0;
  const GInt N = m_grid->size();
  for (GInt q = 0; q < N; ++q) {
    if (excluded.count(q)) continue;
    GDouble x[2] = {m_grid->center(q, 0), m_grid->center(q, 1)}, a[2];
    solution(x, a);
    const GDouble dx = m_u[q * 2] - a[0], dy = m_u[q * 2 + 1] - a[1];
    const GDouble delta = std::sqrt(dx * dx + dy * dy);
    sumError += delta;
    sumErrorSq += delta * delta;
    sumSolution += std::sqrt(a[0] * a[0] + a[1] * a[1]);
    sumSolutionSq += a[0] * a[0] + a[1] * a[1];
    maxError = std::max(delta, maxError);
  }
  const GDouble L2 = (std::sqrt(sumErrorSq) / std::sqrt(sumSolutionSq)) / std::pow((GDouble)N, 1.0 / m_dim);
  const GDouble gre = sumError / sumSolution;
  std::cerr << "Comparing to analytical result " << name << "\nmax. Error: " << maxError << "\navg. L2: " << L2
            << "\nglobal relative error: " << gre << std::endl;
  const GDouble eMax = c.opt_config_value<GDouble>("errorMax", 1.0);
  const GDouble eL2 = c.opt_config_value<GDouble>("errorL2", 1.0);
  const GDouble eGre = c.opt_config_value<GDouble>("errorGRE", 1.0);
  bool failed = false;
  if (eMax < maxError) { failed = true; std::cerr << "Error bounds for the maximum error have failed!!! ( < " << eMax << ")" << std::endl; }
  if (eL2 < L2 || std::isnan(L2)) { failed = true; std::cerr << "Error bounds for the L2 error have failed!!! ( < " << eL2 << ")" << std::endl; }
  if (eGre < gre || std::isnan(gre)) { failed = true; std::cerr << "Error bounds for the global relative error have failed!!! ( < " << eGre << ")" << std::endl; }
  if (failed) TERMM(-1, "Analytical testcase failed");
}

// LBMSolver::output (src/lbm/solver.cpp:323-382): every solution_interval
// steps and when forced, the fields u, v, (w), rho of every cell that passes
// the output filter go to <output_dir><solution_filename>_<step><postfix>.vtp;
// after a divergence, once, the fields as they are with postfix "bdiv".
void LBMSolver::output(bool forced, const std::string& postfix) {
  if (m_diverged && !m_wroteDiverged) {
    m_wroteDiverged = true;
    output(true, "bdiv");
    return;
  }
  if (!((m_timeStep > 0 && m_timeStep % m_solutionInterval == 0) || forced)) return;
  // the reference's output first runs updateMacroscopicValues on the fold
  // after this iteration's step (src/lbm/solver.cpp:336-337), so the files,
  // the vars the atEnd hook sees and those compareToAnalyticalResult reads
  // after the final forced output are the macros after the step
  fetchMacros();
  const GInt N = m_grid->size();
  for (GInt q = 0; q < N; ++q) m_ppU[q] = m_u[q * m_dim];
  // the cells the filter passes, in cell order (VTK::BINARY::writePoints,
  // src/common/IO.h:414-507)
  std::vector<GInt> cells;
  for (GInt q = 0; q < N; ++q)
    if (m_outputCell[q]) cells.push_back(q);
  const GInt n = (GInt)cells.size();
  if (n == 0) {
    std::cerr << "ERROR: Invalid call to encodeLE() with length = 0" << std::endl;
    std::exit(-1);
  }
  static const char* velName[3] = {"U", "V", "W"};  // VAR::VELSTR, src/lbm/variables.h:70
  std::vector<vtk::Field> fields;
  for (GInt d = 0; d < m_dim; ++d) {
    vtk::Field f{velName[d], std::vector<double>((size_t)n)};
    for (GInt k = 0; k < n; ++k) f.values[k] = vtk::textRoundTrip(m_u[cells[k] * m_dim + d]);
    fields.push_back(std::move(f));
  }
  vtk::Field r{"rho", std::vector<double>((size_t)n)};
  for (GInt k = 0; k < n; ++k) r.values[k] = vtk::textRoundTrip(m_rho[cells[k]]);
  fields.push_back(std::move(r));
  std::vector<double> xyz((size_t)n * m_dim);
  for (GInt k = 0; k < n; ++k)
    for (GInt d = 0; d < m_dim; ++d) xyz[k * m_dim + d] = m_grid->center(cells[k], d);
  mkdirs(m_outputDir);
  const std::string base = m_outputDir + m_solutionFile + "_" + std::to_string(m_timeStep) + postfix;
  if (!vtk::writePoints(base, (int)m_dim, n, xyz.data(), fields))
    std::cerr << "WARNING: cannot write " << base << ".vtp" << std::endl;
}

// `lbm --bench`: the reference's LBMSolver::initBenchmark is TERMM("Not
// implemented!") (src/lbm/solver.cpp:39-46); here it times the D3Q19 BGK
// periodic box 256^3 (Taylor-Green vortex u0 = 0.05) in FP32.
auto LBMSolver::runBenchmark() -> GInt {
  const int n = 256;
  lbm_grid_desc g{3, n, n, n, nullptr};
  lbm_params p{LBM_D3Q19, m_opt.storage.empty() ? LBM_STORE_F32 : storageFromString(m_opt.storage), 1.0 / 0.6, {0, 0, 0}};
  check(lbm_create(&g, &p, &m_h), "lbm_create");
  const GInt N = (GInt)n * n * n;
  std::vector<GDouble> rho(N, 1.0), u(N * 3, 0.0);
  const GDouble k = 2 * M_PI / n;
  for (GInt z = 0; z < n; ++z)
    for (GInt y = 0; y < n; ++y)
      for (GInt x = 0; x < n; ++x) {
        const GInt c = x + n * (y + n * z);
        const GDouble X = (x + 0.5) * k, Y = (y + 0.5) * k, Z = (z + 0.5) * k;
        u[c * 3 + 0] = 0.05 * std::sin(X) * std::cos(Y) * std::cos(Z);
        u[c * 3 + 1] = -0.05 * std::cos(X) * std::sin(Y) * std::cos(Z);
      }
  check(lbm_init_fields(m_h, rho.data(), u.data()), "lbm_init_fields");
  check(lbm_step(m_h, 20), "lbm_step");
  double ms = 0;
  const int steps = 200;
  check(lbm_time_steps(m_h, steps, &ms), "lbm_time_steps");
  int64_t ncell;
  int32_t q, bpp;
  lbm_info(m_h, &ncell, &q, &bpp);
  const double s = ms * 1e-3 / steps;
  std::cout << "D3Q19 BGK 256^3 periodic: " << ms / steps << " ms/step, " << N / s / 1e6 << " MLUPS, "
            << 2.0 * q * bpp * N / s / 1e9 << " GB/s algorithmic (" << 2.0 * q * bpp * N / s / 8e12 * 100 << " % of 8 TB/s)"
            << std::endl;
  return 0;
}

}  // namespace lbmh

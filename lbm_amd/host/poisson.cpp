// The reference's Poisson solver (LBMSolver<..., LBEquationType::Poisson>,
// src/lbm/solver.cpp) on the device step library of include/lbm_poisson.h.
#include "poisson.h"

#include <cmath>
#include <set>

#include "expr.h"
#include "vtk.h"

namespace lbmh {

namespace {
constexpr GDouble kDebyeHuckel = 27.79;  // solver.cpp:597 (Debye-Hueckel kappa)

// D1Q3 / D2Q5 / D2Q9 poissonAlpha (src/lbm/constants.h:261,285,309)
auto poissonAlpha(GInt lattice) -> GDouble { return lattice == 5 ? 0.5 : 1.0 / 3.0; }

// analytical::poisson (src/lbm/analytical_solutions.h:68-101); x: the
// (shifted) centre; returns the potential (the second component of the 2D
// solution is 0)
auto analyticPotential(const std::string& name, const GDouble* x, GInt dim) -> GDouble {
  if (name == "poissonCHAI08_1" && dim == 1) {
    const GDouble k = kDebyeHuckel;
    const GDouble epK = std::exp(k), emK = std::exp(-k);
    return (epK - 1.0) / (epK - emK) * std::exp(-k * x[0]) + (1.0 - emK) / (epK - emK) * std::exp(k * x[0]);
  }
  if (name == "poissonSimpleDiffReaction" && dim == 1) {
    const GDouble th = 1.0;
    return std::cosh(th * (1.0 - x[0])) / std::cosh(th);
  }
  if (name == "poissonCHAI08_2" && dim == 2) {
    const GDouble mu = std::sqrt(4 + M_PI * M_PI);
    return std::cos(M_PI * x[0]) * std::sinh(mu * (1 - x[1])) / std::sinh(mu);
  }
  TERMM(-1, "Invalid analyticalSolution :" + name + " selected!");
}
}  // namespace

auto solverEquation(const json& cfg) -> std::string {
  if (cfg.count("solver") && cfg["solver"].count("equation")) return cfg["solver"]["equation"].get<std::string>();
  return "navierstokes";
}

PoissonSolver::~PoissonSolver() {
  if (m_h) lbm_poisson_destroy(m_h);
}

void PoissonSolver::check(int rc, const char* what) {
  if (rc != LBM_OK) TERMM(-1, std::string(what) + ": " + lbm_poisson_last_error(m_h));
}

void PoissonSolver::init(int, char**, std::string config_file) {
  m_config = Configuration::fromFile(config_file);
  if (!m_config.count("solver")) TERMM(-1, "No solver section in " + config_file);
  m_solverCfg = std::make_unique<Configuration>(m_config["solver"], "solver.");
}

void PoissonSolver::transferGrid(const GridInterface& grid) {
  const auto* cg = dynamic_cast<const CartesianGrid*>(&grid);
  if (!cg) TERMM(-1, "Invalid grid type");
  takeGrid(std::make_unique<CartesianGrid>(*cg));
}

void PoissonSolver::takeGrid(std::unique_ptr<CartesianGrid> g) {
  if (!m_solverCfg) TERMM(-1, "init() must be called before the grid is transferred");
  m_grid = std::move(g);
  std::cerr << "Transferring " << m_grid->dim() << "D Grid to LBM solver" << std::endl;
  m_grid->loadGridInplace(*m_solverCfg);
}

// loadConfiguration, Poisson branch (src/lbm/solver.cpp:71-163): tau from
// "relaxation", dt = finest spacing (ma = 1/cs, so dt = spacing / refLength),
// kappa from "equation_application" (collisionStep, solver.cpp:589-600)
void PoissonSolver::loadConfiguration() {
  Configuration& c = *m_solverCfg;
  const std::string method = c.opt_config_value<std::string>("method", "bgk");
  if (method != "bgk") TERMM(-1, "Invalid equation configuration!");
  c.opt_config_value<std::string>("type", "lbm");
  c.opt_config_value<std::string>("equation", "poisson");
  m_dim = m_grid->dim();
  const std::string model = c.opt_config_value<std::string>("model", "D2Q9");
  if (model == "D1Q3") m_lattice = 3;
  else if (model == "D2Q5") m_lattice = 5;
  else if (model == "D2Q9") m_lattice = 9;
  else TERMM(-1, "Unsupported model");
  if ((m_lattice == 3) != (m_dim == 1) || m_dim > 2)
    TERMM(-1, "Invalid configuration the grid dimensionality is not matching!");
  m_infoInterval = c.opt_config_value<GInt>("info_interval", 10);
  m_convInterval = c.opt_config_value<GInt>("conv_interval", m_infoInterval);
  if (m_convInterval < 1) TERMM(-1, "conv_interval must be >= 1");
  c.opt_config_value<GInt>("keep_alive_time", 60);
  m_maxSteps = c.required_config_value<GInt>("maxSteps");
  m_outputDir = c.opt_config_value<std::string>("output_dir", "out/");
  if (m_outputDir.empty() || m_outputDir.back() != '/') m_outputDir += '/';
  m_solutionInterval = c.opt_config_value<GInt>("solution_interval", 100);
  if (m_solutionInterval < 1) TERMM(-1, "solution_interval must be >= 1");
  m_solutionFile = c.opt_config_value<std::string>("solution_filename", "solution");
  {
    json fc = c.getObject("cellFilter");
    if (fc.empty()) fc = json{{"cellFilter", "leafCells"}};
    if (!fc.count("cellFilter")) TERMM(-1, "Invalid output configuration missing \"cellFilter\" key");
    const std::string f = fc["cellFilter"].get<std::string>();
    m_outputAll = true;
    if (f == "targetLvl") {
      if (!fc.count("outputLvl")) TERMM(-1, "The required configuration value is missing: outputLvl");
      m_outputAll = fc["outputLvl"].get<GInt>() == m_grid->partitionLvl();
    } else if (f != "highestLvl" && f != "lowestLvl" && f != "partitionLvl" && f != "leafCells") {
      TERMM(-1, "Unknown output filter " + f);
    }
  }
  m_ppConf = c.getObject("postprocessing");
  c.opt_config_value<bool>("assumeAxisAligned", true);
  const GDouble refLength = c.opt_config_value<GDouble>("refLength", 1.0);
  const GDouble finestSpacing = 1.0 / (std::pow((GDouble)m_grid->size(), 1.0 / m_dim) - 1);
  m_relaxTime = c.required_config_value<GDouble>("relaxation");
  m_omega = 1.0 / m_relaxTime;
  m_dt = finestSpacing / refLength;
  m_convergence = c.opt_config_value<GDouble>("convergence", 1e-12);
  const std::string app = c.opt_config_value<std::string>("equation_application", "debye_huckel");
  if (app == "simple_diff_reaction") m_kappa = c.required_config_value<GDouble>("equation_th");
  else if (app == "debye_huckel") m_kappa = kDebyeHuckel;
  else TERMM(-1, "Invalid model configuration!");
  const GDouble diffusivity = poissonAlpha(m_lattice) * 1.0 * (0.5 - m_relaxTime) * m_dt;
  m_source = m_dt * diffusivity * m_kappa * m_kappa;
  setupBndryCnds(c.getObject("boundary"));
  std::cerr << "<<<<<<<<<<<<>>>>>>>>>>>>>\nLBM Type BGK\nLBM Model " << model << "\nEquation Poisson\n"
            << "No. of variables 1\nNo. Leaf cells: " << m_grid->size() << "\nMax Mesh Level: " << m_grid->maxLvl()
            << "\nNo. Bnd cells: " << m_grid->noBndCells() << "\nRelaxation Time: " << m_relaxTime
            << "\nTimestep: " << m_dt << "\nOmega: " << m_omega << "\n+++++++++++++++++++++++++" << std::endl;
}

// LBMBndManager::setupBndryCnds (src/lbm/bnd/bnd.h:69-133) for the Poisson
// objects: 'dirichlet'/'neem' (values: a number, [number] or an expression
// of x, y, z evaluated at each surface cell's centre,
// bnd_dirichlet.h:264-286) and 'neumann'/'neem'
void PoissonSolver::setupBndryCnds(const json& cfg) {
  int order = 0;
  for (auto g = cfg.begin(); g != cfg.end(); ++g) {
    const GInt noBnds = (GInt)g.value().size();
    for (auto f = g.value().begin(); f != g.value().end(); ++f) {
      const std::string name = noBnds > 1 ? g.key() + "_" + f.key() : g.key();
      const json& props = f.value();
      if (!m_grid->hasBndrySurface(name)) TERMM(-1, "Invalid bndryId \"" + name + "\"");
      const Surface& srf = m_grid->bndrySurface(name);
      if (srf.size() == 0) {
        std::cerr << "WARNING: Skipping " << name << " no valid cells!" << std::endl;
        continue;
      }
      if (!props.count("type")) TERMM(-1, "The required configuration value is missing: type");
      const std::string type = props["type"].get<std::string>();
      if (type != "dirichlet" && type != "neumann") TERMM(-1, "Invalid bndCndType: " + type);
      if (!props.count("model")) TERMM(-1, "The required configuration value is missing: model");
      const std::string model = props["model"].get<std::string>();
      // the reference's bounce-back Dirichlet object terminates at its
      // first apply for the Poisson equation (bnd_dirichlet.h:95-103)
      if (type == "dirichlet" && model == "bounceback")
        TERMM(-1, "Dirichlet bounce-back for the Poisson equation: this is incorrect!");
      if (model != "neem")
        TERMM(-1, "The Poisson solver of this build has the 'neem' model for '" + type + "' faces (got '" + model + "')");
      Bnd b;
      b.surface = name;
      b.kind = type == "dirichlet" ? LBM_POISSON_DIRICHLET : LBM_POISSON_NEUMANN;
      b.order = order++;
      if (!props.count("value")) TERMM(-1, "The required configuration value is missing: value");
      const json& v = props["value"];
      const auto& cells = srf.getCellList();
      if (v.is_string() || (v.is_array() && !v.empty() && v[0].is_string())) {
        if (v.is_array()) TERMM(-1, "Impl");
        std::unique_ptr<expr::Expression> e;
        try {
          e = std::make_unique<expr::Expression>(v.get<std::string>());
        } catch (const std::exception& ex) {
          TERMM(-1, ex.what());
        }
        for (GInt cell : cells)
          b.value.push_back(e->eval(m_grid->center(cell, 0), m_dim > 1 ? m_grid->center(cell, 1) : 0.0, 0.0));
      } else {
        const GDouble val = v.is_array() ? v[0].get<GDouble>() : v.get<GDouble>();
        b.value.assign(cells.size(), val);
      }
      if (props.count("setAnalyticalValue") && props["setAnalyticalValue"].get<bool>() && m_dim == 1)
        for (size_t k = 0; k < cells.size(); ++k) {
          const GDouble x = m_grid->center(cells[k], 0);
          b.value[k] = analyticPotential("poissonCHAI08_1", &x, 1);
        }
      m_bnd.push_back(std::move(b));
    }
  }
}

void PoissonSolver::fetchPotential() {
  const GInt S = m_grid->structSize();
  std::vector<GDouble> phi(S);
  check(lbm_poisson_get(m_h, phi.data()), "lbm_poisson_get");
  const GInt N = m_grid->size();
  m_phi.assign(N, 0.0);
  for (GInt q = 0; q < N; ++q) m_phi[q] = phi[m_grid->structIndex(q)];
}

auto PoissonSolver::run() -> GInt {
  if (!m_grid) TERMM(-1, "No grid transferred to the LBM solver");
  std::cerr << m_grid->dim() << "D LBM Solver started ||>" << std::endl;
  loadConfiguration();
  m_pp.init(m_ppConf, *m_grid);
  if (m_grid->levels().size() != 1)
    TERMM(-1, "The Poisson solver of this build runs single-level grids (partitionLevel = uniformLevel = maxRfnmtLvl)");
  const GInt nx = m_grid->extent(0), ny = m_grid->extent(1), N = m_grid->size();
  if (nx * ny != N) TERMM(-1, "The Poisson solver needs a full structured window (a box)");
  lbm_poisson_params p{(int32_t)m_lattice, (int32_t)nx, (int32_t)ny, m_omega, m_source};
  check(lbm_poisson_create(&p, &m_h), "lbm_poisson_create");
  // faces of the window: each boundary object covers one face; its values
  // go along the face in structured order
  std::vector<lbm_poisson_face> faces;
  std::vector<std::vector<GDouble>> vals;
  vals.reserve(m_bnd.size());
  for (const Bnd& b : m_bnd) {
    const Surface& srf = m_grid->bndrySurface(b.surface);
    const auto& cells = srf.getCellList();
    // the face: the missing direction all cells of the surface share
    int face = -1;
    for (int d = 0; d < 2 * m_dim && face < 0; ++d) {
      bool all = true;
      for (GInt c : cells) all = all && m_grid->neighbor(c, d) == INVALID_CELLID;
      if (all) face = d;
    }
    if (face < 0) TERMM(-1, "Boundary surface " + b.surface + " is not a face of the window");
    const GInt len = face < 2 ? ny : nx;
    vals.emplace_back(len, 0.0);
    for (size_t k = 0; k < cells.size(); ++k) {
      const GInt s = m_grid->structIndex(cells[k]);
      vals.back()[face < 2 ? s / nx : s % nx] = b.value.empty() ? 0.0 : b.value[k];
    }
    faces.push_back({face, b.kind, b.order, vals.back().data()});
  }
  check(lbm_poisson_set_faces(m_h, faces.data(), (int32_t)faces.size()), "lbm_poisson_set_faces");
  // initialCondition (solver.cpp:267-295): vars = 0, the Dirichlet objects'
  // initCnd writes their values (the Neumann object's is empty)
  std::vector<GDouble> vars(N, 0.0), phi0(nx * ny, 0.0);
  for (const Bnd& b : m_bnd) {
    if (b.kind != LBM_POISSON_DIRICHLET) continue;
    const auto& cells = m_grid->bndrySurface(b.surface).getCellList();
    for (size_t k = 0; k < cells.size(); ++k) vars[cells[k]] = b.value[k];
  }
  for (GInt q = 0; q < N; ++q) phi0[m_grid->structIndex(q)] = vars[q];
  check(lbm_poisson_init(m_h, phi0.data()), "lbm_poisson_init");
  m_phi = vars;
  const bool ppStep = m_pp.active(PPHook::BeforeTimestep) || m_pp.active(PPHook::AfterTimestep);
  m_pp.execute(PPHook::AtStart, *m_grid, m_phi);

  const auto t0 = std::chrono::steady_clock::now();
  auto lastInfo = t0;
  GInt lastStep = 0;
  bool converged = false, haveSum = false;
  GDouble lastSum = 0;
  const GInt I = m_convInterval;
  for (m_timeStep = 0; m_timeStep < m_maxSteps && !converged; ++m_timeStep) {
    if (m_timeStep > 0 && m_timeStep % m_infoInterval == 0) {
      const auto now = std::chrono::steady_clock::now();
      const double dt = std::chrono::duration<double>(now - lastInfo).count();
      if (!m_opt.quiet) std::cerr << m_timeStep << "/" << m_maxSteps << " " << (m_timeStep - lastStep) / std::max(dt, 1e-9) << "it/s \n";
      lastInfo = now;
      lastStep = m_timeStep;
    }
    // convergenceCondition (solver.cpp:233-263): sumAbsDiff of the one
    // variable, formed by the check launch of the previous iteration
    if (m_timeStep > 0 && m_timeStep % I == 0 && haveSum) {
      std::cerr << m_timeStep << ": dP=" << lastSum << " " << std::endl;
      if (m_timeStep > 1 && lastSum < m_convergence) {
        std::cerr << "Reached convergence to: " << lastSum << std::endl;
        converged = true;
      } else if (m_timeStep > 1 && (std::isnan(lastSum) || std::isinf(lastSum))) {
        std::cerr << "Solution diverged!" << std::endl;
        m_diverged = true;
        converged = true;
      }
    }
    // iteration t runs launch k = t + 1, the step that takes the state from
    // t to t + 1 steps; the check at iteration k reads vars and varsold as
    // they stand after it, so launch k forms that sum when k % I == 0
    const GInt k = m_timeStep + 1;
    m_pp.execute(PPHook::BeforeTimestep, *m_grid, m_phi);
    if (ppStep) fetchPotential();
    if (k % I == 0) {
      const int rc = lbm_poisson_step_check(m_h, &lastSum);
      if (rc != LBM_DIVERGED) check(rc, "lbm_poisson_step_check");
      haveSum = true;
    } else {
      // without per-step hooks the launches up to the next check launch,
      // output or the last step run as one call: iterations t .. t+n-1,
      // none of them the check launch's iteration (so no convergence test
      // falls inside) and none of t .. t+n-2 writing output
      GInt n = 1;
      if (!ppStep && !converged) {
        const GInt t = m_timeStep;
        const GInt nextCheck = (k / I + 1) * I - 1;  // iteration of the next check launch
        GInt nextOut = std::max<GInt>(t, 1);
        nextOut = (nextOut + m_solutionInterval - 1) / m_solutionInterval * m_solutionInterval;
        n = std::min({nextCheck - t, nextOut - t + 1, m_maxSteps - t});
        n = std::max<GInt>(n, 1);
      }
      check(lbm_poisson_step(m_h, (int32_t)n), "lbm_poisson_step");
      m_timeStep += n - 1;
    }
    m_pp.execute(PPHook::AfterTimestep, *m_grid, m_phi);
    output(m_timeStep == m_maxSteps - 1 || converged, "");
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::cerr << "Main loop: " << m_timeStep << " steps in " << secs << " s" << std::endl;
  fetchPotential();
  m_pp.execute(PPHook::AtEnd, *m_grid, m_phi);
  if (m_diverged) TERMM(-1, "Solution diverged");
  if (m_solverCfg->has_config_value("analyticalSolution")) compareToAnalyticalResult();
  for (const auto& key : m_solverCfg->unusedConfigValues()) std::cerr << "WARNING: unused configuration value " << key << std::endl;
  std::cerr << "LBM Solver finished <||" << std::endl;
  return 0;
}

// compareToAnalyticalResult (solver.cpp:388-482) with the Poisson variable:
// v = (P, 0); shiftCenter (solver.cpp:484-500) maps the centres of
// poissonCHAI08_1 onto equidistant points of (0, 1)
void PoissonSolver::compareToAnalyticalResult() {
  Configuration& c = *m_solverCfg;
  const std::string name = c.required_config_value<std::string>("analyticalSolution");
  std::set<GInt> excluded;
  if (c.has_config_value("analyticalSolutionExcludeSurface"))
    for (const auto& s : c.required_config_value<std::vector<std::string>>("analyticalSolutionExcludeSurface"))
      for (GInt q : m_grid->bndrySurface(s).getCellList()) excluded.insert(q);
  const GInt N = m_grid->size();
  const GDouble extent = std::abs(m_grid->center(0, 0) - m_grid->center(N - 1, 0));
  GDouble maxError = 0, sumError = 0, sumErrorSq = 0, sumSolution = 0, sumSolutionSq = 0;
  for (GInt q = 0; q < N; ++q) {
    if (excluded.count(q)) continue;
    GDouble x[2] = {m_grid->center(q, 0), m_dim > 1 ? m_grid->center(q, 1) : 0.0};
    if (name == "poissonCHAI08_1") {
      if (m_dim > 1) TERMM(-1, "FIX ME");
      const GDouble pos = std::floor(x[0] / (extent / (N - 1)));
      x[0] = pos * 1.0 / (N - 1);
    }
    const GDouble a = analyticPotential(name, x, m_dim);
    const GDouble delta = std::abs(m_phi[q] - a);
    sumError += delta;
    sumErrorSq += delta * delta;
    sumSolution += std::abs(a);
    sumSolutionSq += a * a;
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

// output (solver.cpp:323-382), Poisson: the potential as the field "V",
// recomputed from the state after this iteration's step
void PoissonSolver::output(bool forced, const std::string& postfix) {
  if (m_diverged && !m_wroteDiverged) {
    m_wroteDiverged = true;
    output(true, "bdiv");
    return;
  }
  if (!((m_timeStep > 0 && m_timeStep % m_solutionInterval == 0) || forced)) return;
  fetchPotential();
  const GInt N = m_grid->size();
  if (!m_outputAll) {
    std::cerr << "ERROR: Invalid call to encodeLE() with length = 0" << std::endl;
    std::exit(-1);
  }
  std::vector<vtk::Field> fields;
  vtk::Field f{"V", std::vector<double>((size_t)N)};
  for (GInt q = 0; q < N; ++q) f.values[q] = vtk::textRoundTrip(m_phi[q]);
  fields.push_back(std::move(f));
  std::vector<double> xyz((size_t)N * m_dim);
  for (GInt q = 0; q < N; ++q)
    for (GInt d = 0; d < m_dim; ++d) xyz[q * m_dim + d] = m_grid->center(q, d);
  mkdirs(m_outputDir);
  const std::string base = m_outputDir + m_solutionFile + "_" + std::to_string(m_timeStep) + postfix;
  if (!vtk::writePoints(base, (int)m_dim, N, xyz.data(), fields))
    std::cerr << "WARNING: cannot write " << base << ".vtp" << std::endl;
}

}  // namespace lbmh

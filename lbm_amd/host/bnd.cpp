#include "bnd.h"

#include <map>

namespace lbmh {

void LBMBndManager::setupBndryCnds(const json& bndConfig, CartesianGrid& grid) {
  const GInt dim = grid.dim();
  for (auto g = bndConfig.begin(); g != bndConfig.end(); ++g) {
    const std::string geometry = g.key();
    const GInt noBnds = (GInt)g.value().size();
    for (auto f = g.value().begin(); f != g.value().end(); ++f) {
      const std::string name = noBnds > 1 ? geometry + "_" + f.key() : geometry;
      const json& props = f.value();
      if (!grid.hasBndrySurface(name)) TERMM(-1, "Invalid bndryId \"" + name + "\"");
      Surface& srf = grid.bndrySurface(name);
      if (srf.size() == 0) {
        std::cerr << "WARNING: Skipping " << name << " no valid cells!" << std::endl;
        continue;
      }
      const std::string type = required_value<std::string>(props, "type");
      // handleDummyBnd (bnd.h:147-159)
      if (!opt_value<bool>(props, "generateBndry", true)) {
        m_bndrys.emplace_back(std::make_unique<LBMBnd_dummy>(&srf));
        continue;
      }
      if (type == "periodic") {
        required_value<std::string>(props, "connection");
        m_bndrys.emplace_back(std::make_unique<LBMBnd_Periodic>(&srf, props));
      } else if (type == "wall") {
        const auto model = required_value<std::string>(props, "model");
        if (model == "bounceback") {
          m_bndrys.emplace_back(std::make_unique<LBMBnd_wallBB>(&srf, props, dim));
        } else if (model == "equilibrium") {
          m_bndrys.emplace_back(std::make_unique<LBMBnd_wallWetnode>(&srf, props, dim, LBM_FACE_WALL_EQ));
        } else if (model == "neem") {
          m_bndrys.emplace_back(std::make_unique<LBMBnd_wallWetnode>(&srf, props, dim, LBM_FACE_WALL_NEEM));
        } else if (model == "nebb") {
          m_bndrys.emplace_back(std::make_unique<LBMBnd_wallWetnode>(&srf, props, dim, LBM_FACE_WALL_NEBB));
        } else {
          TERMM(-1, "Invalid wall boundary model: " + model);
        }
      } else if (type == "pressure") {
        const auto model = opt_value<std::string>(props, "model", "antibounceback");
        if (model == "zouhe") {
          m_bndrys.emplace_back(std::make_unique<LBMBnd_ZouHe>(&srf, props, dim, false));
        } else if (model == "neem") {
          m_bndrys.emplace_back(std::make_unique<LBMBnd_NEEM>(&srf, props, dim, false));
        } else {
          auto p = std::make_unique<LBMBnd_Pressure>(&srf, props);
          p->setRhoVar(dim);
          m_bndrys.emplace_back(std::move(p));
        }
      } else if (type == "inlet" || type == "outlet") {
        TERMM(-1, "Broken");  // as in the reference (bnd.h:174-183)
      } else if (type == "dirichlet") {
        const auto model = required_value<std::string>(props, "model");
        if (model == "bounceback") {
          m_bndrys.emplace_back(std::make_unique<LBMBnd_DirichletBB>(&srf, props, dim));
        } else if (model == "zouhe") {
          m_bndrys.emplace_back(std::make_unique<LBMBnd_ZouHe>(&srf, props, dim, true));
        } else if (model == "neem") {
          m_bndrys.emplace_back(std::make_unique<LBMBnd_NEEM>(&srf, props, dim, true));
        } else {
          TERMM(-1, "Dirichlet boundary model '" + model + "' is not part of this build");
        }
      } else if (type == "neumann") {
        TERMM(-1, "Neumann boundaries (Poisson) are not part of this build");
      } else {
        TERMM(-1, "Invalid bndCndType: " + type);
      }
    }
  }
}

namespace {
auto sameRule(const BndRule& a, const BndRule& b) -> bool {
  if (a.kind != b.kind || a.order != b.order) return false;
  for (int d = 0; d < 3; ++d)
    if (std::abs(a.value[d] - b.value[d]) > 1e-15) return false;
  return std::abs(a.rho - b.rho) <= 1e-15;
}
}  // namespace

auto LBMBndManager::lower(const CartesianGrid& grid) const -> DevicePlan {
  DevicePlan plan;
  const GInt dim = grid.dim();
  plan.nx = grid.extent(0);
  plan.ny = grid.extent(1);
  plan.nz = grid.extent(2);
  plan.solid.assign(grid.structSize(), 0);
  for (GInt s = 0; s < grid.structSize(); ++s) {
    if (grid.cellAt(s) == INVALID_CELLID) {
      plan.solid[s] = LBM_CELL_SOLID;
      plan.anySolid = true;
    }
  }
  const std::array<GInt, 3> n{{plan.nx, plan.ny, plan.nz}};
  std::map<int, std::vector<BndRule>> rules;
  int order = 0;
  for (const auto& b : m_bndrys) {
    ++order;
    const Surface* s = b->surface();
    for (GInt c : s->getCellList()) {
      const GInt dir = s->dirOf(c);
      const GInt a = dir / 2;
      const GInt si = grid.structIndex(c);
      const GInt co[3] = {si % n[0], (si / n[0]) % n[1], si / (n[0] * n[1])};
      const bool onFace = (dir % 2 == 0) ? co[a] == 0 : co[a] == n[a] - 1;
      BndRule r = b->lower(c);
      r.order = order;
      if (onFace) {
        rules[(int)dir].push_back(r);
      } else {
        // missing neighbour inside the window: an interior body
        if (r.kind != LBM_FACE_WALL || r.value[0] != 0 || r.value[1] != 0 || r.value[2] != 0)
          TERMM(-1, "Only no-slip bounce-back is supported on interior bodies");
      }
    }
  }
  std::array<int, 6> kind{{-1, -1, -1, -1, -1, -1}};
  for (int f = 0; f < 2 * dim; ++f) {
    auto it = rules.find(f);
    if (it == rules.end() || it->second.empty()) continue;
    const BndRule& r0 = it->second.front();
    for (const auto& r : it->second)
      if (!sameRule(r, r0))
        TERMM(-1, "Face " + std::to_string(f) + " carries different boundary rules; the structured device path needs one rule per face");
    kind[f] = r0.kind;
    lbm_bc_desc d;
    d.face = f;
    d.kind = r0.kind;
    for (int k = 0; k < 3; ++k) d.value[k] = r0.value[k];
    d.rho = r0.rho;
    // later boundary objects overwrite earlier ones in the reference; a
    // periodic pressure face writes before propagation (preApply), so every
    // other object overwrites it
    d.priority = r0.kind == LBM_FACE_PERIODIC_PRESSURE ? 0 : r0.order;
    if (r0.kind != LBM_FACE_PERIODIC) plan.faces.push_back(d);
  }
  for (int a = 0; a < dim; ++a) {
    const bool pLink = grid.periodicAxis(a);
    const bool p0 = kind[2 * a] == LBM_FACE_PERIODIC || (kind[2 * a] < 0 && pLink);
    const bool p1 = kind[2 * a + 1] == LBM_FACE_PERIODIC || (kind[2 * a + 1] < 0 && pLink);
    if (p0 != p1) TERMM(-1, "Invalid periodic setup: both faces of an axis must be periodic");
    if (p0) {
      // dummy periodic surfaces need grid links that are a wrap of the window
      for (int s = 0; s < 2; ++s) {
        auto it = rules.find(2 * a + s);
        bool viaBc = false;
        if (it != rules.end())
          for (auto& r : it->second) viaBc |= r.periodicViaBc;
        if (!viaBc && !pLink) TERMM(-1, "Periodic links do not wrap the grid along axis " + std::to_string(a));
      }
      continue;
    }
    for (int s = 0; s < 2; ++s) {
      if (kind[2 * a + s] >= 0) continue;
      std::cerr << "WARNING: face " << 2 * a + s << " has no boundary condition; treated as a no-slip wall" << std::endl;
      lbm_bc_desc d{};
      d.face = 2 * a + s;
      d.kind = LBM_FACE_WALL;
      d.rho = 1.0;
      d.priority = 0;
      plan.faces.push_back(d);
    }
  }
  return plan;
}

}  // namespace lbmh

// Boundary conditions of the drop-in host: the reference's LBMBndInterface /
// LBMBndManager shapes (src/lbm/bnd/bnd_interface.h:7-23, src/lbm/bnd/bnd.h:17-282)
// and their lowering onto the device's face descriptors.
//
// The reference applies each boundary object on the CPU through per-element
// std::function accessors.  Those calls cannot cross into device code, so
// every object here instead reports the rule it applies to the missing
// populations of its surface cells (lower()), and the manager compiles the
// rules into one lbm_bc_desc per face of the structured window plus the
// solid mask for interior bodies.  initCnd keeps the reference's host role:
// it writes the boundary's initial values into the host variable array.
#pragma once
#include <functional>
#include <memory>

#include "../../include/lbm_hip.h"
#include "grid.h"

namespace lbmh {

using VarAccess = std::function<GDouble&(GInt cellId, GInt varId)>;

struct BndRule {
  int kind = LBM_FACE_WALL;  // LBM_FACE_*
  int order = 0;             // position of the boundary object in application order
  std::array<GDouble, 3> value{{0, 0, 0}};
  GDouble rho = 1.0;
  bool periodicViaBc = false;  // 'periodic' with generateBndry = true
};

class LBMBndInterface {
 public:
  LBMBndInterface() = default;
  virtual ~LBMBndInterface() = default;
  LBMBndInterface(const LBMBndInterface&) = delete;
  auto operator=(const LBMBndInterface&) -> LBMBndInterface& = delete;

  virtual void initCnd(const VarAccess& vars) = 0;
  // rule for the missing population of surface cell c in direction dir
  [[nodiscard]] virtual auto lower(GInt cellId) const -> BndRule = 0;
  [[nodiscard]] virtual auto surface() const -> const Surface* = 0;
  [[nodiscard]] virtual auto isDummy() const -> bool { return false; }
};

class LBMBnd_dummy : public LBMBndInterface {
 public:
  explicit LBMBnd_dummy(const Surface* s) : m_s(s) {}
  void initCnd(const VarAccess&) override {}
  [[nodiscard]] auto lower(GInt) const -> BndRule override {
    BndRule r;
    r.kind = LBM_FACE_PERIODIC;
    return r;
  }
  [[nodiscard]] auto surface() const -> const Surface* override { return m_s; }
  [[nodiscard]] auto isDummy() const -> bool override { return true; }

 private:
  const Surface* m_s;
};

// 'periodic' handled as a boundary object (generateBndry = true): copies the
// outgoing populations to the linked cell, i.e. a wrap (bnd_periodic.h:102-119);
// with 'pressure' the copied set is re-based on that density
// (feq(p, u) + f - feq), a periodic pair with a density jump.
class LBMBnd_Periodic : public LBMBndInterface {
 public:
  LBMBnd_Periodic(const Surface* s, const json& props) : m_s(s) {
    if (props.count("pressure")) {
      m_pressure = required_value<GDouble>(props, "pressure");
      m_hasPressure = true;
    }
  }
  void initCnd(const VarAccess&) override {}
  [[nodiscard]] auto lower(GInt) const -> BndRule override {
    BndRule r;
    if (m_hasPressure) {
      r.kind = LBM_FACE_PERIODIC_PRESSURE;
      r.rho = m_pressure;
    } else {
      r.kind = LBM_FACE_PERIODIC;
      r.periodicViaBc = true;
    }
    return r;
  }
  [[nodiscard]] auto surface() const -> const Surface* override { return m_s; }

 private:
  const Surface* m_s;
  GDouble m_pressure = 1.0;
  bool m_hasPressure = false;
};

// 'dirichlet'/'bounceback' (bnd_dirichlet.h:20-127): halfway bounce-back
// with the Ladd velocity term 2/cs^2 w (c.u), rho0 = 1; initCnd sets u.
class LBMBnd_DirichletBB : public LBMBndInterface {
 public:
  LBMBnd_DirichletBB(const Surface* s, const json& props, GInt dim) : m_s(s), m_dim(dim) {
    auto v = required_value<std::vector<GDouble>>(props, "value");
    if ((GInt)v.size() < dim) TERMM(-1, "dirichlet value needs at least dim entries");
    for (GInt d = 0; d < dim; ++d) m_value[d] = v[d];
  }
  void initCnd(const VarAccess& vars) override {
    for (GInt c : m_s->getCellList())
      for (GInt d = 0; d < m_dim; ++d) vars(c, d) = m_value[d];
  }
  [[nodiscard]] auto lower(GInt) const -> BndRule override {
    BndRule r;
    r.kind = LBM_FACE_WALL;
    r.value = m_value;
    return r;
  }
  [[nodiscard]] auto surface() const -> const Surface* override { return m_s; }

 private:
  const Surface* m_s;
  GInt m_dim;
  std::array<GDouble, 3> m_value{{0, 0, 0}};
};

// 'wall'/'bounceback' (+ tangentialVelocity), bnd_wall.h:8-95.  The wall
// velocity is V times the tangent (n_y, n_x) of the cell's normal, which is
// what the reference's per-population dirichlet values amount to (2D only in
// the reference; the same tangent rule is not defined in 3D).
class LBMBnd_wallBB : public LBMBndInterface {
 public:
  LBMBnd_wallBB(const Surface* s, const json& props, GInt dim) : m_s(s), m_dim(dim) {
    m_tangentialV = opt_value<GDouble>(props, "tangentialVelocity", 0.0);
    if (std::abs(m_tangentialV) > GDoubleEps && dim != 2) TERMM(-1, "Not implemented");
  }
  void initCnd(const VarAccess&) override {}
  [[nodiscard]] auto lower(GInt cellId) const -> BndRule override {
    BndRule r;
    r.kind = LBM_FACE_WALL;
    if (std::abs(m_tangentialV) > GDoubleEps) {
      const GDouble* n = m_s->normal_p(cellId);
      r.value = {{m_tangentialV * n[1], m_tangentialV * n[0], 0.0}};
    }
    return r;
  }
  [[nodiscard]] auto surface() const -> const Surface* override { return m_s; }

 private:
  const Surface* m_s;
  GInt m_dim;
  GDouble m_tangentialV = 0;
};

// Wet-node walls: 'wall' with model 'equilibrium' (LBMBnd_wallEq,
// bnd_wall.h:100-185), 'neem' (LBMBnd_wallNEEM, bnd_wall.h:187-300) or
// 'nebb' (LBMBnd_wallNEBB, bnd_wall.h:303-479, D2Q9 only); optional
// 'velocity' (no-slip without it).  initCnd is empty, as in the reference.
class LBMBnd_wallWetnode : public LBMBndInterface {
 public:
  LBMBnd_wallWetnode(const Surface* s, const json& props, GInt dim, int kind) : m_s(s), m_kind(kind) {
    if (kind == LBM_FACE_WALL_NEBB && dim != 2) TERMM(-1, "Not implemented for this distribution!");
    if (props.count("velocity")) {
      m_hasVelocity = true;
      auto v = required_value<std::vector<GDouble>>(props, "velocity");
      if ((GInt)v.size() < dim) TERMM(-1, "wall velocity needs dim entries");
      for (GInt d = 0; d < dim; ++d) m_value[d] = v[d];
    }
  }
  void initCnd(const VarAccess&) override {}
  [[nodiscard]] auto lower(GInt cellId) const -> BndRule override {
    BndRule r;
    r.kind = m_kind;
    r.value = m_value;
    // without 'velocity' the reference runs apply_0 (bnd_wall.h:353-400)
    if (m_kind == LBM_FACE_WALL_NEBB && !m_hasVelocity) r.kind = LBM_FACE_WALL_NEBB0;
    if (m_kind == LBM_FACE_WALL_NEBB) {
      // the reference reflects the normal population with the normal
      // velocity of the cell; this build takes the wall's, so only
      // tangential wall velocities are accepted
      const GDouble* n = m_s->normal_p(cellId);
      for (GInt d = 0; d < 3; ++d)
        if (n[d] != 0 && m_value[d] != 0) TERMM(-1, "A NEBB wall velocity must be tangential to the wall");
    }
    return r;
  }
  [[nodiscard]] auto surface() const -> const Surface* override { return m_s; }

 private:
  const Surface* m_s;
  int m_kind;
  bool m_hasVelocity = false;
  std::array<GDouble, 3> m_value{{0, 0, 0}};
};

// 'pressure' anti-bounce-back (bnd_pressure.h:11-113); 3D form new.
class LBMBnd_Pressure : public LBMBndInterface {
 public:
  LBMBnd_Pressure(const Surface* s, const json& props) : m_s(s) {
    m_p = required_value<GDouble>(props, "pressure");
  }
  void initCnd(const VarAccess& vars) override {
    for (GInt c : m_s->getCellList()) vars(c, m_rhoVar) = m_p;
  }
  void setRhoVar(GInt v) { m_rhoVar = v; }
  [[nodiscard]] auto lower(GInt) const -> BndRule override {
    BndRule r;
    r.kind = LBM_FACE_PRESSURE_ABB;
    r.rho = m_p;
    return r;
  }
  [[nodiscard]] auto surface() const -> const Surface* override { return m_s; }

 private:
  const Surface* m_s;
  GDouble m_p = 1.0;
  GInt m_rhoVar = 2;
};

// Zou-He velocity ('dirichlet', model 'zouhe') or pressure ('pressure',
// model 'zouhe') face: new in this build (the reference lists Zou-He as
// planned only).
class LBMBnd_ZouHe : public LBMBndInterface {
 public:
  LBMBnd_ZouHe(const Surface* s, const json& props, GInt dim, bool velocity) : m_s(s), m_dim(dim), m_vel(velocity) {
    if (velocity) {
      auto v = required_value<std::vector<GDouble>>(props, "value");
      for (GInt d = 0; d < dim && d < (GInt)v.size(); ++d) m_value[d] = v[d];
    } else {
      m_p = required_value<GDouble>(props, "pressure");
    }
  }
  void initCnd(const VarAccess& vars) override {
    for (GInt c : m_s->getCellList()) {
      if (m_vel)
        for (GInt d = 0; d < m_dim; ++d) vars(c, d) = m_value[d];
      else
        vars(c, m_dim) = m_p;
    }
  }
  [[nodiscard]] auto lower(GInt) const -> BndRule override {
    BndRule r;
    r.kind = m_vel ? LBM_FACE_VELOCITY_ZH : LBM_FACE_PRESSURE_ZH;
    r.value = m_value;
    r.rho = m_p;
    return r;
  }
  [[nodiscard]] auto surface() const -> const Surface* override { return m_s; }

 private:
  const Surface* m_s;
  GInt m_dim;
  bool m_vel;
  std::array<GDouble, 3> m_value{{0, 0, 0}};
  GDouble m_p = 1.0;
};

// Non-equilibrium extrapolation faces: 'pressure' with model 'neem'
// (prescribed density) and 'dirichlet' with model 'neem' (prescribed
// velocity).  New in this build.  Every population of a face cell is rebuilt
// from its inward neighbour.
class LBMBnd_NEEM : public LBMBndInterface {
 public:
  LBMBnd_NEEM(const Surface* s, const json& props, GInt dim, bool velocity) : m_s(s), m_dim(dim), m_vel(velocity) {
    if (velocity) {
      auto v = required_value<std::vector<GDouble>>(props, "value");
      for (GInt d = 0; d < dim && d < (GInt)v.size(); ++d) m_value[d] = v[d];
    } else {
      m_p = required_value<GDouble>(props, "pressure");
    }
  }
  void initCnd(const VarAccess& vars) override {
    for (GInt c : m_s->getCellList()) {
      if (m_vel)
        for (GInt d = 0; d < m_dim; ++d) vars(c, d) = m_value[d];
      else
        vars(c, m_dim) = m_p;
    }
  }
  [[nodiscard]] auto lower(GInt) const -> BndRule override {
    BndRule r;
    r.kind = m_vel ? LBM_FACE_VELOCITY_NEEM : LBM_FACE_PRESSURE_NEEM;
    r.value = m_value;
    r.rho = m_p;
    return r;
  }
  [[nodiscard]] auto surface() const -> const Surface* override { return m_s; }

 private:
  const Surface* m_s;
  GInt m_dim;
  bool m_vel;
  std::array<GDouble, 3> m_value{{0, 0, 0}};
  GDouble m_p = 1.0;
};

// Device plan compiled from the boundary objects.
struct DevicePlan {
  GInt nx = 1, ny = 1, nz = 1;
  std::vector<uint8_t> solid;  // structured order, 1 = not a fluid cell
  std::vector<lbm_bc_desc> faces;
  bool anySolid = false;
};

class LBMBndManager {
 public:
  // LBMBndManager::setupBndryCnds (bnd.h:71-142)
  void setupBndryCnds(const json& bndConfig, CartesianGrid& grid);
  void initCndBnd(const VarAccess& vars) {
    for (auto& b : m_bndrys) b->initCnd(vars);
  }
  // compile into the structured device description
  [[nodiscard]] auto lower(const CartesianGrid& grid) const -> DevicePlan;
  [[nodiscard]] auto noBndrys() const -> GInt { return (GInt)m_bndrys.size(); }

 private:
  std::vector<std::unique_ptr<LBMBndInterface>> m_bndrys;
};

}  // namespace lbmh

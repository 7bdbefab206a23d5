// Analytic geometries of the grid generator: box, cube and sphere with the
// reference's point-inside and cell-cut tests, grouped into bodies with
// optional subtraction (src/geometry.h:640-1075).
//
// The reference implements the cut tests in 2D only (3D box and sphere are
// TERMM, src/geometry.h:690-692,796-798).  The 3D forms here extend the same
// tests to three axes: a cell is cut by a box when it lies within half a cell
// of one face plane and within the face's extent widened by half a cell;
// a cell is cut by a sphere when its centre lies within r +- sqrt(2) h/2 of
// the sphere centre and not all 2^D corners are inside.
#pragma once
#include <map>
#include <memory>
#include <unordered_map>

#include "common.h"

namespace lbmh {

struct BoundingBox {
  int dim = 0;
  std::array<double, 3> min{{0, 0, 0}}, max{{0, 0, 0}};
};

class Geometry {
 public:
  Geometry(const json& j, std::string name, int dim)
      : m_name(std::move(name)),
        m_body(opt_value<std::string>(j, "body", "unique")),
        m_subtract(opt_value<bool>(j, "subtract", false)),
        m_dim(dim) {}
  virtual ~Geometry() = default;
  [[nodiscard]] virtual auto pointIsInside(const double* x) const -> bool = 0;
  [[nodiscard]] virtual auto cutWithCell(const double* c, double len) const -> bool = 0;
  [[nodiscard]] virtual auto boundingBox() const -> BoundingBox = 0;
  [[nodiscard]] auto name() const -> const std::string& { return m_name; }
  [[nodiscard]] auto body() const -> const std::string& { return m_body; }
  auto body() -> std::string& { return m_body; }
  [[nodiscard]] auto subtract() const -> bool { return m_subtract; }

 protected:
  std::string m_name, m_body;
  bool m_subtract;
  int m_dim;
};

class GeomBox : public Geometry {
 public:
  GeomBox(const json& j, const std::string& name, int dim) : Geometry(j, name, dim) {
    auto A = j.at("A").get<std::vector<double>>();
    auto B = j.at("B").get<std::vector<double>>();
    if ((int)A.size() != dim || (int)B.size() != dim) TERMM(-1, "Invalid dimensionality given for box corner points");
    for (int d = 0; d < dim; ++d) { m_A[d] = A[d]; m_B[d] = B[d]; }
  }
  // src/geometry.h:755-762
  [[nodiscard]] auto pointIsInside(const double* x) const -> bool override {
    for (int d = 0; d < m_dim; ++d)
      if (m_A[d] > x[d] || m_B[d] < x[d]) return false;
    return true;
  }
  // src/geometry.h:764-809 (2D); same rule per axis for 3D
  [[nodiscard]] auto cutWithCell(const double* c, double len) const -> bool override {
    const double h = 0.5 * len;
    if (m_dim == 2) {
      if (m_A[0] - h <= c[0] && m_B[0] + h >= c[0]) {
        if (std::abs(c[1] - m_A[1]) <= h) return true;
        if (std::abs(c[1] - m_B[1]) <= h) return true;
      }
      if (m_A[1] - h <= c[1] && m_B[1] + h >= c[1]) {
        if (std::abs(c[0] - m_A[0]) <= h) return true;
        if (std::abs(c[0] - m_B[0]) <= h) return true;
      }
      return false;
    }
    for (int a = 0; a < m_dim; ++a) {
      bool within = true;
      for (int b = 0; b < m_dim; ++b)
        if (b != a && !(m_A[b] - h <= c[b] && m_B[b] + h >= c[b])) within = false;
      if (!within) continue;
      if (std::abs(c[a] - m_A[a]) <= h) return true;
      if (std::abs(c[a] - m_B[a]) <= h) return true;
    }
    return false;
  }
  [[nodiscard]] auto boundingBox() const -> BoundingBox override {
    BoundingBox b;
    b.dim = m_dim;
    for (int d = 0; d < m_dim; ++d) { b.min[d] = m_A[d]; b.max[d] = m_B[d]; }
    return b;
  }

 private:
  std::array<double, 3> m_A{{0, 0, 0}}, m_B{{0, 0, 0}};
};

class GeomCube : public Geometry {
 public:
  GeomCube(const json& j, const std::string& name, int dim) : Geometry(j, name, dim) {
    auto c = j.at("center").get<std::vector<double>>();
    for (int d = 0; d < dim; ++d) m_c[d] = c.at(d);
    m_len = j.at("length").get<double>();
  }
  // src/geometry.h:866-882
  [[nodiscard]] auto pointIsInside(const double* x) const -> bool override {
    for (int d = 0; d < m_dim; ++d)
      if (std::abs(x[d] - m_c[d]) > m_len) return false;
    return true;
  }
  [[nodiscard]] auto cutWithCell(const double* c, double len) const -> bool override {
    for (int d = 0; d < m_dim; ++d)
      if (std::abs(c[d] - m_c[d]) > m_len + len) return false;
    return true;
  }
  [[nodiscard]] auto boundingBox() const -> BoundingBox override {
    BoundingBox b;
    b.dim = m_dim;
    const double r = std::sqrt((double)m_dim) * m_len;
    for (int d = 0; d < m_dim; ++d) { b.min[d] = m_c[d] - r; b.max[d] = m_c[d] + r; }
    return b;
  }

 private:
  std::array<double, 3> m_c{{0, 0, 0}};
  double m_len = 0;
};

class GeomSphere : public Geometry {
 public:
  GeomSphere(const json& j, const std::string& name, int dim) : Geometry(j, name, dim) {
    auto c = j.at("center").get<std::vector<double>>();
    for (int d = 0; d < dim; ++d) m_c[d] = c.at(d);
    m_r = j.at("radius").get<double>();
  }
  [[nodiscard]] auto dist(const double* x) const -> double {
    double s = 0;
    for (int d = 0; d < m_dim; ++d) s += (x[d] - m_c[d]) * (x[d] - m_c[d]);
    return std::sqrt(s);
  }
  // src/geometry.h:660-662
  [[nodiscard]] auto pointIsInside(const double* x) const -> bool override { return dist(x) < m_r + GDoubleEps; }
  // src/geometry.h:664-698 (2D corner count), 3D extension with 8 corners
  [[nodiscard]] auto cutWithCell(const double* c, double len) const -> bool override {
    const double distance = dist(c);
    const double half = 0.5 * len;
    const double surr = std::sqrt(2.0) * half;
    if (distance <= m_r + surr && distance >= m_r - surr) {
      int inside = 0;
      const int ncorner = 1 << m_dim;
      for (int k = 0; k < ncorner; ++k) {
        double v[3];
        for (int d = 0; d < m_dim; ++d) v[d] = (c[d] - half) + len * ((k >> d) & 1);
        if (dist(v) <= m_r) ++inside;
      }
      if (inside < ncorner) return true;
    }
    return false;
  }
  [[nodiscard]] auto boundingBox() const -> BoundingBox override {
    BoundingBox b;
    b.dim = m_dim;
    for (int d = 0; d < m_dim; ++d) { b.min[d] = m_c[d] - m_r; b.max[d] = m_c[d] + m_r; }
    return b;
  }

 private:
  std::array<double, 3> m_c{{0, 0, 0}};
  double m_r = 0;
};

// GeometryManager, src/geometry.h:922-1075: objects in configuration order
// (sorted keys), bodies keyed by name, subtraction inside a body.
class GeometryManager {
 public:
  void setup(const json& geometry, int dim) {
    for (auto it = geometry.begin(); it != geometry.end(); ++it) {
      const std::string name = it.key();
      const json& g = it.value();
      if (!g.count("type")) TERMM(-1, "Malformed json: No \"type\" given!");
      const std::string type = g.at("type").get<std::string>();
      if (type == "sphere") m_obj.emplace_back(std::make_unique<GeomSphere>(g, name, dim));
      else if (type == "box") m_obj.emplace_back(std::make_unique<GeomBox>(g, name, dim));
      else if (type == "cube") m_obj.emplace_back(std::make_unique<GeomCube>(g, name, dim));
      else if (type == "stl") TERMM(-1, "STL geometries are not supported by this build");
      // unknown types are skipped like the reference
    }
    for (auto& g : m_obj) {
      if (g->body() == "unique") g->body() = g->name();
      m_bodies[g->body()].push_back(g.get());
    }
  }
  [[nodiscard]] auto noObjects() const -> GInt { return (GInt)m_obj.size(); }
  [[nodiscard]] auto pointIsInside(const double* x) const -> bool {
    for (const auto& [key, list] : m_bodies) {
      bool hasSub = false;
      for (auto* o : list) hasSub |= o->subtract();
      for (auto* o : list) {
        if (o->pointIsInside(x)) {
          if (!hasSub) return true;
          if (o->subtract()) return false;
          for (auto* s : list)
            if (s->subtract() && s->pointIsInside(x)) return false;
          return true;
        }
      }
    }
    return false;
  }
  [[nodiscard]] auto cutWithCell(const double* c, double len) const -> bool {
    for (auto& o : m_obj)
      if (o->cutWithCell(c, len)) return true;
    return false;
  }
  [[nodiscard]] auto cutWithCell(const std::string& name, const double* c, double len) const -> bool {
    for (auto& o : m_obj)
      if (o->name() == name && o->cutWithCell(c, len)) return true;
    return false;
  }
  [[nodiscard]] auto isSubtracted(const std::string& name) const -> bool {
    for (auto& o : m_obj)
      if (o->name() == name) return o->subtract();
    return false;
  }
  [[nodiscard]] auto boundingBox(int dim) const -> BoundingBox {
    BoundingBox b;
    b.dim = dim;
    for (size_t i = 0; i < m_obj.size(); ++i) {
      const BoundingBox t = m_obj[i]->boundingBox();
      for (int d = 0; d < dim; ++d) {
        if (b.min[d] > t.min[d] || i == 0) b.min[d] = t.min[d];
        if (b.max[d] < t.max[d] || i == 0) b.max[d] = t.max[d];
      }
    }
    return b;
  }

 private:
  std::vector<std::unique_ptr<Geometry>> m_obj;
  std::map<std::string, std::vector<Geometry*>> m_bodies;
};

}  // namespace lbmh

// Common host types, the fatal-error convention and JSON configuration access
// of the drop-in host layer.
//
// TERMM mirrors the reference's term()/TERMM (src/common/term.h:6-40): print
// "Error in <location>: <message>", then end the process with status 255
// (the observable status of the reference's MPI_Abort(-1)).
// Configuration mirrors the accessors of src/common/configuration.h:15-469:
// required/optional values, scoped sections and the unused-key report.
#pragma once
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <limits>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include <json.hpp>  // nlohmann::json 3.x (system copy); std::map keys => sorted iteration like the reference

namespace lbmh {

using GInt = int64_t;
using GDouble = double;
using GString = std::string;
using json = nlohmann::json;

constexpr GInt INVALID_CELLID = -1;
constexpr GDouble GDoubleEps = std::numeric_limits<GDouble>::epsilon();
constexpr GInt MAX_LVL = 100;

[[noreturn]] inline void term(int code, const std::string& where, const std::string& msg) {
  std::cerr << "\nError in " << where << ": " << msg << "\n\nProgram is aborting!!\n" << std::flush;
  std::exit(code == 0 ? 0 : 255);
}
#define TERMM(code, msg) ::lbmh::term((code), std::string(__FILE__) + ":" + std::to_string(__LINE__), (msg))

// direction ids of the reference, include/common/constants.h:104-125
inline auto dirIdString2Id(const std::string& d) -> GInt {
  if (d == "-x") return 0;
  if (d == "+x") return 1;
  if (d == "-y") return 2;
  if (d == "+y") return 3;
  if (d == "-z") return 4;
  if (d == "+z") return 5;
  TERMM(-1, "Invalid direction " + d);
}

// Scoped configuration access with unused-key tracking.
class Configuration {
 public:
  Configuration() = default;
  explicit Configuration(json j, std::string scope = "") : m_json(std::move(j)), m_scope(std::move(scope)) {}

  static auto fromFile(const std::string& file) -> json {
    std::ifstream in(file);
    if (!in) TERMM(-1, "Configuration file not found: " + file);
    json j;
    try {
      in >> j;
    } catch (const std::exception& e) {
      TERMM(-1, std::string("Malformed configuration file: ") + e.what());
    }
    return j;
  }

  [[nodiscard]] auto has_config_value(const std::string& key) const -> bool { return m_json.count(key) > 0; }

  template <typename T>
  auto required_config_value(const std::string& key) -> T {
    if (!has_config_value(key)) TERMM(-1, "The required configuration value is missing: " + m_scope + key);
    m_used.insert(key);
    try {
      return m_json[key].get<T>();
    } catch (const std::exception& e) {
      TERMM(-1, "Invalid configuration value " + m_scope + key + ": " + e.what());
    }
  }

  template <typename T>
  auto opt_config_value(const std::string& key, const T& def) -> T {
    if (!has_config_value(key)) return def;
    m_used.insert(key);
    try {
      return m_json[key].get<T>();
    } catch (const std::exception& e) {
      TERMM(-1, "Invalid configuration value " + m_scope + key + ": " + e.what());
    }
  }

  auto getObject(const std::string& key) -> json {
    m_used.insert(key);
    return m_json.count(key) ? m_json[key] : json::object();
  }

  // true if any object below this level has key == value (has_any_key_value)
  [[nodiscard]] auto has_any_key_value(const std::string& key, const std::string& value) const -> bool {
    return anyKeyValue(m_json, key, value);
  }

  [[nodiscard]] auto raw() const -> const json& { return m_json; }

  [[nodiscard]] auto unusedConfigValues() const -> std::vector<std::string> {
    std::vector<std::string> out;
    for (auto it = m_json.begin(); it != m_json.end(); ++it)
      if (!m_used.count(it.key())) out.push_back(m_scope + it.key());
    return out;
  }

 private:
  static auto anyKeyValue(const json& j, const std::string& key, const std::string& value) -> bool {
    if (j.is_object()) {
      for (auto it = j.begin(); it != j.end(); ++it) {
        if (it.key() == key && it.value().is_string() && it.value().get<std::string>() == value) return true;
        if (anyKeyValue(it.value(), key, value)) return true;
      }
    }
    return false;
  }
  json m_json;
  std::string m_scope;
  std::set<std::string> m_used;
};

// config::opt_config_value / required_config_value on a plain json object
template <typename T>
auto opt_value(const json& j, const std::string& key, const T& def) -> T {
  if (!j.count(key)) return def;
  return j.at(key).get<T>();
}
template <typename T>
auto required_value(const json& j, const std::string& key) -> T {
  if (!j.count(key)) TERMM(-1, "The required configuration value is missing: " + key);
  return j.at(key).get<T>();
}

}  // namespace lbmh

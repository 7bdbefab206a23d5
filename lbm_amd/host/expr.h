// Arithmetic expressions of boundary values, e.g. "cos(pi*x)" or
// "-sinh(sqrt(4+pi^2)*(1-y))/sinh(sqrt(4+pi^2))", evaluated at a cell
// centre.  The reference compiles such strings with exprtk over the
// variables x, y, z and exprtk's constants (src/common/mathexpr.h:8-40,
// read by config::required_math_expression, src/common/configuration.h:49-56).
// This is a recursive-descent evaluator of the subset the configurations
// use: numbers, x y z, pi, + - * / ^ (power, right associative, above unary
// minus as in exprtk: -x^2 = -(x^2)), parentheses and the functions below.
#pragma once
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace expr {

class Expression {
 public:
  explicit Expression(std::string text) : m_text(std::move(text)) {
    m_pos = 0;
    m_root = parseSum();
    skip();
    if (m_pos != m_text.size()) fail("unexpected '" + m_text.substr(m_pos, 1) + "'");
  }

  // value at the point (x, y, z)
  [[nodiscard]] auto eval(double x, double y = 0, double z = 0) const -> double {
    const double v[3] = {x, y, z};
    return evalNode(*m_root, v);
  }

 private:
  enum class Op { Num, Var, Neg, Add, Sub, Mul, Div, Pow, Call };
  struct Node {
    Op op = Op::Num;
    double value = 0;
    int var = 0;
    std::string fn;
    std::unique_ptr<Node> a, b;
  };

  std::string m_text;
  size_t m_pos = 0;
  std::unique_ptr<Node> m_root;

  [[noreturn]] void fail(const std::string& why) const {
    throw std::invalid_argument("expression \"" + m_text + "\": " + why);
  }
  void skip() {
    while (m_pos < m_text.size() && std::isspace((unsigned char)m_text[m_pos])) ++m_pos;
  }
  auto peek() -> char {
    skip();
    return m_pos < m_text.size() ? m_text[m_pos] : '\0';
  }
  static auto make(Op op, std::unique_ptr<Node> a, std::unique_ptr<Node> b = nullptr) -> std::unique_ptr<Node> {
    auto n = std::make_unique<Node>();
    n->op = op;
    n->a = std::move(a);
    n->b = std::move(b);
    return n;
  }

  // sum := product (('+' | '-') product)*
  auto parseSum() -> std::unique_ptr<Node> {
    auto n = parseProduct();
    for (char c = peek(); c == '+' || c == '-'; c = peek()) {
      ++m_pos;
      n = make(c == '+' ? Op::Add : Op::Sub, std::move(n), parseProduct());
    }
    return n;
  }
  // product := unary (('*' | '/') unary)*
  auto parseProduct() -> std::unique_ptr<Node> {
    auto n = parseUnary();
    for (char c = peek(); c == '*' || c == '/'; c = peek()) {
      ++m_pos;
      n = make(c == '*' ? Op::Mul : Op::Div, std::move(n), parseUnary());
    }
    return n;
  }
  // unary := ('-' | '+') unary | power
  auto parseUnary() -> std::unique_ptr<Node> {
    const char c = peek();
    if (c == '-' || c == '+') {
      ++m_pos;
      auto n = parseUnary();
      return c == '-' ? make(Op::Neg, std::move(n)) : std::move(n);
    }
    return parsePower();
  }
  // power := atom ('^' unary)?   (right associative)
  auto parsePower() -> std::unique_ptr<Node> {
    auto n = parseAtom();
    if (peek() == '^') {
      ++m_pos;
      n = make(Op::Pow, std::move(n), parseUnary());
    }
    return n;
  }
  auto parseAtom() -> std::unique_ptr<Node> {
    const char c = peek();
    if (c == '(') {
      ++m_pos;
      auto n = parseSum();
      if (peek() != ')') fail("missing ')'");
      ++m_pos;
      return n;
    }
    if (std::isdigit((unsigned char)c) || c == '.') {
      const char* s = m_text.c_str() + m_pos;
      char* end = nullptr;
      const double v = std::strtod(s, &end);
      if (end == s) fail("bad number");
      m_pos += (size_t)(end - s);
      auto n = std::make_unique<Node>();
      n->op = Op::Num;
      n->value = v;
      return n;
    }
    if (std::isalpha((unsigned char)c) || c == '_') {
      size_t e = m_pos;
      while (e < m_text.size() && (std::isalnum((unsigned char)m_text[e]) || m_text[e] == '_')) ++e;
      const std::string name = m_text.substr(m_pos, e - m_pos);
      m_pos = e;
      if (peek() == '(') {
        ++m_pos;
        auto n = std::make_unique<Node>();
        n->op = Op::Call;
        n->fn = name;
        n->a = parseSum();
        if (peek() == ',') {
          ++m_pos;
          n->b = parseSum();
        }
        if (peek() != ')') fail("missing ')' after the arguments of " + name);
        ++m_pos;
        if (!known(name, n->b != nullptr)) fail("unknown function " + name);
        return n;
      }
      auto n = std::make_unique<Node>();
      if (name == "x" || name == "y" || name == "z") {
        n->op = Op::Var;
        n->var = name[0] - 'x';
      } else if (name == "pi") {
        n->op = Op::Num;
        n->value = M_PI;
      } else {
        fail("unknown symbol " + name);
      }
      return n;
    }
    fail(c ? std::string("unexpected '") + c + "'" : "unexpected end");
  }
  static auto known(const std::string& f, bool two) -> bool {
    static const char* one[] = {"sin", "cos", "tan", "sinh", "cosh", "tanh", "exp", "log", "sqrt", "abs", "asin",
                                "acos", "atan"};
    if (two) return f == "pow" || f == "atan2" || f == "min" || f == "max";
    for (const char* g : one)
      if (f == g) return true;
    return false;
  }
  static auto call(const std::string& f, double a, double b) -> double {
    if (f == "sin") return std::sin(a);
    if (f == "cos") return std::cos(a);
    if (f == "tan") return std::tan(a);
    if (f == "sinh") return std::sinh(a);
    if (f == "cosh") return std::cosh(a);
    if (f == "tanh") return std::tanh(a);
    if (f == "exp") return std::exp(a);
    if (f == "log") return std::log(a);
    if (f == "sqrt") return std::sqrt(a);
    if (f == "abs") return std::fabs(a);
    if (f == "asin") return std::asin(a);
    if (f == "acos") return std::acos(a);
    if (f == "atan") return std::atan(a);
    if (f == "pow") return std::pow(a, b);
    if (f == "atan2") return std::atan2(a, b);
    if (f == "min") return std::fmin(a, b);
    return std::fmax(a, b);
  }
  static auto evalNode(const Node& n, const double* v) -> double {
    switch (n.op) {
      case Op::Num: return n.value;
      case Op::Var: return v[n.var];
      case Op::Neg: return -evalNode(*n.a, v);
      case Op::Add: return evalNode(*n.a, v) + evalNode(*n.b, v);
      case Op::Sub: return evalNode(*n.a, v) - evalNode(*n.b, v);
      case Op::Mul: return evalNode(*n.a, v) * evalNode(*n.b, v);
      case Op::Div: return evalNode(*n.a, v) / evalNode(*n.b, v);
      case Op::Pow: return std::pow(evalNode(*n.a, v), evalNode(*n.b, v));
      case Op::Call: return call(n.fn, evalNode(*n.a, v), n.b ? evalNode(*n.b, v) : 0.0);
    }
    return 0;
  }
};

}  // namespace expr

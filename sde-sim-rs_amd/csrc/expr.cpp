// expr.cpp -- see expr.hpp. Grammar and semantics restated from fasteval 0.2.4 as used by
// src/func.rs:18-41 (parse + compile + eval with a BTreeMap<String, f64> namespace).
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <sstream>

#include "expr.hpp"

namespace sdesim {

namespace {

struct FnInfo { const char* name; Fn fn; int min_args; int max_args; };
const FnInfo kFns[] = {
    {"min", Fn::Min, 1, 1 << 20}, {"max", Fn::Max, 1, 1 << 20}, {"e", Fn::E, 0, 0},
    {"pi", Fn::Pi, 0, 0}, {"log", Fn::Log, 1, 2}, {"round", Fn::Round, 1, 2},
    {"int", Fn::Int, 1, 1}, {"ceil", Fn::Ceil, 1, 1}, {"floor", Fn::Floor, 1, 1},
    {"abs", Fn::Abs, 1, 1}, {"sign", Fn::Sign, 1, 1}, {"sin", Fn::Sin, 1, 1},
    {"cos", Fn::Cos, 1, 1}, {"tan", Fn::Tan, 1, 1}, {"asin", Fn::Asin, 1, 1},
    {"acos", Fn::Acos, 1, 1}, {"atan", Fn::Atan, 1, 1}, {"sinh", Fn::Sinh, 1, 1},
    {"cosh", Fn::Cosh, 1, 1}, {"tanh", Fn::Tanh, 1, 1}, {"asinh", Fn::Asinh, 1, 1},
    {"acosh", Fn::Acosh, 1, 1}, {"atanh", Fn::Atanh, 1, 1},
};

// Binding power of each binary operator: one level per operator, low to high.
int precedence(Op op) {
  switch (op) {
    case Op::Or: return 1;
    case Op::And: return 2;
    case Op::Ne: return 3;
    case Op::Eq: return 4;
    case Op::Ge: return 5;
    case Op::Le: return 6;
    case Op::Gt: return 7;
    case Op::Lt: return 8;
    case Op::Add: return 9;
    case Op::Mul: return 11;  // '*' ; '/' is 12 and is lowered to Mul(a, Inv(b))
    case Op::Mod: return 13;
    case Op::Pow: return 14;
    default: return 0;
  }
}

ExprPtr node(Op op) {
  ExprPtr e(new Expr());
  e->op = op;
  return e;
}

class Parser {
 public:
  explicit Parser(const std::string& s) : s_(s) {}

  ExprPtr parse() {
    ExprPtr e = binary(1);
    skip();
    if (i_ != s_.size()) fail("unparsed tokens remaining: \"" + s_.substr(i_) + "\"");
    return e;
  }

 private:
  const std::string& s_;
  size_t i_ = 0;

  [[noreturn]] void fail(const std::string& m) { throw ParseError(m); }
  void skip() { while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_; }
  bool starts(const char* t) const { return s_.compare(i_, std::strlen(t), t) == 0; }

  // Peek the next binary operator: returns its precedence (0 if none), the lowered op and
  // the token length. '-' is precedence 10 and '/' precedence 12.
  int peek_op(Op& op, size_t& len, bool& is_sub, bool& is_div) {
    skip();
    is_sub = is_div = false;
    if (i_ >= s_.size()) return 0;
    struct T { const char* tok; Op op; int prec; };
    static const T table[] = {
        {"||", Op::Or, 1}, {"&&", Op::And, 2}, {"!=", Op::Ne, 3}, {"==", Op::Eq, 4},
        {">=", Op::Ge, 5}, {"<=", Op::Le, 6}, {">", Op::Gt, 7}, {"<", Op::Lt, 8},
        {"+", Op::Add, 9}, {"-", Op::Add, 10}, {"*", Op::Mul, 11}, {"/", Op::Mul, 12},
        {"%", Op::Mod, 13}, {"^", Op::Pow, 14},
    };
    for (const T& t : table) {
      if (starts(t.tok)) {
        op = t.op;
        len = std::strlen(t.tok);
        is_sub = (t.tok[0] == '-');
        is_div = (t.tok[0] == '/');
        return t.prec;
      }
    }
    return 0;
  }

  // Precedence climbing: operators at or above `min_prec`.
  ExprPtr binary(int min_prec) {
    ExprPtr lhs = value();
    for (;;) {
      Op op; size_t len; bool is_sub, is_div;
      int prec = peek_op(op, len, is_sub, is_div);
      if (prec == 0 || prec < min_prec) return lhs;
      i_ += len;
      ExprPtr rhs = binary(op == Op::Pow ? prec : prec + 1);
      if (is_sub) { ExprPtr n = node(Op::Neg); n->args.push_back(std::move(rhs)); rhs = std::move(n); }
      if (is_div) { ExprPtr n = node(Op::Inv); n->args.push_back(std::move(rhs)); rhs = std::move(n); }
      ExprPtr b = node(op);
      b->args.push_back(std::move(lhs));
      b->args.push_back(std::move(rhs));
      lhs = std::move(b);
    }
  }

  ExprPtr value() {
    skip();
    if (i_ >= s_.size()) fail("unexpected end of expression");
    char c = s_[i_];
    if (c == '(') {
      ++i_;
      ExprPtr e = binary(1);
      skip();
      if (i_ >= s_.size() || s_[i_] != ')') fail("expected ')'");
      ++i_;
      return e;
    }
    if (c == '-' || c == '!') {
      ++i_;
      ExprPtr n = node(c == '-' ? Op::Neg : Op::Not);
      n->args.push_back(value());
      return n;
    }
    if (c == '+') { ++i_; return value(); }
    if (std::isdigit((unsigned char)c) || c == '.') return number();
    if (std::isalpha((unsigned char)c) || c == '_') return name_or_call();
    fail(std::string("unexpected character '") + c + "'");
  }

  ExprPtr number() {
    size_t b = i_;
    auto digits = [&] { while (i_ < s_.size() && std::isdigit((unsigned char)s_[i_])) ++i_; };
    digits();
    if (i_ < s_.size() && s_[i_] == '.') { ++i_; digits(); }
    if (i_ < s_.size() && (s_[i_] == 'e' || s_[i_] == 'E')) {
      size_t mark = i_++;
      if (i_ < s_.size() && (s_[i_] == '+' || s_[i_] == '-')) ++i_;
      if (i_ < s_.size() && std::isdigit((unsigned char)s_[i_])) digits(); else i_ = mark;
    }
    std::string tok = s_.substr(b, i_ - b);
    if (tok == ".") fail("invalid number");
    double v = std::strtod(tok.c_str(), nullptr);
    if (i_ < s_.size()) {
      double scale = 0.0;
      switch (s_[i_]) {
        case 'k': case 'K': scale = 1e3; break;
        case 'M': scale = 1e6; break;
        case 'G': scale = 1e9; break;
        case 'T': scale = 1e12; break;
        case 'm': scale = 1e-3; break;
        case 'u': scale = 1e-6; break;
        case 'n': scale = 1e-9; break;
        case 'p': scale = 1e-12; break;
        default: break;
      }
      bool word = i_ + 1 < s_.size() && (std::isalnum((unsigned char)s_[i_ + 1]) || s_[i_ + 1] == '_');
      if (scale != 0.0 && !word) { v *= scale; ++i_; }
    }
    ExprPtr n = node(Op::Const);
    n->value = v;
    return n;
  }

  ExprPtr name_or_call() {
    size_t b = i_;
    while (i_ < s_.size() && (std::isalnum((unsigned char)s_[i_]) || s_[i_] == '_')) ++i_;
    std::string name = s_.substr(b, i_ - b);
    skip();
    if (i_ >= s_.size() || s_[i_] != '(') {
      ExprPtr v = node(Op::Var);
      v->name = name;
      return v;
    }
    ++i_;
    ExprPtr call = node(Op::Call);
    call->name = name;
    skip();
    if (i_ < s_.size() && s_[i_] == ')') {
      ++i_;
    } else {
      for (;;) {
        call->args.push_back(binary(1));
        skip();
        if (i_ < s_.size() && s_[i_] == ',') { ++i_; continue; }
        if (i_ < s_.size() && s_[i_] == ')') { ++i_; break; }
        fail("expected ',' or ')' in call of " + name);
      }
    }
    call->fn = Fn::Lookup;
    for (const FnInfo& f : kFns) {
      if (name == f.name) {
        int n = (int)call->args.size();
        if (n < f.min_args || n > f.max_args) fail("wrong number of arguments for " + name + "()");
        call->fn = f.fn;
      }
    }
    return call;
  }
};

bool feq(double a, double b) { return std::fabs(a - b) <= 8.0 * 2.220446049250313e-16; }

double apply_fn(Fn fn, const std::vector<double>& a) {
  switch (fn) {
    case Fn::Min: { double r = a[0]; for (size_t k = 1; k < a.size(); ++k) r = std::fmin(r, a[k]); return r; }
    case Fn::Max: { double r = a[0]; for (size_t k = 1; k < a.size(); ++k) r = std::fmax(r, a[k]); return r; }
    case Fn::E: return 2.718281828459045;
    case Fn::Pi: return 3.141592653589793;
    case Fn::Log: return a.size() == 1 ? std::log10(a[0]) : std::log(a[1]) / std::log(a[0]);
    case Fn::Round: return a.size() == 1 ? std::round(a[0]) : std::round(a[1] / a[0]) * a[0];
    case Fn::Int: return std::trunc(a[0]);
    case Fn::Ceil: return std::ceil(a[0]);
    case Fn::Floor: return std::floor(a[0]);
    case Fn::Abs: return std::fabs(a[0]);
    case Fn::Sign: return std::isnan(a[0]) ? a[0] : std::copysign(1.0, a[0]);
    case Fn::Sin: return std::sin(a[0]);
    case Fn::Cos: return std::cos(a[0]);
    case Fn::Tan: return std::tan(a[0]);
    case Fn::Asin: return std::asin(a[0]);
    case Fn::Acos: return std::acos(a[0]);
    case Fn::Atan: return std::atan(a[0]);
    case Fn::Sinh: return std::sinh(a[0]);
    case Fn::Cosh: return std::cosh(a[0]);
    case Fn::Tanh: return std::tanh(a[0]);
    case Fn::Asinh: return std::asinh(a[0]);
    case Fn::Acosh: return std::acosh(a[0]);
    case Fn::Atanh: return std::atanh(a[0]);
    case Fn::Lookup: break;
  }
  throw ParseError("unsupported function");
}

void fold(ExprPtr& e) {
  for (auto& a : e->args) fold(a);
  if (e->op == Op::Const || e->op == Op::Var) return;
  if (e->op == Op::Call && e->fn == Fn::Lookup) return;
  for (auto& a : e->args)
    if (a->op != Op::Const) return;
  std::map<std::string, double> none;
  double v = eval_host(*e, none);
  ExprPtr c = node(Op::Const);
  c->value = v;
  e = std::move(c);
}

std::string lit(double v, bool f32) {
  if (std::isnan(v)) return f32 ? "__builtin_nanf(\"\")" : "__builtin_nan(\"\")";
  if (std::isinf(v)) {
    std::string s = f32 ? "__builtin_inff()" : "__builtin_inf()";
    return v < 0 ? "(-" + s + ")" : s;
  }
  char buf[64];
  // %a is exact: the device sees the same double the host folded.
  std::snprintf(buf, sizeof buf, "%a", v);
  std::string s = buf;
  if (f32) return "((float)" + s + ")";
  return "(" + s + ")";
}

}  // namespace

ExprPtr parse_expression(const std::string& text) {
  Parser p(text);
  ExprPtr e = p.parse();
  fold(e);
  return e;
}

double eval_host(const Expr& e, const std::map<std::string, double>& ns) {
  auto arg = [&](size_t k) { return eval_host(*e.args[k], ns); };
  auto truthy = [](double x) { return !feq(x, 0.0); };
  switch (e.op) {
    case Op::Const: return e.value;
    case Op::Var: {
      auto it = ns.find(e.name);
      if (it == ns.end()) throw std::runtime_error("undefined variable '" + e.name + "'");
      return it->second;
    }
    case Op::Neg: return -arg(0);
    case Op::Not: return truthy(arg(0)) ? 0.0 : 1.0;
    case Op::Inv: return 1.0 / arg(0);
    case Op::Or: { double a = arg(0); return truthy(a) ? a : arg(1); }
    case Op::And: { double a = arg(0); return truthy(a) ? arg(1) : a; }
    case Op::Ne: return feq(arg(0), arg(1)) ? 0.0 : 1.0;
    case Op::Eq: return feq(arg(0), arg(1)) ? 1.0 : 0.0;
    case Op::Ge: return arg(0) >= arg(1) ? 1.0 : 0.0;
    case Op::Le: return arg(0) <= arg(1) ? 1.0 : 0.0;
    case Op::Gt: return arg(0) > arg(1) ? 1.0 : 0.0;
    case Op::Lt: return arg(0) < arg(1) ? 1.0 : 0.0;
    case Op::Add: return arg(0) + arg(1);
    case Op::Mul: return arg(0) * arg(1);
    case Op::Mod: return std::fmod(arg(0), arg(1));
    case Op::Pow: return std::pow(arg(0), arg(1));
    case Op::Call: {
      std::vector<double> a;
      for (size_t k = 0; k < e.args.size(); ++k) a.push_back(arg(k));
      if (e.fn == Fn::Lookup) {
        auto it = ns.find(e.name);
        if (it == ns.end()) throw std::runtime_error("undefined variable '" + e.name + "'");
        return it->second;
      }
      return apply_fn(e.fn, a);
    }
  }
  return 0.0;
}

std::string canonical(const Expr& e) {
  std::ostringstream o;
  switch (e.op) {
    case Op::Const: { char b[64]; std::snprintf(b, sizeof b, "%a", e.value); o << b; break; }
    case Op::Var: o << "$" << e.name; break;
    case Op::Call:
      o << (e.fn == Fn::Lookup ? "@" : "") << e.name << "(";
      for (size_t k = 0; k < e.args.size(); ++k) o << (k ? "," : "") << canonical(*e.args[k]);
      o << ")";
      break;
    default:
      o << "[" << (int)e.op;
      for (auto& a : e.args) o << " " << canonical(*a);
      o << "]";
  }
  return o.str();
}

void collect_names(const Expr& e, std::vector<std::string>& out) {
  if (e.op == Op::Var || (e.op == Op::Call && e.fn == Fn::Lookup)) out.push_back(e.name);
  for (auto& a : e.args) collect_names(*a, out);
}

std::string emit(const Expr& e, const EmitEnv& env) {
  const bool f = env.f32;
  auto A = [&](size_t k) { return emit(*e.args[k], env); };
  const std::string R = f ? "float" : "double";
  switch (e.op) {
    case Op::Const: return lit(e.value, f);
    case Op::Var: return env.resolve(e.name);
    case Op::Neg: return "(-" + A(0) + ")";
    case Op::Not: return "sde_not(" + A(0) + ")";
    case Op::Inv: return "(" + lit(1.0, f) + " / " + A(0) + ")";
    case Op::Or: return "sde_or(" + A(0) + ", " + A(1) + ")";
    case Op::And: return "sde_and(" + A(0) + ", " + A(1) + ")";
    case Op::Ne: return "sde_ne(" + A(0) + ", " + A(1) + ")";
    case Op::Eq: return "sde_eq(" + A(0) + ", " + A(1) + ")";
    case Op::Ge: return "((" + A(0) + " >= " + A(1) + ") ? " + lit(1.0, f) + " : " + lit(0.0, f) + ")";
    case Op::Le: return "((" + A(0) + " <= " + A(1) + ") ? " + lit(1.0, f) + " : " + lit(0.0, f) + ")";
    case Op::Gt: return "((" + A(0) + " > " + A(1) + ") ? " + lit(1.0, f) + " : " + lit(0.0, f) + ")";
    case Op::Lt: return "((" + A(0) + " < " + A(1) + ") ? " + lit(1.0, f) + " : " + lit(0.0, f) + ")";
    case Op::Add: return "(" + A(0) + " + " + A(1) + ")";
    case Op::Mul: return "(" + A(0) + " * " + A(1) + ")";
    case Op::Mod: return "sde_fmod(" + A(0) + ", " + A(1) + ")";
    case Op::Pow: {
      const Expr& ex = *e.args[1];
      if (ex.op == Op::Const) {
        // Exact rewrites of pow for these exponents (IEEE sqrt and one multiply are correctly
        // rounded, as pow is for them); the sign of a zero result may differ for -0.0 bases.
        if (ex.value == 0.5) return "sde_sqrt(" + A(0) + ")";
        if (ex.value == 1.0) return "(" + lit(1.0, f) + " * " + A(0) + ")";
        if (ex.value == 2.0) return "sde_sq(" + A(0) + ")";
      }
      return "sde_pow(" + A(0) + ", " + A(1) + ")";
    }
    case Op::Call: {
      if (e.fn == Fn::Lookup) return env.resolve(e.name);
      auto fn1 = [&](const char* n) { return std::string("sde_") + n + "(" + A(0) + ")"; };
      switch (e.fn) {
        case Fn::Min: case Fn::Max: {
          std::string r = A(0);
          for (size_t k = 1; k < e.args.size(); ++k)
            r = std::string(e.fn == Fn::Min ? "sde_fmin(" : "sde_fmax(") + r + ", " + A(k) + ")";
          return r;
        }
        case Fn::E: return lit(2.718281828459045, f);
        case Fn::Pi: return lit(3.141592653589793, f);
        case Fn::Log:
          if (e.args.size() == 1) return fn1("log10");
          return "(sde_log(" + A(1) + ") / sde_log(" + A(0) + "))";
        case Fn::Round:
          if (e.args.size() == 1) return fn1("round");
          return "(sde_round(" + A(1) + " / " + A(0) + ") * " + A(0) + ")";
        case Fn::Int: return fn1("trunc");
        case Fn::Ceil: return fn1("ceil");
        case Fn::Floor: return fn1("floor");
        case Fn::Abs: return fn1("fabs");
        case Fn::Sign: return fn1("sign");
        case Fn::Sin: return fn1("sin");
        case Fn::Cos: return fn1("cos");
        case Fn::Tan: return fn1("tan");
        case Fn::Asin: return fn1("asin");
        case Fn::Acos: return fn1("acos");
        case Fn::Atan: return fn1("atan");
        case Fn::Sinh: return fn1("sinh");
        case Fn::Cosh: return fn1("cosh");
        case Fn::Tanh: return fn1("tanh");
        case Fn::Asinh: return fn1("asinh");
        case Fn::Acosh: return fn1("acosh");
        case Fn::Atanh: return fn1("atanh");
        default: break;
      }
    }
  }
  throw ParseError("cannot emit expression");
}

}  // namespace sdesim

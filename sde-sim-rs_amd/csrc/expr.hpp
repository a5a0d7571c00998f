// expr.hpp -- coefficient expressions: parser, constant folder and HIP source emitter.
//
// Replaces the reference's `Function` (src/func.rs:5-42), which parses and compiles each
// coefficient with the fasteval crate (0.2.4) and evaluates it against a name -> f64 cache.
// Here an expression is parsed once on the host into a small tree whose variables are bound
// to slots (`t` or a process index); the tree is then emitted as straight-line HIP C++ that the
// fused path kernel inlines, so nothing is interpreted on the device.
//
// Semantics follow fasteval 0.2.4 as documented: binary precedence from low to high
// || && != == >= <= > < + - * / % ^ (each its own level, ^ right-associative, the rest left);
// unary - + ! bind to the following value; `a - b` is a + (-b) and `a / b` is a * (1/b);
// comparisons give 1.0 / 0.0 with == and != using an absolute tolerance of 8 * DBL_EPSILON;
// && and || return one of their operands; log(x) is base 10 and log(b, x) = ln x / ln b;
// round(m, x) = round(x / m) * m; literals accept the SI suffixes k K M G T m u n p; a call of a
// name that is not a builtin reads that name's value and ignores its arguments.
#pragma once
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace sdesim {

struct ParseError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

enum class Op : uint8_t {
  Const, Var, Neg, Not, Inv,
  Or, And, Ne, Eq, Ge, Le, Gt, Lt, Add, Mul, Mod, Pow,
  Call
};

enum class Fn : uint8_t {
  Min, Max, E, Pi, Log, Round, Int, Ceil, Floor, Abs, Sign,
  Sin, Cos, Tan, Asin, Acos, Atan, Sinh, Cosh, Tanh, Asinh, Acosh, Atanh,
  Lookup  // a call of a non-builtin name: value of that name, arguments ignored
};

struct Expr {
  Op op = Op::Const;
  Fn fn = Fn::Lookup;
  double value = 0.0;     // Const
  std::string name;       // Var / Call(Lookup)
  std::vector<std::unique_ptr<Expr>> args;
};
using ExprPtr = std::unique_ptr<Expr>;

// Parse `text` (throws ParseError with a message) and fold constant subtrees.
ExprPtr parse_expression(const std::string& text);

// Canonical text of a tree (used as a memo key and for kernel-cache keys).
std::string canonical(const Expr& e);

// Names the tree reads (variables and lookup calls).
void collect_names(const Expr& e, std::vector<std::string>& out);

// Host evaluation in f64 with a name resolver (used by constant folding and tests).
double eval_host(const Expr& e, const std::map<std::string, double>& ns);

// Emit a HIP C++ expression. `resolve(name)` returns the C++ text of the variable (throws on
// an unknown name); `f32` selects float arithmetic.
struct EmitEnv {
  bool f32 = false;
  std::function<std::string(const std::string&)> resolve;
};
std::string emit(const Expr& e, const EmitEnv& env);

}  // namespace sdesim

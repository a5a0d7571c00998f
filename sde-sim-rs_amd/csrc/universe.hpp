// universe.hpp -- the process model built from the equation DSL.
//
// Replaces proc::util::parse_equations / parse_single_equation / build_incrementor
// (src/proc/util.rs:52-166) and ProcessUniverse / LevyProcess / AlgebraicProcess
// (src/proc/mod.rs:7-90). The grammar, including its quirks, is the reference's:
//   * an equation must split on '=' into exactly two parts;
//   * a left side starting with 'd' is a Levy process named without the 'd'; anything else is
//     an algebraic process whose whole right side is one expression;
//   * each Levy term is "(coefficient) * incrementor"; the text between terms (signs included)
//     is skipped, and parsing stops quietly at a ')' that is not followed by '*';
//   * an incrementor is `dt`, `dW...` (the token up to the next space) or `dN...(lambda)`;
//   * every distinct non-dt incrementor text gets the next stochastic index in order of first
//     appearance (a dN key includes its lambda text).
#pragma once
#include <map>
#include <string>
#include <vector>

#include "expr.hpp"

namespace sdesim {

enum class IncKind { Time = 0, Wiener = 1, Jump = 2 };

struct Term {
  ExprPtr coef;
  IncKind kind = IncKind::Time;
  int index = -1;       // stochastic registry index (Wiener / Jump)
  ExprPtr lambda;       // Jump intensity
};

struct ProcessDef {
  std::string name;
  bool levy = true;
  std::vector<Term> terms;  // Levy
  ExprPtr expr;             // algebraic
};

struct Universe {
  std::vector<ProcessDef> processes;
  std::map<std::string, int> registry;      // process name -> last index with that name
  std::vector<std::string> increments;      // stochastic registry in index order
  std::vector<int> levy, algebraic;         // process indices in equation order
  std::vector<std::string> equations;       // source text (cache keys)

  int P() const { return (int)processes.size(); }
  int M() const { return (int)increments.size(); }
};

// Throws ParseError with the reference's message texts.
Universe parse_equations(const std::vector<std::string>& equations);

}  // namespace sdesim

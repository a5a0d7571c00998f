// universe.cpp -- equation DSL (src/proc/util.rs:16-166) into a Universe (src/proc/mod.rs:61-90).
#include <cctype>

#include "universe.hpp"

namespace sdesim {

namespace {

std::string strip(const std::string& s, bool left, bool right) {
  size_t b = 0, e = s.size();
  if (left) while (b < e && std::isspace((unsigned char)s[b])) ++b;
  if (right) while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
  return s.substr(b, e - b);
}

// Match a balanced "( ... )" group starting at s[at] == '('. On success `inner` holds the
// enclosed text and the return value is the offset just past the closing ')'; std::string::npos
// when the group never closes (src/proc/util.rs:16-38).
size_t match_group(const std::string& s, size_t at, std::string& inner) {
  int depth = 0;
  for (size_t k = at; k < s.size(); ++k) {
    char c = s[k];
    if (c == '(') {
      ++depth;
    } else if (c == ')') {
      if (--depth == 0) {
        inner = s.substr(at + 1, k - at - 1);
        return k + 1;
      }
    }
  }
  return std::string::npos;
}

ExprPtr compile_or(const std::string& text, const std::string& prefix) {
  try {
    return parse_expression(text);
  } catch (const ParseError& e) {
    throw ParseError(prefix + "Parse Error: " + e.what());
  }
}

// src/proc/util.rs:136-166
Term make_term(ExprPtr coef, const std::string& inc, std::vector<std::string>& registry) {
  Term t;
  t.coef = std::move(coef);
  if (inc == "dt") return t;
  int idx = -1;
  for (size_t k = 0; k < registry.size(); ++k)
    if (registry[k] == inc) idx = (int)k;
  if (idx < 0) {
    idx = (int)registry.size();
    registry.push_back(inc);
  }
  t.index = idx;
  if (inc.compare(0, 2, "dW") == 0) {
    t.kind = IncKind::Wiener;
    return t;
  }
  if (inc.compare(0, 2, "dN") == 0) {
    size_t open = inc.find('(');
    if (open == std::string::npos) throw ParseError("Missing '(' in dN incrementor");
    std::string inner;
    if (match_group(inc, open, inner) == std::string::npos)
      throw ParseError("Unbalanced parentheses in jump term: " + inc);
    std::string lam = strip(inner, true, true);
    t.kind = IncKind::Jump;
    t.lambda = compile_or(lam, "Math error in jump lambda '" + lam + "': ");
    return t;
  }
  throw ParseError("Unknown incrementor type: " + inc);
}

// src/proc/util.rs:68-134
ProcessDef parse_one(const std::string& eq, std::vector<std::string>& registry) {
  size_t eq1 = eq.find('=');
  if (eq1 == std::string::npos || eq.find('=', eq1 + 1) != std::string::npos) throw ParseError("Missing '='");
  std::string lhs = strip(eq.substr(0, eq1), true, true);
  std::string rhs = strip(eq.substr(eq1 + 1), true, true);
  ProcessDef p;
  bool levy = !lhs.empty() && lhs[0] == 'd';
  p.name = levy ? lhs.substr(1) : lhs;
  p.levy = levy;
  if (!levy) {
    p.expr = compile_or(rhs, "");
    return p;
  }
  std::string rest = rhs;
  for (;;) {
    size_t open = rest.find('(');
    if (open == std::string::npos) break;
    std::string coef;
    size_t after = match_group(rest, open, coef);
    if (after == std::string::npos) throw ParseError("Unbalanced parentheses in coefficient");
    std::string tail = strip(rest.substr(after), true, false);
    if (tail.empty() || tail[0] != '*') break;
    std::string inc_text = strip(tail.substr(1), true, false);
    std::string inc, next;
    if (inc_text.compare(0, 2, "dN") == 0) {
      size_t lp = inc_text.find('(');
      if (lp == std::string::npos) throw ParseError("dN missing opening bracket");
      std::string lam;
      size_t end = match_group(inc_text, lp, lam);
      if (end == std::string::npos) throw ParseError("Unbalanced parentheses in dN intensity");
      inc = inc_text.substr(0, end);
      next = inc_text.substr(end);
    } else {
      size_t sp = inc_text.find(' ');
      if (sp == std::string::npos) sp = inc_text.size();
      inc = inc_text.substr(0, sp);
      next = inc_text.substr(sp);
    }
    ExprPtr c = compile_or(strip(coef, true, true), "Math error in coefficient: ");
    p.terms.push_back(make_term(std::move(c), inc, registry));
    rest = next;
  }
  return p;
}

}  // namespace

Universe parse_equations(const std::vector<std::string>& equations) {
  Universe u;
  u.equations = equations;
  for (const std::string& e : equations) u.processes.push_back(parse_one(e, u.increments));
  for (size_t k = 0; k < u.processes.size(); ++k) {
    u.registry[u.processes[k].name] = (int)k;
    (u.processes[k].levy ? u.levy : u.algebraic).push_back((int)k);
  }
  return u;
}

}  // namespace sdesim

// sdesim_oracle.cpp -- CPU restatement of the sde-sim-rs simulator (TEST INFRASTRUCTURE ONLY).
//
// This file is the oracle: the reference the GPU product is compared against, and the timed CPU
// baseline. It is never linked into, loaded by or called from the product library. Only the
// tests, the smoke check and the benchmark's CPU-baseline leg load it (through ctypes).
//
// It restates the reference's algorithm with the reference's own structure, so it is also a
// faithful "reference CPU path" for timing:
//   * per-scenario filtration `raw[T*P]`, time-major          (src/filtration.rs:22-64)
//   * a name-keyed variable cache refreshed only on time change  (src/filtration.rs:70-79, src/func.rs:32-42)
//   * a tree-walking expression interpreter with name lookup    (src/func.rs:5-42; fasteval semantics, see below)
//   * Incrementor objects dt / dW / dN                           (src/proc/increment.rs:26-157)
//   * A&S 26.2.23 inverse normal and sequential inverse Poisson  (src/proc/increment.rs:160-200)
//   * PseudoRng: ChaCha8 seeded by seed_from_u64, m f64 per step  (src/rng/pseudo.rs:7-60)
//   * SobolRng: Joe-Kuo Gray-code points + per-path f64 random shift (src/rng/sobol.rs:7-79)
//   * Euler and "Runge-Kutta" steps                              (src/sim/euler.rs:5-37, src/sim/runge_kutta.rs:5-107)
//   * simulate(): seed s+seed per scenario, threads over scenarios (src/sim/mod.rs:20-92)
//   * equation DSL parser                                        (src/proc/util.rs:16-166)
//
// Two additions that have no reference counterpart are restated here from their published
// definitions because the product implements them: the digital (XOR) shift of Sobol points
// (one 32-bit mask per dimension, shared by all paths of one replicate) and an "injected
// uniforms" generator used to drive both sides with identical inputs.
//
// Parity status of the third-party pieces (their sources are not available offline):
//   * ChaCha core: pinned by the RFC 8439 ChaCha20 block vector (rounds is a parameter).
//   * rand_core seed_from_u64 (PCG32 expansion) and rand's f64 conversion `(u64 >> 11) * 2^-53`:
//     restated from the published rand_core / rand 0.9 algorithms; parity unpinned.
//   * Sobol points: pinned against scipy.stats.qmc.Sobol(bits=32) on the same Joe-Kuo table.
//     The sobol crate's starting point (skip(5) => first point index 5) and f64 conversion
//     (x * 2^-32) are restated assumptions; parity unpinned.
//   * fasteval expression semantics beyond ordinary arithmetic: restated from fasteval 0.2.4's
//     documented grammar (precedence ^ % / * - + < > <= >= == != && ||, unary operators bind to
//     the following value, `a-b` as a+(-b), `a/b` as a*(1/b), log = log10, == with 8*eps
//     tolerance, SI suffixes on literals); parity unpinned beyond the closed-form checks.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <map>
#include <unordered_map>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>
#include <atomic>
#include <algorithm>

namespace orc {

// ------------------------------------------------------------------------------------------
// ChaCha (rand_chacha layout: words 12-13 = 64-bit block counter, 14-15 = 64-bit stream id)
// ------------------------------------------------------------------------------------------
static inline uint32_t rotl(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }

static void chacha_block(const uint32_t in[16], int rounds, uint32_t out[16]) {
  uint32_t x[16];
  for (int i = 0; i < 16; ++i) x[i] = in[i];
  auto qr = [&](int a, int b, int c, int d) {
    x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 16);
    x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 12);
    x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 8);
    x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 7);
  };
  for (int r = 0; r < rounds; r += 2) {
    qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15);
    qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14);
  }
  for (int i = 0; i < 16; ++i) out[i] = x[i] + in[i];
}

// One PCG32 output as rand_core's seed_from_u64 draws it: advance the LCG (multiplier
// 6364136223846793005, increment `inc`) first, then the XSH-RR output function of the new state.
static uint32_t pcg32_advance_output(uint64_t& state, uint64_t inc) {
  state = state * 6364136223846793005ULL + inc;
  uint32_t xorshifted = (uint32_t)(((state >> 18) ^ state) >> 27);
  uint32_t rot = (uint32_t)(state >> 59);
  return (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
}

// rand_core::SeedableRng::seed_from_u64: PCG32 with increment 11634580027462260723, one u32 per
// 4 seed bytes, little endian.
static void seed_from_u64(uint64_t state, uint32_t key[8]) {
  for (int i = 0; i < 8; ++i) key[i] = pcg32_advance_output(state, 11634580027462260723ULL);
}

// ChaCha8Rng::seed_from_u64(seed) as a u32 word stream; next_f64 = rand's StandardUniform.
struct ChaCha8Stream {
  uint32_t st[16];
  uint32_t buf[16];
  int pos = 16;
  uint64_t counter = 0;
  explicit ChaCha8Stream(uint64_t seed) {
    st[0] = 0x61707865; st[1] = 0x3320646e; st[2] = 0x79622d32; st[3] = 0x6b206574;
    seed_from_u64(seed, st + 4);
    st[12] = st[13] = st[14] = st[15] = 0;
  }
  uint32_t next_u32() {
    if (pos >= 16) {
      st[12] = (uint32_t)counter; st[13] = (uint32_t)(counter >> 32);
      chacha_block(st, 8, buf);
      ++counter; pos = 0;
    }
    return buf[pos++];
  }
  uint64_t next_u64() { uint64_t lo = next_u32(); uint64_t hi = next_u32(); return (hi << 32) | lo; }
  double next_f64() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }
};

// ------------------------------------------------------------------------------------------
// Sobol (Joe-Kuo D6 table as shipped by scipy: poly includes leading and trailing terms)
// ------------------------------------------------------------------------------------------
struct SobolTable {
  int dims = 0;
  std::vector<uint32_t> v;  // [dims][32], v[d*32+k] multiplies bit k of gray(n)
};
static SobolTable g_sobol;

static void sobol_build(int dims, const int64_t* poly, const int64_t* vinit /* [dims][18] */) {
  g_sobol.dims = dims;
  g_sobol.v.assign((size_t)dims * 32, 0);
  for (int d = 0; d < dims; ++d) {
    uint64_t m[33];
    if (d == 0) {
      for (int k = 1; k <= 32; ++k) m[k] = 1;
    } else {
      int64_t p = poly[d];
      int s = 0;
      while ((p >> (s + 1)) != 0) ++s;  // degree
      for (int k = 1; k <= s && k <= 32; ++k) m[k] = (uint64_t)vinit[(size_t)d * 18 + (k - 1)];
      for (int k = s + 1; k <= 32; ++k) {
        uint64_t val = m[k - s] ^ (m[k - s] << s);
        for (int i = 1; i < s; ++i)
          if ((p >> (s - i)) & 1) val ^= (m[k - i] << i);
        m[k] = val;
      }
    }
    for (int k = 1; k <= 32; ++k) g_sobol.v[(size_t)d * 32 + (k - 1)] = (uint32_t)(m[k] << (32 - k));
  }
}

static inline uint32_t sobol_point(uint64_t n, int d) {
  uint64_t g = n ^ (n >> 1);
  uint32_t x = 0;
  const uint32_t* vd = &g_sobol.v[(size_t)d * 32];
  for (int k = 0; g != 0 && k < 32; ++k, g >>= 1)
    if (g & 1) x ^= vd[k];
  return x;
}

// ------------------------------------------------------------------------------------------
// Inverse CDFs (src/proc/increment.rs:160-200)
// ------------------------------------------------------------------------------------------
static double inv_normal(double p) {
  double t = (p < 0.5) ? std::sqrt(-2.0 * std::log(p)) : std::sqrt(-2.0 * std::log(1.0 - p));
  const double c0 = 2.515517, c1 = 0.802853, c2 = 0.010328;
  const double d1 = 1.432788, d2 = 0.189269, d3 = 0.001308;
  double x = t - ((c2 * t + c1) * t + c0) / (((d3 * t + d2) * t + d1) * t + 1.0);
  return (p < 0.5) ? -x : x;
}

static uint64_t inv_poisson(double u, double lambda) {
  if (lambda <= 0.0) return 0;
  double p = std::exp(-lambda);
  double f = p;
  uint64_t k = 0;
  while (u > f && k < 200) {
    k += 1;
    p *= lambda / (double)k;
    f += p;
  }
  return k;
}

// ------------------------------------------------------------------------------------------
// Expression interpreter (fasteval semantics as documented in the header)
// ------------------------------------------------------------------------------------------
struct EvalError : std::runtime_error { using std::runtime_error::runtime_error; };

struct Node {
  enum Kind { Const, Var, Neg, Not, Inv, Bin, Call } kind;
  double c = 0;
  std::string name;  // Var / Call
  int op = 0;        // Bin: operator id
  std::vector<std::unique_ptr<Node>> kids;
};
using NodeP = std::unique_ptr<Node>;

enum BinOp { OR = 1, AND, NE, EQ, GE, LE, GT, LT, ADD, SUB, MUL, DIV, MOD, POW };

static bool f_eq(double a, double b) { return std::fabs(a - b) <= 8.0 * 2.220446049250313e-16; }
static bool truthy(double a) { return !f_eq(a, 0.0); }

typedef std::map<std::string, double> Namespace;

static double eval(const Node* n, const Namespace& ns) {
  switch (n->kind) {
    case Node::Const: return n->c;
    case Node::Var: {
      auto it = ns.find(n->name);
      if (it == ns.end()) throw EvalError("Undefined(\"" + n->name + "\")");
      return it->second;
    }
    case Node::Neg: return -eval(n->kids[0].get(), ns);
    case Node::Not: return truthy(eval(n->kids[0].get(), ns)) ? 0.0 : 1.0;
    case Node::Inv: return 1.0 / eval(n->kids[0].get(), ns);
    case Node::Bin: {
      double a = eval(n->kids[0].get(), ns);
      if (n->op == AND) return truthy(a) ? eval(n->kids[1].get(), ns) : a;
      if (n->op == OR) return truthy(a) ? a : eval(n->kids[1].get(), ns);
      double b = eval(n->kids[1].get(), ns);
      switch (n->op) {
        case ADD: return a + b;
        case MUL: return a * b;
        case MOD: return std::fmod(a, b);
        case POW: return std::pow(a, b);
        case LT: return a < b ? 1.0 : 0.0;
        case LE: return a <= b ? 1.0 : 0.0;
        case GT: return a > b ? 1.0 : 0.0;
        case GE: return a >= b ? 1.0 : 0.0;
        case EQ: return f_eq(a, b) ? 1.0 : 0.0;
        case NE: return f_eq(a, b) ? 0.0 : 1.0;
      }
      throw EvalError("bad op");
    }
    case Node::Call: {
      const std::string& f = n->name;
      std::vector<double> a;
      for (auto& k : n->kids) a.push_back(eval(k.get(), ns));
      size_t na = a.size();
      if (f == "min" || f == "max") {
        double r = a[0];
        for (size_t i = 1; i < na; ++i) r = (f == "min") ? std::fmin(r, a[i]) : std::fmax(r, a[i]);
        return r;
      }
      if (f == "e") return 2.718281828459045;
      if (f == "pi") return 3.141592653589793;
      if (f == "log") return na == 1 ? std::log10(a[0]) : std::log(a[1]) / std::log(a[0]);
      if (f == "round") return na == 1 ? std::round(a[0]) : std::round(a[1] / a[0]) * a[0];
      if (f == "int") return std::trunc(a[0]);
      if (f == "ceil") return std::ceil(a[0]);
      if (f == "floor") return std::floor(a[0]);
      if (f == "abs") return std::fabs(a[0]);
      if (f == "sign") return std::isnan(a[0]) ? a[0] : std::copysign(1.0, a[0]);
      if (f == "sin") return std::sin(a[0]);
      if (f == "cos") return std::cos(a[0]);
      if (f == "tan") return std::tan(a[0]);
      if (f == "asin") return std::asin(a[0]);
      if (f == "acos") return std::acos(a[0]);
      if (f == "atan") return std::atan(a[0]);
      if (f == "sinh") return std::sinh(a[0]);
      if (f == "cosh") return std::cosh(a[0]);
      if (f == "tanh") return std::tanh(a[0]);
      if (f == "asinh") return std::asinh(a[0]);
      if (f == "acosh") return std::acosh(a[0]);
      if (f == "atanh") return std::atanh(a[0]);
      // A call of any other name is a namespace lookup that ignores its arguments.
      auto it = ns.find(f);
      if (it == ns.end()) throw EvalError("Undefined(\"" + f + "\")");
      return it->second;
    }
  }
  throw EvalError("bad node");
}

struct ExprParser {
  const std::string& s;
  size_t i = 0;
  explicit ExprParser(const std::string& str) : s(str) {}
  void ws() { while (i < s.size() && std::isspace((unsigned char)s[i])) ++i; }
  [[noreturn]] void fail(const std::string& m) { throw std::runtime_error(m); }

  static NodeP mk(Node::Kind k) { NodeP n(new Node()); n->kind = k; return n; }

  NodeP value() {
    ws();
    if (i >= s.size()) fail("unexpected end of expression");
    char ch = s[i];
    if (ch == '(') {
      ++i;
      NodeP e = expr();
      ws();
      if (i >= s.size() || s[i] != ')') fail("expected ')'");
      ++i;
      return e;
    }
    if (ch == '-') { ++i; NodeP n = mk(Node::Neg); n->kids.push_back(value()); return n; }
    if (ch == '+') { ++i; return value(); }
    if (ch == '!') { ++i; NodeP n = mk(Node::Not); n->kids.push_back(value()); return n; }
    if (std::isdigit((unsigned char)ch) || ch == '.') return number();
    if (std::isalpha((unsigned char)ch) || ch == '_') {
      size_t b = i;
      while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '_')) ++i;
      std::string name = s.substr(b, i - b);
      ws();
      if (i < s.size() && s[i] == '(') {
        ++i;
        NodeP n = mk(Node::Call);
        n->name = name;
        ws();
        if (i < s.size() && s[i] == ')') { ++i; }
        else {
          for (;;) {
            n->kids.push_back(expr());
            ws();
            if (i < s.size() && s[i] == ',') { ++i; continue; }
            if (i < s.size() && s[i] == ')') { ++i; break; }
            fail("expected ',' or ')' in call of " + name);
          }
        }
        size_t na = n->kids.size();
        if ((name == "e" || name == "pi") && na != 0) fail(name + "() takes no arguments");
        if ((name == "min" || name == "max") && na < 1) fail(name + "() needs arguments");
        if ((name == "log" || name == "round") && (na < 1 || na > 2)) fail(name + "() takes 1 or 2 arguments");
        static const char* unary[] = {"int", "ceil", "floor", "abs", "sign", "sin", "cos", "tan", "asin",
                                      "acos", "atan", "sinh", "cosh", "tanh", "asinh", "acosh", "atanh"};
        for (const char* u : unary)
          if (name == u && na != 1) fail(name + "() takes 1 argument");
        return n;
      }
      NodeP n = mk(Node::Var);
      n->name = name;
      return n;
    }
    fail(std::string("unexpected character '") + ch + "'");
  }

  NodeP number() {
    size_t b = i;
    while (i < s.size() && std::isdigit((unsigned char)s[i])) ++i;
    if (i < s.size() && s[i] == '.') { ++i; while (i < s.size() && std::isdigit((unsigned char)s[i])) ++i; }
    if (i < s.size() && (s[i] == 'e' || s[i] == 'E')) {
      size_t save = i;
      ++i;
      if (i < s.size() && (s[i] == '+' || s[i] == '-')) ++i;
      if (i < s.size() && std::isdigit((unsigned char)s[i])) { while (i < s.size() && std::isdigit((unsigned char)s[i])) ++i; }
      else i = save;
    }
    std::string tok = s.substr(b, i - b);
    if (tok == ".") fail("bad number");
    double v = std::strtod(tok.c_str(), nullptr);
    if (i < s.size()) {
      double mult = 0;
      char c = s[i];
      switch (c) {
        case 'k': case 'K': mult = 1e3; break;
        case 'M': mult = 1e6; break;
        case 'G': mult = 1e9; break;
        case 'T': mult = 1e12; break;
        case 'm': mult = 1e-3; break;
        case 'u': mult = 1e-6; break;
        case 'n': mult = 1e-9; break;
        case 'p': mult = 1e-12; break;
      }
      bool ident_follows = (i + 1 < s.size()) && (std::isalnum((unsigned char)s[i + 1]) || s[i + 1] == '_');
      if (mult != 0 && !ident_follows) { v *= mult; ++i; }
    }
    NodeP n = mk(Node::Const);
    n->c = v;
    return n;
  }

  bool binop(int& op, size_t& len) {
    ws();
    if (i >= s.size()) return false;
    auto at = [&](const char* t) { return s.compare(i, std::strlen(t), t) == 0; };
    if (at("||")) { op = OR; len = 2; return true; }
    if (at("&&")) { op = AND; len = 2; return true; }
    if (at("!=")) { op = NE; len = 2; return true; }
    if (at("==")) { op = EQ; len = 2; return true; }
    if (at(">=")) { op = GE; len = 2; return true; }
    if (at("<=")) { op = LE; len = 2; return true; }
    if (at(">")) { op = GT; len = 1; return true; }
    if (at("<")) { op = LT; len = 1; return true; }
    if (at("+")) { op = ADD; len = 1; return true; }
    if (at("-")) { op = SUB; len = 1; return true; }
    if (at("*")) { op = MUL; len = 1; return true; }
    if (at("/")) { op = DIV; len = 1; return true; }
    if (at("%")) { op = MOD; len = 1; return true; }
    if (at("^")) { op = POW; len = 1; return true; }
    return false;
  }

  // Flat "value (op value)*" list, reduced by precedence: highest level first; all levels
  // reduce left to right except ^, which reduces right to left.
  NodeP expr() {
    std::vector<NodeP> vals;
    std::vector<int> ops;
    vals.push_back(value());
    int op; size_t len;
    while (binop(op, len)) {
      i += len;
      ops.push_back(op);
      vals.push_back(value());
    }
    for (int level = POW; level >= OR; --level) {
      if (level == POW) {
        for (int k = (int)ops.size() - 1; k >= 0; --k)
          if (ops[k] == POW) combine(vals, ops, k);
      } else {
        for (size_t k = 0; k < ops.size();) {
          if (ops[k] == level) combine(vals, ops, k); else ++k;
        }
      }
    }
    return std::move(vals[0]);
  }

  static void combine(std::vector<NodeP>& vals, std::vector<int>& ops, size_t k) {
    int op = ops[k];
    NodeP a = std::move(vals[k]), b = std::move(vals[k + 1]);
    NodeP n = mk(Node::Bin);
    if (op == SUB) {  // a - b == a + (-b)
      NodeP nb = mk(Node::Neg); nb->kids.push_back(std::move(b)); b = std::move(nb); op = ADD;
    } else if (op == DIV) {  // a / b == a * (1/b)
      NodeP nb = mk(Node::Inv); nb->kids.push_back(std::move(b)); b = std::move(nb); op = MUL;
    }
    n->op = op;
    n->kids.push_back(std::move(a));
    n->kids.push_back(std::move(b));
    vals[k] = std::move(n);
    vals.erase(vals.begin() + k + 1);
    ops.erase(ops.begin() + k);
  }

  NodeP parse() {
    NodeP e = expr();
    ws();
    if (i != s.size()) fail("unparsed tokens remaining: \"" + s.substr(i) + "\"");
    return e;
  }
};

// Constant folding of fully constant subtrees (evaluation order unchanged).
static void fold(NodeP& n) {
  for (auto& k : n->kids) fold(k);
  if (n->kind == Node::Const || n->kind == Node::Var) return;
  if (n->kind == Node::Call) {
    static const char* pure[] = {"min", "max", "e", "pi", "log", "round", "int", "ceil", "floor", "abs", "sign",
                                 "sin", "cos", "tan", "asin", "acos", "atan", "sinh", "cosh", "tanh",
                                 "asinh", "acosh", "atanh"};
    bool known = false;
    for (const char* p : pure) if (n->name == p) known = true;
    if (!known) return;
  }
  for (auto& k : n->kids) if (k->kind != Node::Const) return;
  Namespace empty;
  double v = eval(n.get(), empty);
  NodeP c(new Node());
  c->kind = Node::Const;
  c->c = v;
  n = std::move(c);
}

struct Function {  // src/func.rs:5-42
  NodeP root;
  std::string text;
  explicit Function(const std::string& expr) : text(expr) {
    ExprParser p(expr);
    root = p.parse();
    fold(root);
  }
};

// ------------------------------------------------------------------------------------------
// Process universe and DSL parser (src/proc/mod.rs:7-90, src/proc/util.rs:16-166)
// ------------------------------------------------------------------------------------------
struct Incr { int kind; int idx; std::unique_ptr<Function> lambda; };  // kind 0 dt, 1 dW, 2 dN
struct Process {
  std::string name;
  bool levy;
  std::vector<std::unique_ptr<Function>> coefs;
  std::vector<Incr> incs;
};
struct Universe {
  std::vector<Process> procs;
  std::map<std::string, int> registry;      // name -> last index with that name
  std::vector<std::string> stoch_names;     // stochastic registry in index order
  std::vector<int> levy_idx, alg_idx;
};

struct ParseError : std::runtime_error { using std::runtime_error::runtime_error; };

// delimited(char('('), balanced_parens, char(')')) on `in` (which starts with '(').
// Returns false when the parentheses do not balance. `inside` is the enclosed text and
// `rest` the offset just past the closing parenthesis.
static bool delimited_parens(const std::string& in, size_t start, std::string& inside, size_t& rest) {
  if (start >= in.size() || in[start] != '(') return false;
  int depth = 0;
  for (size_t k = start; k < in.size(); ++k) {
    if (in[k] == '(') ++depth;
    else if (in[k] == ')') {
      --depth;
      if (depth == 0) { inside = in.substr(start + 1, k - start - 1); rest = k + 1; return true; }
    }
  }
  return false;
}

static std::string trim(const std::string& s) {
  size_t b = 0, e = s.size();
  while (b < e && std::isspace((unsigned char)s[b])) ++b;
  while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
  return s.substr(b, e - b);
}
static std::string trim_start(const std::string& s) {
  size_t b = 0;
  while (b < s.size() && std::isspace((unsigned char)s[b])) ++b;
  return s.substr(b);
}

static std::unique_ptr<Function> make_function(const std::string& text) {
  try {
    return std::unique_ptr<Function>(new Function(text));
  } catch (const std::runtime_error& e) {
    throw ParseError(std::string("Parse Error: ") + e.what());
  }
}

static Incr build_incrementor(const std::string& inc, std::vector<std::string>& reg) {
  Incr r;
  r.idx = -1;
  if (inc == "dt") { r.kind = 0; return r; }
  int idx = -1;
  for (size_t k = 0; k < reg.size(); ++k) if (reg[k] == inc) idx = (int)k;
  if (idx < 0) { idx = (int)reg.size(); reg.push_back(inc); }
  r.idx = idx;
  if (inc.rfind("dW", 0) == 0) { r.kind = 1; return r; }
  if (inc.rfind("dN", 0) == 0) {
    size_t b = inc.find('(');
    if (b == std::string::npos) throw ParseError("Missing '(' in dN incrementor");
    std::string inside; size_t rest;
    if (!delimited_parens(inc, b, inside, rest)) throw ParseError("Unbalanced parentheses in jump term: " + inc);
    std::string lam = trim(inside);
    try {
      r.lambda = std::unique_ptr<Function>(new Function(lam));
    } catch (const std::runtime_error& e) {
      throw ParseError("Math error in jump lambda '" + lam + "': Parse Error: " + e.what());
    }
    r.kind = 2;
    return r;
  }
  throw ParseError("Unknown incrementor type: " + inc);
}

static Process parse_single(const std::string& eq, std::vector<std::string>& reg) {
  std::vector<std::string> parts;
  size_t b = 0;
  for (size_t k = 0; k <= eq.size(); ++k)
    if (k == eq.size() || eq[k] == '=') { parts.push_back(eq.substr(b, k - b)); b = k + 1; }
  if (parts.size() != 2) throw ParseError("Missing '='");
  std::string lhs = trim(parts[0]), rhs = trim(parts[1]);
  Process P;
  P.name = (!lhs.empty() && lhs[0] == 'd') ? lhs.substr(1) : lhs;
  if (!lhs.empty() && lhs[0] == 'd') {
    P.levy = true;
    std::string cur = rhs;
    for (;;) {
      size_t st = cur.find('(');
      if (st == std::string::npos) break;
      std::string coeff; size_t after;
      if (!delimited_parens(cur, st, coeff, after)) throw ParseError("Unbalanced parentheses in coefficient");
      std::string ta = trim_start(cur.substr(after));
      if (ta.empty() || ta[0] != '*') break;
      std::string as = trim_start(ta.substr(1));
      std::string inc, remaining;
      if (as.rfind("dN", 0) == 0) {
        size_t ds = as.find('(');
        if (ds == std::string::npos) throw ParseError("dN missing opening bracket");
        std::string inside; size_t rest;
        if (!delimited_parens(as, ds, inside, rest)) throw ParseError("Unbalanced parentheses in dN intensity");
        inc = as.substr(0, rest);
        remaining = as.substr(rest);
      } else {
        size_t e = as.find(' ');
        if (e == std::string::npos) e = as.size();
        inc = as.substr(0, e);
        remaining = as.substr(e);
      }
      std::unique_ptr<Function> f;
      try {
        f = std::unique_ptr<Function>(new Function(trim(coeff)));
      } catch (const std::runtime_error& e) {
        throw ParseError(std::string("Math error in coefficient: Parse Error: ") + e.what());
      }
      Incr in = build_incrementor(inc, reg);
      P.coefs.push_back(std::move(f));
      P.incs.push_back(std::move(in));
      cur = remaining;
    }
  } else {
    P.levy = false;
    P.coefs.push_back(make_function(rhs));
  }
  return P;
}

static std::unique_ptr<Universe> parse_equations(const std::vector<std::string>& eqs) {
  std::unique_ptr<Universe> U(new Universe());
  for (auto& e : eqs) U->procs.push_back(parse_single(e, U->stoch_names));
  for (size_t k = 0; k < U->procs.size(); ++k) {
    U->registry[U->procs[k].name] = (int)k;
    (U->procs[k].levy ? U->levy_idx : U->alg_idx).push_back((int)k);
  }
  return U;
}

// ------------------------------------------------------------------------------------------
// Random number generators (src/rng/*.rs)
// ------------------------------------------------------------------------------------------
struct BaseRng { virtual double sample(size_t time_idx, size_t k) = 0; virtual ~BaseRng() {} };

struct PseudoRng : BaseRng {  // src/rng/pseudo.rs:7-60
  ChaCha8Stream rng;
  size_t m;
  bool cached = false;
  size_t cache_t = 0;
  std::vector<double> vals;
  PseudoRng(uint64_t seed, size_t m_) : rng(seed), m(m_) {}
  double sample(size_t t, size_t k) override {
    if (!(cached && cache_t == t)) {
      vals.clear();
      for (size_t i = 0; i < m; ++i) vals.push_back(rng.next_f64());
      cached = true; cache_t = t;
    }
    if (k >= vals.size()) throw EvalError("RNG Index " + std::to_string(k) + " out of bounds (max " + std::to_string(m) + ")");
    return vals[k];
  }
};

struct TableRng : BaseRng {  // Sobol (shift or XOR) and injected uniforms: values[t*m + k]
  size_t m;
  std::vector<double> vals;
  double sample(size_t t, size_t k) override {
    size_t j = t * m + k;
    if (k >= m || j >= vals.size()) throw EvalError("RNG index out of bounds");
    return vals[j];
  }
};

// ------------------------------------------------------------------------------------------
// Filtration and schemes (src/filtration.rs, src/sim/*.rs)
// ------------------------------------------------------------------------------------------
struct Filtration {
  const Universe& U;
  const std::vector<double>& times;
  size_t P;
  std::vector<double> raw;
  std::unordered_map<double, size_t> time_registry;
  double cache_time;
  Namespace cache;
  Filtration(const Universe& u, const std::vector<double>& ts, const std::map<std::string, double>& init)
      : U(u), times(ts), P(u.procs.size()), raw(ts.size() * u.procs.size(), 0.0) {
    for (size_t i = 0; i < ts.size(); ++i) time_registry[ts[i]] = i;
    for (auto& kv : init) {
      auto it = U.registry.find(kv.first);
      if (it != U.registry.end()) set(0, it->second, kv.second);
    }
    refresh(times[0]);
  }
  double get(size_t t, size_t p) const { return raw[t * P + p]; }
  void set(size_t t, size_t p, double v) { raw[t * P + p] = v; }
  void refresh(double time) {
    cache_time = time;
    cache["t"] = time;
    auto it = time_registry.find(time);
    size_t ti = (it == time_registry.end()) ? 0 : it->second;
    for (auto& kv : U.registry) cache[kv.first] = get(ti, kv.second);
  }
  double eval(const Function& f, double t) {  // Function::eval
    if (!(t == cache_time)) refresh(t);
    return orc::eval(f.root.get(), cache);
  }
};

struct Steps {
  std::vector<double> dts, sqrt_dts;
};

static double sample_inc(const Incr& in, size_t t_idx, Filtration& F, BaseRng& rng, const Steps& st) {
  switch (in.kind) {
    case 0: return st.dts[t_idx];
    case 1: { double q = rng.sample(t_idx, in.idx); return st.sqrt_dts[t_idx] * inv_normal(q); }
    default: {
      double u = rng.sample(t_idx, in.idx);
      double lam = F.eval(*in.lambda, F.times[t_idx]) * st.dts[t_idx];
      return (double)inv_poisson(u, lam);
    }
  }
}

// compat 0 = reference; compat 1 = "textbook": the cache is refreshed from the settled row
// t_idx before the first evaluation of the step (algebraic values included).
static void euler_iteration(Filtration& F, const Universe& U, size_t t_idx, BaseRng& rng, const Steps& st, int compat) {
  double tc = F.times[t_idx], tn = F.times[t_idx + 1];
  if (compat == 1) F.refresh(tc);
  for (int p : U.levy_idx) {
    const Process& pr = U.procs[p];
    double val = F.get(t_idx, p);
    for (size_t k = 0; k < pr.incs.size(); ++k) {
      double c = F.eval(*pr.coefs[k], tc);
      double x = sample_inc(pr.incs[k], t_idx, F, rng, st);
      val += c * x;
    }
    F.set(t_idx + 1, p, val);
  }
  for (int p : U.alg_idx) F.set(t_idx + 1, p, F.eval(*U.procs[p].coefs[0], tn));
}

// compat 0 = reference (src/sim/runge_kutta.rs:5-107);
// compat 1 = "textbook": the cache is refreshed from the settled row t_idx before stage 1 and
// from the settled row t_idx+1 before the algebraic pass.
static void rk_iteration(Filtration& F, const Universe& U, size_t t_idx, BaseRng& rng, const Steps& st, int compat) {
  size_t P = U.procs.size();
  double tc = F.times[t_idx], tn = F.times[t_idx + 1];
  double dt = tn - tc, sdt = std::sqrt(dt);
  double sk = rng.sample(t_idx, 0) > 0.5 ? 1.0 : -1.0;
  if (compat == 1) F.refresh(tc);
  std::vector<std::vector<double>> incs(P);
  for (size_t p = 0; p < P; ++p)
    if (U.procs[p].levy)
      for (auto& in : U.procs[p].incs) incs[p].push_back(sample_inc(in, t_idx, F, rng, st));
  std::vector<double> xt(P), k1(P, 0.0), k2(P, 0.0);
  for (size_t p = 0; p < P; ++p) xt[p] = F.get(t_idx, p);
  for (size_t p = 0; p < P; ++p)
    if (U.procs[p].levy)
      for (size_t k = 0; k < incs[p].size(); ++k) k1[p] += F.eval(*U.procs[p].coefs[k], tc) * incs[p][k];
  for (size_t p = 0; p < P; ++p)
    if (U.procs[p].levy) {
      double pert = 0.0;
      for (size_t k = 0; k < U.procs[p].incs.size(); ++k)
        if (U.procs[p].incs[k].kind == 1) pert += F.eval(*U.procs[p].coefs[k], tc) * sk * sdt;
      F.set(t_idx + 1, p, xt[p] + k1[p] + pert);
    }
  for (size_t p = 0; p < P; ++p)
    if (U.procs[p].levy)
      for (size_t k = 0; k < incs[p].size(); ++k) k2[p] += F.eval(*U.procs[p].coefs[k], tn) * incs[p][k];
  for (int p : U.levy_idx) F.set(t_idx + 1, p, xt[p] + 0.5 * (k1[p] + k2[p]));
  if (compat == 1) F.refresh(tn);
  for (int p : U.alg_idx) F.set(t_idx + 1, p, F.eval(*U.procs[p].coefs[0], tn));
}

}  // namespace orc

// ------------------------------------------------------------------------------------------
// C interface for ctypes
// ------------------------------------------------------------------------------------------
using namespace orc;

static void put_err(char* err, int len, const std::string& m) {
  if (err && len > 0) { std::snprintf(err, (size_t)len, "%s", m.c_str()); }
}

extern "C" {

void oracle_chacha_block(const uint32_t in[16], int rounds, uint32_t out[16]) { chacha_block(in, rounds, out); }
void oracle_seed_from_u64(uint64_t seed, uint32_t key[8]) { seed_from_u64(seed, key); }
// n outputs of pcg32_advance_output from `state` with increment `inc` (the seed expansion's step).
void oracle_pcg32(uint64_t state, uint64_t inc, int n, uint32_t* out) {
  for (int i = 0; i < n; ++i) out[i] = pcg32_advance_output(state, inc);
}

// n consecutive f64 draws of ChaCha8Rng::seed_from_u64(seed), starting at draw `first`.
void oracle_pseudo_f64(uint64_t seed, uint64_t first, int64_t n, double* out) {
  ChaCha8Stream s(seed);
  for (uint64_t j = 0; j < first; ++j) s.next_u64();
  for (int64_t j = 0; j < n; ++j) out[j] = s.next_f64();
}

void oracle_sobol_init(int dims, const int64_t* poly, const int64_t* vinit) { sobol_build(dims, poly, vinit); }
int oracle_sobol_dims() { return g_sobol.dims; }
void oracle_sobol_directions(int d, uint32_t out[32]) { for (int k = 0; k < 32; ++k) out[k] = g_sobol.v[(size_t)d * 32 + k]; }
// u32 points [n][dims] for point indices first..first+n-1 (Gray-code order).
void oracle_sobol_points(int dims, uint64_t first, int64_t n, uint32_t* out) {
  for (int64_t i = 0; i < n; ++i)
    for (int d = 0; d < dims; ++d) out[(size_t)i * dims + d] = sobol_point(first + (uint64_t)i, d);
}
// The per-dimension XOR masks of the digital shift drawn from seed: word d of ChaCha8(seed).
void oracle_xor_masks(uint64_t seed, int dims, uint32_t* out) {
  ChaCha8Stream s(seed);
  for (int d = 0; d < dims; ++d) out[d] = s.next_u32();
}

double oracle_inv_normal(double p) { return inv_normal(p); }
uint64_t oracle_inv_poisson(double u, double lambda) { return inv_poisson(u, lambda); }

// Evaluate one expression with a namespace of (name, value) pairs. Returns 0 on success,
// 1 on a parse error, 2 on an evaluation error.
int oracle_eval_expr(const char* expr, const char** names, const double* vals, int n, double* out, char* err, int errlen) {
  std::unique_ptr<Function> f;
  try { f.reset(new Function(expr)); } catch (const std::exception& e) { put_err(err, errlen, e.what()); return 1; }
  Namespace ns;
  for (int k = 0; k < n; ++k) ns[names[k]] = vals[k];
  try { *out = eval(f->root.get(), ns); } catch (const std::exception& e) { put_err(err, errlen, e.what()); return 2; }
  return 0;
}

// Universe summary: P, m, process names ('\n'-joined), algebraic flags. Returns 0 / 1 (parse error).
int oracle_parse(const char** eqs, int n_eq, int* P, int* m, char* names, int names_len, int* is_alg, char* err, int errlen) {
  std::vector<std::string> v(eqs, eqs + n_eq);
  try {
    auto U = parse_equations(v);
    *P = (int)U->procs.size();
    *m = (int)U->stoch_names.size();
    std::string j;
    for (size_t k = 0; k < U->procs.size(); ++k) { if (k) j += "\n"; j += U->procs[k].name; if (is_alg) is_alg[k] = U->procs[k].levy ? 0 : 1; }
    put_err(names, names_len, j);
    return 0;
  } catch (const std::exception& e) { put_err(err, errlen, e.what()); return 1; }
}

// Full simulation. rng: 0 pseudo (ChaCha8, path s seeded s+seed), 1 sobol, 2 injected uniforms
// (u[s][t][k], t < T-1). scramble (sobol only): 0 = the reference's per-path f64 random shift,
// 1 = digital XOR shift with masks from ChaCha8(scramble_seed) and u = (y + 0.5) * 2^-32.
// Sobol point of scenario s is sobol_first + s. compat: 0 reference, 1 textbook (RK only).
// out: [S][T][P] f64. Returns 0 ok, 1 parse/argument error, 2 simulation error.
int oracle_simulate(const char** eqs, int n_eq, const double* times, int T, const char** init_names,
                    const double* init_vals, int n_init, int64_t S, int scheme, int rng_kind, uint64_t seed,
                    int scramble, uint64_t scramble_seed, int64_t sobol_first, const double* injected, int compat,
                    int threads, double* out, char* err, int errlen) {
  std::vector<std::string> v(eqs, eqs + n_eq);
  std::unique_ptr<Universe> U;
  try { U = parse_equations(v); } catch (const std::exception& e) { put_err(err, errlen, e.what()); return 1; }
  std::vector<double> ts(times, times + T);
  std::map<std::string, double> init;
  for (int k = 0; k < n_init; ++k) init[init_names[k]] = init_vals[k];
  size_t P = U->procs.size(), m = U->stoch_names.size();
  size_t N = (size_t)T - 1;
  size_t dims = N * m;
  Steps st;
  for (size_t k = 0; k + 1 < ts.size(); ++k) { st.dts.push_back(ts[k + 1] - ts[k]); st.sqrt_dts.push_back(std::sqrt(ts[k + 1] - ts[k])); }
  if (rng_kind == 1 && dims > 0 && (int)dims > g_sobol.dims) { put_err(err, errlen, "sobol dimension exceeds table"); return 1; }
  std::vector<uint32_t> masks;
  if (rng_kind == 1 && scramble == 1) { masks.resize(dims); oracle_xor_masks(scramble_seed, (int)dims, masks.data()); }
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  std::atomic<int64_t> next(0);
  std::atomic<int> failed(0);
  std::string fail_msg;
  auto worker = [&]() {
    for (;;) {
      int64_t s = next.fetch_add(1);
      if (s >= S || failed.load()) return;
      try {
        Filtration F(*U, ts, init);
        std::unique_ptr<BaseRng> r;
        if (rng_kind == 0) {
          r.reset(new PseudoRng((uint64_t)s + seed, m));
        } else {
          TableRng* tr = new TableRng();
          tr->m = m;
          tr->vals.resize(dims);
          if (rng_kind == 2) {
            for (size_t j = 0; j < dims; ++j) tr->vals[j] = injected[(size_t)s * dims + j];
          } else if (scramble == 0) {
            ChaCha8Stream sh((uint64_t)s + seed);
            uint64_t n = (uint64_t)(sobol_first + s);
            for (size_t d = 0; d < dims; ++d) {
              double x = (double)sobol_point(n, (int)d) * (1.0 / 4294967296.0);
              double y = x + sh.next_f64();
              tr->vals[d] = y - std::trunc(y);
            }
          } else {
            uint64_t n = (uint64_t)(sobol_first + s);
            for (size_t d = 0; d < dims; ++d) {
              uint32_t y = sobol_point(n, (int)d) ^ masks[d];
              tr->vals[d] = ((double)y + 0.5) * (1.0 / 4294967296.0);
            }
          }
          r.reset(tr);
        }
        for (size_t t = 0; t < N; ++t) {
          if (scheme == 0) euler_iteration(F, *U, t, *r, st, compat);
          else rk_iteration(F, *U, t, *r, st, compat);
        }
        std::memcpy(out + (size_t)s * ts.size() * P, F.raw.data(), sizeof(double) * F.raw.size());
      } catch (const std::exception& e) {
        if (failed.fetch_add(1) == 0) fail_msg = e.what();
        return;
      }
    }
  };
  std::vector<std::thread> pool;
  for (int k = 0; k < threads; ++k) pool.emplace_back(worker);
  for (auto& th : pool) th.join();
  if (failed.load()) { put_err(err, errlen, fail_msg); return 2; }
  return 0;
}

}  // extern "C"

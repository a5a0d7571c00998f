// codegen.cpp -- emit the per-path step of a universe as straight-line HIP C++.
//
// The reference evaluates every coefficient through Function::eval (src/func.rs:32-42), which
// refreshes a name -> value cache from the filtration only when the evaluation time differs
// from the cached one (src/filtration.rs:70-79). Which values a coefficient sees therefore
// depends on the order of evaluations inside a step:
//   * Euler (src/sim/euler.rs:5-37): Levy coefficients at t_n, then algebraic processes at
//     t_{n+1}; the algebraic pass refreshes from row n+1 before any algebraic value of that row
//     is set, and the next step's Levy coefficients reuse that cache.
//   * Runge-Kutta (src/sim/runge_kutta.rs:5-107): stage 1 at t_n, a probe row written into
//     n+1, stage 2 at t_{n+1} (which refreshes from the probe row), then the settled row; the
//     next step's stage 1 and the algebraic pass reuse the probe cache.
// The emitter replays that control flow on a symbolic cache (which row and which generation it
// holds) and writes an explicit register copy wherever the reference would refresh. Because
// the sequence of evaluations is the same every step, the symbolic run is exact. Values that
// the reference computes twice from the same cache and time (e.g. RK stage-2 coefficients at
// t_{n+1} and the next step's stage-1 coefficients at the same time on the same probe cache)
// are carried in registers across steps instead of being recomputed.
#include <map>
#include <set>
#include <sstream>

#include "codegen.hpp"

namespace sdesim {

namespace {

enum class At { N, N1, Other };

struct Carried {
  std::string key;   // canonical text of the expression
  const Expr* expr;
};

class StepEmitter {
 public:
  StepEmitter(const Universe& u, const CodegenOptions& o) : u_(u), o_(o) {
    R_ = o.f32 ? "float" : "double";
  }

  // Emit one step given the cache state at entry. `carried` lists values available in K[i]
  // when entry == At::N. On return, end_at() / end_values() describe the cache at exit.
  std::string step(At entry, const std::vector<Carried>& carried) {
    out_.str("");
    memo_.clear();
    psi_.clear();
    at_ = entry;
    gen_ = 0;
    tmp_ = 0;
    if (entry == At::N) {
      for (size_t i = 0; i < carried.size(); ++i) memo_[key(carried[i].key, false)] = "K[" + std::to_string(i) + "]";
    }
    const int P = u_.P();
    for (int p = 0; p < P; ++p) line(R_ + " y" + std::to_string(p) + " = " + R_ + "(0);");
    if (o_.scheme == 0) euler(); else rk();
    finish_row();
    end_at_ = at_;
    end_gen_ = gen_;
    end_values_.clear();
    if (at_ == At::N1) {
      for (auto& kv : memo_) {
        if (kv.first.gen == gen_ && kv.first.n1) end_values_.push_back(kv.first.canon);
      }
    }
    return out_.str();
  }

  At end_at() const { return end_at_; }
  const std::vector<std::string>& end_values() const { return end_values_; }

  // Expression text of a carried value at the end of the step just emitted.
  std::string end_value(const std::string& canon) const {
    MemoKey k{canon, true, end_gen_};
    auto it = memo_.find(k);
    return it == memo_.end() ? std::string() : it->second;
  }

  // Emit the prologue: the cache equals the initial row at t0, and each carried value is its
  // expression evaluated there.
  std::string prologue(const std::vector<Carried>& carried) {
    std::ostringstream o;
    for (size_t i = 0; i < carried.size(); ++i)
      o << "    K[" << i << "] = " << emit_expr(*carried[i].expr, "tc", "C") << ";\n";
    return o.str();
  }

  std::string emit_expr(const Expr& e, const std::string& tvar, const std::string& arr) {
    EmitEnv env;
    env.f32 = o_.f32;
    env.resolve = [&](const std::string& name) -> std::string {
      auto it = u_.registry.find(name);
      if (it != u_.registry.end()) return arr + "[" + std::to_string(it->second) + "]";
      if (name == "t") return tvar;
      if (o_.steps) throw UndefinedName("undefined variable '" + name + "'");
      return R_ + "(0)";
    };
    return emit(e, env);
  }

 private:
  struct MemoKey {
    std::string canon;
    bool n1;
    int gen;
    bool operator<(const MemoKey& o) const {
      if (canon != o.canon) return canon < o.canon;
      if (n1 != o.n1) return n1 < o.n1;
      return gen < o.gen;
    }
  };

  const Universe& u_;
  const CodegenOptions& o_;
  std::string R_;
  std::ostringstream out_;
  std::map<MemoKey, std::string> memo_;
  std::set<int> psi_;
  At at_ = At::Other;
  int gen_ = 0, tmp_ = 0;
  At end_at_ = At::Other;
  int end_gen_ = 0;
  std::vector<std::string> end_values_;

  MemoKey key(const std::string& canon, bool n1) const { return MemoKey{canon, n1, gen_}; }
  void line(const std::string& s) { out_ << "    " << s << "\n"; }
  std::string y(int p) const { return "y" + std::to_string(p); }

  void refresh(bool n1) {
    const int P = u_.P();
    for (int p = 0; p < P; ++p) line("C[" + std::to_string(p) + "] = " + (n1 ? y(p) : "X[" + std::to_string(p) + "]") + ";");
    at_ = n1 ? At::N1 : At::N;
    ++gen_;
  }

  // Function::eval at t_n (n1 = false) or t_{n+1} (n1 = true).
  std::string eval(const Expr& e, bool n1) {
    At want = n1 ? At::N1 : At::N;
    if (at_ != want) refresh(n1);
    std::string canon = canonical(e);
    MemoKey k = key(canon, n1);
    auto it = memo_.find(k);
    if (it != memo_.end()) return it->second;
    std::string name = "v" + std::to_string(tmp_++);
    line("const " + R_ + " " + name + " = " + emit_expr(e, n1 ? "tn" : "tc", "C") + ";");
    memo_[k] = name;
    return name;
  }

  // The Wiener increment sqrt(dt) * Psi(u_idx), evaluated once per step.
  std::string dw(int idx) {
    std::string w = "w" + std::to_string(idx);
    if (!psi_.count(idx)) {
      line("const " + R_ + " " + w + " = sde_dw(u[" + std::to_string(idx) + "], sdt);");
      psi_.insert(idx);
    }
    return w;
  }

  // Incrementor::sample (src/proc/increment.rs:46-157).
  std::string sample(const Term& t) {
    switch (t.kind) {
      case IncKind::Time: return "dt";
      case IncKind::Wiener: return dw(t.index);
      case IncKind::Jump: {
        std::string lam = eval(*t.lambda, false);
        std::string name = "j" + std::to_string(tmp_++);
        line("const " + R_ + " " + name + " = (" + R_ + ")sde_inv_poisson(sde_uval(u[" + std::to_string(t.index) +
             "]), (double)(" + lam + " * dt));");
        return name;
      }
    }
    return "dt";
  }

  void finish_row() {
    const int P = u_.P();
    for (int p = 0; p < P; ++p) line("X[" + std::to_string(p) + "] = " + y(p) + ";");
  }

  void algebraic_pass() {
    for (int a : u_.algebraic) {
      std::string v = eval(*u_.processes[a].expr, true);
      line(y(a) + " = " + v + ";");
    }
  }

  void euler() {
    if (o_.compat == 1) refresh(false);
    for (int p : u_.levy) {
      const ProcessDef& pd = u_.processes[p];
      std::string acc = "a" + std::to_string(p);
      if (o_.f32) {
        // f32: the terms are summed first and the sum is added to the state strictly.
        bool first = true;
        for (const Term& t : pd.terms) {
          std::string c = eval(*t.coef, false);
          std::string x = sample(t);
          if (first) line(R_ + " " + acc + " = " + c + " * " + x + ";");
          else line(acc + " = " + acc + " + " + c + " * " + x + ";");
          first = false;
        }
        if (first) line(y(p) + " = X[" + std::to_string(p) + "];");
        else line(y(p) + " = sde_add_strict(X[" + std::to_string(p) + "], " + acc + ");");
        continue;
      }
      line(R_ + " " + acc + " = X[" + std::to_string(p) + "];");
      for (const Term& t : pd.terms) {
        std::string c = eval(*t.coef, false);
        std::string x = sample(t);
        line(acc + " = " + acc + " + " + c + " * " + x + ";");
      }
      line(y(p) + " = " + acc + ";");
    }
    algebraic_pass();
  }

  void rk() {
    if (u_.M() == 0) throw Panic("RNG Index 0 out of bounds (max 0)");
    if (o_.compat == 1) refresh(false);
    const int P = u_.P();
    std::vector<std::vector<std::string>> d(P);
    for (int p = 0; p < P; ++p) {
      const ProcessDef& pd = u_.processes[p];
      if (!pd.levy) continue;
      for (const Term& t : pd.terms) d[p].push_back(sample(t));
    }
    // Sums start at their first term rather than at 0 + first term (the reference's
    // `k1 = 0.0; k1 += c * d`): the values are identical except that an exact -0.0 term stays
    // -0.0, and the device can fuse every product into an FMA. (c * sk) * sdt is evaluated as
    // c * (sk * sdt), exactly equal because sk = +-1 (sk = 1 when u_0 > 1/2, else -1).
    line("const " + R_ + " sks = sde_signed(u[0], sdt);");
    for (int p = 0; p < P; ++p) {
      const ProcessDef& pd = u_.processes[p];
      if (!pd.levy) continue;
      std::string k1 = "k1_" + std::to_string(p);
      line(R_ + " " + k1 + " = " + R_ + "(0);");
      for (size_t k = 0; k < pd.terms.size(); ++k) {
        std::string c = eval(*pd.terms[k].coef, false);
        if (k == 0) line(k1 + " = " + c + " * " + d[p][k] + ";");
        else line(k1 + " = " + k1 + " + " + c + " * " + d[p][k] + ";");
      }
    }
    for (int p = 0; p < P; ++p) {
      const ProcessDef& pd = u_.processes[p];
      if (!pd.levy) continue;
      std::string pe = "pe" + std::to_string(p);
      line(R_ + " " + pe + " = " + R_ + "(0);");
      bool first = true;
      for (const Term& t : pd.terms) {
        if (t.kind != IncKind::Wiener) continue;
        std::string c = eval(*t.coef, false);
        if (first) line(pe + " = " + c + " * sks;");
        else line(pe + " = " + pe + " + " + c + " * sks;");
        first = false;
      }
      line(y(p) + " = X[" + std::to_string(p) + "] + k1_" + std::to_string(p) + " + " + pe + ";");
    }
    for (int p = 0; p < P; ++p) {
      const ProcessDef& pd = u_.processes[p];
      if (!pd.levy) continue;
      std::string k2 = "k2_" + std::to_string(p);
      line(R_ + " " + k2 + " = " + R_ + "(0);");
      for (size_t k = 0; k < pd.terms.size(); ++k) {
        std::string c = eval(*pd.terms[k].coef, true);
        if (k == 0) line(k2 + " = " + c + " * " + d[p][k] + ";");
        else line(k2 + " = " + k2 + " + " + c + " * " + d[p][k] + ";");
      }
    }
    for (int p : u_.levy) {
      const std::string inc = R_ + "(0.5) * (k1_" + std::to_string(p) + " + k2_" + std::to_string(p) + ")";
      if (o_.f32) line(y(p) + " = sde_add_strict(X[" + std::to_string(p) + "], " + inc + ");");
      else line(y(p) + " = X[" + std::to_string(p) + "] + " + inc + ";");
    }
    if (o_.compat == 1) refresh(true);
    algebraic_pass();
  }
};

}  // namespace

std::string generate_kernel_source(const Universe& u, const CodegenOptions& o) {
  const std::string R = o.f32 ? "float" : "double";
  const int P = u.P(), M = u.M();
  StepEmitter em(u, o);

  // Pass 1: find the cache state a step leaves behind and the values it holds.
  em.step(At::Other, {});
  At entry = At::Other;
  std::vector<Carried> carried;
  std::vector<std::unique_ptr<Expr>> keep;
  if (em.end_at() == At::N1) {
    entry = At::N;
    // Re-parse carried expressions from their canonical form is not possible; collect them
    // from the universe instead (every evaluated expression is a coefficient, a lambda or an
    // algebraic expression).
    std::map<std::string, const Expr*> all;
    for (const ProcessDef& pd : u.processes) {
      if (pd.expr) all[canonical(*pd.expr)] = pd.expr.get();
      for (const Term& t : pd.terms) {
        all[canonical(*t.coef)] = t.coef.get();
        if (t.lambda) all[canonical(*t.lambda)] = t.lambda.get();
      }
    }
    for (const std::string& c : em.end_values()) {
      auto it = all.find(c);
      if (it != all.end()) carried.push_back(Carried{c, it->second});
    }
  }
  // Pass 2: the steady-state step with carried values.
  std::string body = em.step(entry, carried);
  std::ostringstream tail;
  for (size_t i = 0; i < carried.size(); ++i) {
    std::string v = em.end_value(carried[i].key);
    if (!v.empty()) tail << "    K[" << i << "] = " << v << ";\n";
  }
  std::string pro = (entry == At::N) ? em.prologue(carried) : std::string();
  const int NK = carried.empty() ? 1 : (int)carried.size();

  std::ostringstream s;
  s << "#include \"sdesim_device.hpp\"\n\n";
  s << "struct SdeModel {\n";
  s << "  typedef " << R << " real;\n";
  s << "  static constexpr int P = " << P << ";\n";
  s << "  static constexpr int M = " << M << ";\n";
  s << "  static constexpr int NK = " << NK << ";\n";
  s << "  static constexpr bool HAS_PAYOFF = " << (o.payoff.empty() ? "false" : "true") << ";\n";
  s << "  __device__ static __forceinline__ void prologue(real (&X)[P], real (&C)[P], real (&K)[NK], real tc) {\n";
  s << pro;
  s << "  }\n";
  s << "  template <class UT>\n";
  s << "  __device__ static __forceinline__ void step(real (&X)[P], real (&C)[P], real (&K)[NK], real tc, real tn,\n";
  s << "                                              real dt, real sdt, const UT (&u)[" << (M > 0 ? M : 1) << "]) {\n";
  // FP32 kernels let the compiler reassociate the step's arithmetic (for example 2 * (a - v)
  // into one FMA); FP64 kernels keep the reference's evaluation order.
  if (o.f32) s << "#pragma clang fp reassociate(on)\n";
  s << body << tail.str();
  s << "  }\n";
  s << "  __device__ static __forceinline__ real payoff(const real (&X)[P], real tt) {\n";
  if (!o.payoff.empty()) {
    ExprPtr pe = parse_expression(o.payoff);
    EmitEnv env;
    env.f32 = o.f32;
    env.resolve = [&](const std::string& name) -> std::string {
      auto it = u.registry.find(name);
      if (it != u.registry.end()) return "X[" + std::to_string(it->second) + "]";
      if (name == "t") return "tt";
      throw UndefinedName("undefined variable '" + name + "' in payoff");
    };
    s << "    return " << emit(*pe, env) << ";\n";
  } else {
    s << "    return real(0);\n";
  }
  s << "  }\n";
  s << "};\n\n";
  s << "extern \"C\" __global__ void __launch_bounds__(" << o.block << ") " << o.name << "(SimArgs a) {\n";
  s << "  sde_path_kernel<SdeModel, " << o.rng << ", " << o.out << ", " << o.ppl << ", " << o.nsl << ", "
    << (o.uniform_grid ? "true" : "false") << ", " << (o.wave_aligned ? "true" : "false") << ">(a);\n";
  s << "}\n";
  return s.str();
}

std::string kernel_name(const CodegenOptions& o) {
  static const char* rng[] = {"pseudo", "sobolshift", "sobolxor", "injected", "none"};
  std::string out = o.out == 1 ? "filt" : o.out == 2 ? "term" : "filt_term";
  return std::string("sdesim_") + (o.scheme == 0 ? "em" : "rk") + (o.f32 ? "_f32_" : "_f64_") +
         rng[o.rng >= 0 && o.rng <= 4 ? o.rng : 4] + "_" + out;
}

}  // namespace sdesim

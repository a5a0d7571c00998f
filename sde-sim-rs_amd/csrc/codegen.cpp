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
//
// FP32 kernels also move constant factors from coefficients to increments, where they cost
// less (the values are those of the same expressions at the same points; only the association
// of the products changes):
//   * a dt term c * r(t, C) becomes r * (c * dt), and c * dt is loop-invariant on a uniform grid;
//   * Wiener terms of one process whose coefficients are constant multiples c_k * r of one
//     residual r become r * g with g = sum_k c_k * dW_k formed once per step, and their RK probe
//     terms r * ((sum_k c_k) * sk * sqrt(dt));
//   * an increment whose every use is such a group member with one constant c > 0 is formed as
//     sqrt(dt) * c * Psi(u) directly (the multiplier takes its sign from u, so c must be > 0).
#include <cstdio>
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
    if (o.f32) plan_terms();
  }

  // Residual expressions made by the FP32 term plans, by canonical text (values carried across
  // steps may be residuals).
  void add_residuals(std::map<std::string, const Expr*>& all) const {
    for (const auto& terms : plan_)
      for (const TermPlan& tp : terms)
        if (tp.r) all[canonical(*tp.r)] = tp.r;
  }

  // Emit one step given the cache state at entry. `carried` lists values available in K[i]
  // when entry == At::N. On return, end_at() / end_values() describe the cache at exit.
  std::string step(At entry, const std::vector<Carried>& carried) {
    out_.str("");
    memo_.clear();
    psi_.clear();
    gvar_.clear();
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

  // FP32 term plans, per process and term: kind 0 = the coefficient as written, 1 = a dt term
  // r * (c * dt), 2 = a member of Wiener group `group` (r * g).
  struct TermPlan {
    int kind = 0;
    double c = 1.0;
    const Expr* r = nullptr;
    int group = -1;
  };
  struct Group {
    const Expr* r;
    std::vector<std::pair<double, int>> members;  // (c_k, increment index)
    double csum = 0.0;
  };
  struct Split {
    double c;
    const Expr* r;  // nullptr: the expression is the constant c
  };
  std::vector<std::vector<TermPlan>> plan_;
  std::vector<Group> groups_;
  std::vector<ExprPtr> store_;
  std::map<int, std::string> gvar_;
  std::map<int, double> prescale_;  // increment index -> the constant folded into its multiplier

  static ExprPtr clone(const Expr& e) {
    ExprPtr c = std::make_unique<Expr>();
    c->op = e.op;
    c->fn = e.fn;
    c->value = e.value;
    c->name = e.name;
    for (const ExprPtr& a : e.args) c->args.push_back(clone(*a));
    return c;
  }

  // e = c * r with c constant: peels constant factors off products and negations.
  Split split(const Expr& e) {
    if (e.op == Op::Const) return {e.value, nullptr};
    if (e.op == Op::Neg) {
      Split a = split(*e.args[0]);
      return {-a.c, a.r};
    }
    if (e.op == Op::Mul) {
      Split a = split(*e.args[0]), b = split(*e.args[1]);
      if (!a.r) return {a.c * b.c, b.r};
      if (!b.r) return {a.c * b.c, a.r};
      if (a.c == 1.0 && b.c == 1.0) return {1.0, &e};
      ExprPtr m = std::make_unique<Expr>();
      m->op = Op::Mul;
      m->args.push_back(clone(*a.r));
      m->args.push_back(clone(*b.r));
      store_.push_back(std::move(m));
      return {a.c * b.c, store_.back().get()};
    }
    return {1.0, &e};
  }

  void plan_terms() {
    plan_.assign(u_.P(), {});
    for (int p = 0; p < u_.P(); ++p) {
      const ProcessDef& pd = u_.processes[p];
      if (!pd.levy) continue;
      std::vector<Split> sp;
      std::map<std::string, std::vector<size_t>> by_r;
      for (size_t k = 0; k < pd.terms.size(); ++k) {
        const Term& t = pd.terms[k];
        sp.push_back(split(*t.coef));
        plan_[p].push_back(TermPlan{});
        if (!sp[k].r) continue;
        if (t.kind == IncKind::Time && sp[k].c != 1.0) plan_[p][k] = TermPlan{1, sp[k].c, sp[k].r, -1};
        if (t.kind == IncKind::Wiener) by_r[canonical(*sp[k].r)].push_back(k);
      }
      for (auto& kv : by_r) {
        if (kv.second.size() < 2) continue;
        Group g{sp[kv.second[0]].r, {}, 0.0};
        for (size_t k : kv.second) {
          g.members.push_back({sp[k].c, pd.terms[k].index});
          g.csum += sp[k].c;
          plan_[p][k] = TermPlan{2, sp[k].c, g.r, (int)groups_.size()};
        }
        groups_.push_back(g);
      }
    }
    // Increments used only as group members with one constant c > 0.
    std::map<int, std::set<double>> consts;
    std::set<int> plain;
    for (int p = 0; p < u_.P(); ++p) {
      const ProcessDef& pd = u_.processes[p];
      if (!pd.levy) continue;
      for (size_t k = 0; k < pd.terms.size(); ++k) {
        if (pd.terms[k].kind != IncKind::Wiener) continue;
        if (plan_[p][k].kind == 2) consts[pd.terms[k].index].insert(plan_[p][k].c);
        else plain.insert(pd.terms[k].index);
      }
    }
    for (auto& kv : consts)
      if (!plain.count(kv.first) && kv.second.size() == 1 && *kv.second.begin() > 0.0)
        prescale_[kv.first] = *kv.second.begin();
  }

  static std::string flit(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "((float)%a)", v);
    return b;
  }

  // g = sum_k c_k * dW_k of a Wiener group, once per step.
  std::string group_var(int gid) {
    auto it = gvar_.find(gid);
    if (it != gvar_.end()) return it->second;
    const Group& g = groups_[gid];
    std::string expr;
    for (const auto& m : g.members)
      expr += (expr.empty() ? "" : " + ") + (prescale_.count(m.second) ? dw(m.second) : flit(m.first) + " * " + dw(m.second));
    std::string name = "g" + std::to_string(gid);
    line("const " + R_ + " " + name + " = " + expr + ";");
    gvar_[gid] = name;
    return name;
  }

  // Emits sum = (the products of process p's terms at t_n (n1 = false) or t_{n+1}, increments
  // d[k]), each line right after the values it reads; `first` is the text that starts the sum
  // ("R sum = " or "sum = "). Returns the number of products.
  int sum_terms(int p, bool n1, const std::vector<std::string>& d, const std::string& sum, const std::string& first) {
    const ProcessDef& pd = u_.processes[p];
    std::set<int> done;
    int n = 0;
    for (size_t k = 0; k < pd.terms.size(); ++k) {
      const TermPlan tp = plan_.empty() ? TermPlan{} : plan_[p][k];
      std::string term;
      if (tp.kind == 0) {
        term = eval(*pd.terms[k].coef, n1) + " * " + d[k];
      } else if (tp.kind == 1) {
        term = eval(*tp.r, n1) + " * (" + flit(tp.c) + " * dt)";
      } else {
        if (done.count(tp.group)) continue;
        done.insert(tp.group);
        std::string r = eval(*tp.r, n1);
        term = r + " * " + group_var(tp.group);
      }
      if (n == 0) line(first + term + ";");
      else line(sum + " = " + sum + " + " + term + ";");
      ++n;
    }
    return n;
  }

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
      auto pc = prescale_.find(idx);
      const std::string m = pc == prescale_.end() ? std::string("sdt") : "(" + flit(pc->second) + " * sdt)";
      line("const " + R_ + " " + w + " = sde_dw(u[" + std::to_string(idx) + "], " + m + ");");
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
        std::vector<std::string> d;
        for (const Term& t : pd.terms) d.push_back(sample(t));
        if (sum_terms(p, false, d, acc, R_ + " " + acc + " = ") == 0) line(y(p) + " = X[" + std::to_string(p) + "];");
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
      sum_terms(p, false, d[p], k1, k1 + " = ");
    }
    for (int p = 0; p < P; ++p) {
      const ProcessDef& pd = u_.processes[p];
      if (!pd.levy) continue;
      std::string pe = "pe" + std::to_string(p);
      line(R_ + " " + pe + " = " + R_ + "(0);");
      bool first = true;
      std::set<int> done;
      for (size_t k = 0; k < pd.terms.size(); ++k) {
        const Term& t = pd.terms[k];
        if (t.kind != IncKind::Wiener) continue;
        const TermPlan tp = plan_.empty() ? TermPlan{} : plan_[p][k];
        std::string term;
        if (tp.kind == 2) {
          if (done.count(tp.group)) continue;
          done.insert(tp.group);
          term = eval(*tp.r, false) + " * (" + flit(groups_[tp.group].csum) + " * sks)";
        } else {
          term = eval(*t.coef, false) + " * sks";
        }
        if (first) line(pe + " = " + term + ";");
        else line(pe + " = " + pe + " + " + term + ";");
        first = false;
      }
      line(y(p) + " = X[" + std::to_string(p) + "] + k1_" + std::to_string(p) + " + " + pe + ";");
    }
    for (int p = 0; p < P; ++p) {
      const ProcessDef& pd = u_.processes[p];
      if (!pd.levy) continue;
      std::string k2 = "k2_" + std::to_string(p);
      line(R_ + " " + k2 + " = " + R_ + "(0);");
      sum_terms(p, true, d[p], k2, k2 + " = ");
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
    em.add_residuals(all);
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

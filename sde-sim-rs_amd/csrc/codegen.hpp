// codegen.hpp -- HIP source for one universe + options (compiled at run time with hipRTC).
#pragma once
#include <string>
#include <vector>

#include "universe.hpp"

namespace sdesim {

struct CodegenOptions {
  int scheme = 0;       // 0 Euler-Maruyama, 1 "Runge-Kutta first order"
  int compat = 0;       // 0 reference cache semantics, 1 textbook (refresh from settled rows)
  bool f32 = false;     // f32 state and coefficients (f64 uniforms and estimator sums)
  int rng = 0;          // SDE_RNG_* in sdesim_device.hpp
  int out = 1;          // SDE_OUT_* bit mask
  int ppl = 1;          // paths per lane
  int block = 256;      // threads per block
  int nsl = 1;          // Sobol table slices (Gray-code bits covered: 6 + 8 (nsl - 1))
  bool uniform_grid = false;  // every dt and sqrt(dt) is bitwise equal in the kernel's precision
  bool wave_aligned = false;  // Sobol index of path 0 is a multiple of 64 and blocks hold whole waves
  bool steps = true;    // false when the time grid has a single point (nothing is evaluated)
  std::string payoff;   // expression over process names and t at the last time point
  std::string name = "sdesim_kernel";  // entry point name
};

struct UndefinedName : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// A condition under which the reference panics (surfaced by PyO3 as PanicException).
struct Panic : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Returns the complete kernel source (entry point `o.name`). Throws ParseError for a bad
// payoff, UndefinedName when an evaluated expression names an unknown variable, and
// Panic for the reference's run-time panics (RK with no stochastic increment).
std::string generate_kernel_source(const Universe& u, const CodegenOptions& o);

// Descriptive entry point name: sdesim_<em|rk>_<f32|f64>_<rng>_<filt|term|filt_term>.
std::string kernel_name(const CodegenOptions& o);

}  // namespace sdesim

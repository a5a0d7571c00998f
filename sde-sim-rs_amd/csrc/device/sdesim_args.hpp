// sdesim_args.hpp -- the argument block of the fused path kernel and its mode constants.
// Plain C++ (no HIP): included by the host runtime and by the device code (ahead-of-time and
// hipRTC builds), so both sides agree on the layout.
#pragma once

typedef unsigned int sde_u32;
typedef unsigned long long sde_u64;

// ------------------------------------------------------------------------------------------
// Kernel argument block (identical layout on host and device; 8-byte fields first)
// ------------------------------------------------------------------------------------------
struct SimArgs {
  const double* times;       // [T]
  const double* dts;         // [N]  t[n+1] - t[n]
  const double* sdts;        // [N]  sqrt(dt)
  const float* times_f;      // [T]  f32 copies for f32 kernels
  const float* dts_f;        // [N]
  const float* sdts_f;       // [N]
  const double* x0;          // [P]
  const sde_u32* sobol_tab;  // [dims][64 + 256 * (nslices - 1)] sliced XOR tables; in XOR mode
                             // slice 0 of dimension d also holds the complement of its mask;
                             // wave-aligned XOR mode: slice 0 only, as [N][64][M]
  const sde_u32* sobol_up;   // wave-aligned XOR mode: slices 1.. as [nslices - 1][256][N * M]
  const float* tail_tab;     // [1233] |z| / sqrt(2 ln 2) of XOR uniforms with qi < 1233 (|z| > 5)
  const double* injected;    // [N * M][S] injected uniforms (test hook)
  void* filtration;          // [T][P][S] real, or null
  double* partials;          // [gridDim.x][2 * KQ] (sum, sum of squares) of shifted values, or null
  double dt_u, sdt_u;        // the step when every dt (and sqrt dt) is bitwise equal (f64 kernels)
  float dtf_u, sdtf_u;       // the same for f32 kernels
  long long S;               // paths
  sde_u64 seed;              // path s uses seed + s (wrapping)
  long long sobol_first;     // Sobol point index of path 0
  int N;                     // steps
  int T;                     // time points
  int nslices;               // Sobol table slices
  int sobol_tab_words;       // entries in sobol_tab (bounds of its buffer resource)
  int sobol_dims;            // wave-aligned XOR mode: N * M, the row length of sobol_up
};

#define SDE_RNG_PSEUDO 0
#define SDE_RNG_SOBOL_SHIFT 1
#define SDE_RNG_SOBOL_XOR 2
#define SDE_RNG_INJECTED 3
#define SDE_RNG_NONE 4

#define SDE_OUT_FILTRATION 1
#define SDE_OUT_TERMINAL 2

// ------------------------------------------------------------------------------------------
// Batched 1-D instances (batch_kernels.hip)
// ------------------------------------------------------------------------------------------
// One SDE instance of a batch (built-in model, run-time parameters, uniform grid).
struct SdeInstance {
  double p0, p1, p2;   // GBM: mu, sigma; OU / CIR: kappa, theta, sigma
  double x0;
  double dt, sdt;      // f64 grid
  float dtf, sdtf;     // f32 grid
  sde_u64 seed;        // path s uses seed + s
  long long paths;
  int kind;            // 0 GBM, 1 OU, 2 CIR
  int steps;
};

// One work item: up to 64 * PPL paths of one instance.
struct SdeItem {
  int instance;
  int chunk;           // paths [chunk * 64 * PPL, ...)
  int slot;            // partial slot (instance-major order)
  int pad;
};


// ------------------------------------------------------------------------------------------
// Correlated GBM basket, price and adjoint (basket_kernels.hip)
// ------------------------------------------------------------------------------------------
struct BasketArgs {
  double x0[8], mu[8], sigma[8], weight[8];
  double chol[64];            // lower Cholesky factor of the correlation, row-major 8 x 8
  double dt, sdt, strike, width;
  sde_u64 seed;               // path s uses seed + s
  long long paths;
  void* states;               // adjoint storage: the tape [steps][D][paths] real, or checkpoints
                              // [slots][D][paths] real (one slot per 19 steps)
  double* partials;           // [gridDim.x][NQ]
  int steps;
  int assets;                 // the basket's asset count A: step n uses draws n * A .. n * A + A - 1
};

// ------------------------------------------------------------------------------------------
// Forward-rate model, price and curve sensitivities (lmm_kernels.hip); 40 forwards, 3 factors
// ------------------------------------------------------------------------------------------
struct LmmArgs {
  double L0[40];                      // initial forwards
  const double* pard;                 // device [40] volatilities then [40][3] loadings (f64)
  const float* parf;                  // the same in f32
  double dt, sdt, tau, strike;
  long long paths;
  const sde_u32* sobol_tab;           // slice 0 as [steps][64][3], the mask complements folded in
  const sde_u32* sobol_up;            // slices 1-2 as [2][256][steps * 3]
  void* checkpoints;                  // [slots][40][paths] real, one slot per 16 steps
  void* tape;                         // [16][40][paths] real, the states of one segment
  double* partials;                   // [gridDim.x][42]
  int steps;
  int steps_per_period;
  int expiry_index;
  int sobol_tab_words;
  int sobol_dims;                     // steps * 3, the row length of sobol_up
  int zero;                           // always 0: makes per-step parameter reads step-dependent
};

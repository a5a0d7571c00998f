/* sdesim.h -- C ABI of the MI355X SDE path simulator (libsdesim.so).
 *
 * The reference exposes two surfaces: the PyO3 function sde_sim_rs.simulate
 * (src/py_binding.rs:8-56) and the Rust crate API sim::simulate (src/sim/mod.rs:20-92) with
 * proc::util::parse_equations (src/proc/util.rs:52-66). Per-step trait calls cannot cross to a
 * GPU, so this ABI is coarse: parse once, plan once (compile + upload constants), run many times.
 *
 * Conventions: plain C types and pointers only; every function that can fail returns an int
 * status (SDESIM_OK, SDESIM_EVALUE for what the reference reports as ValueError / a parse error,
 * SDESIM_ERUNTIME for a RuntimeError or a device failure, SDESIM_EPANIC where the reference
 * panics: unknown scheme, undefined variable, RK without a stochastic increment, Sobol
 * dimension or length beyond the table, empty time grid) and leaves a message in a
 * thread-local buffer read by sdesim_last_error(). No C++ exception crosses the ABI. Handles are
 * owned by the library and freed by the caller with the matching free function. Distinct
 * handles may be used from different threads concurrently; the run-time compiled kernel cache
 * is guarded by a mutex. All device work of a call is issued on the given HIP stream (NULL =
 * the null stream); functions that return host data synchronise that stream.
 */
#ifndef SDESIM_H
#define SDESIM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDESIM_OK 0
#define SDESIM_EVALUE 1
#define SDESIM_ERUNTIME 2
#define SDESIM_EPANIC 3 /* a condition under which the reference panics (PyO3 PanicException) */

#define SDESIM_EULER 0
#define SDESIM_RUNGE_KUTTA 1

#define SDESIM_PSEUDO 0   /* ChaCha8Rng::seed_from_u64(seed + s) per path (src/rng/pseudo.rs) */
#define SDESIM_SOBOL 1    /* Sobol points; scramble selects the randomisation (src/rng/sobol.rs) */
#define SDESIM_INJECTED 2 /* caller-provided uniforms [N*m][S] (test hook, no reference counterpart) */

#define SDESIM_SCRAMBLE_SHIFT 0 /* the reference's per-path f64 random shift (src/rng/sobol.rs:62-79) */
#define SDESIM_SCRAMBLE_XOR 1   /* digital XOR shift shared by all paths of a replicate */

#define SDESIM_F64 0
#define SDESIM_F32 1

#define SDESIM_COMPAT_REFERENCE 0 /* the reference's name-cache semantics (stale RK / algebraic values) */
#define SDESIM_COMPAT_TEXTBOOK 1  /* coefficients see the settled rows */

#define SDESIM_OUT_FILTRATION 1
#define SDESIM_OUT_TERMINAL 2

typedef struct sdesim_universe sdesim_universe;
typedef struct sdesim_plan sdesim_plan;

typedef struct sdesim_options {
  int32_t scheme;         /* SDESIM_EULER | SDESIM_RUNGE_KUTTA */
  int32_t rng;            /* SDESIM_PSEUDO | SDESIM_SOBOL | SDESIM_INJECTED */
  int32_t scramble;       /* SDESIM_SCRAMBLE_SHIFT | SDESIM_SCRAMBLE_XOR (Sobol only) */
  int32_t precision;      /* SDESIM_F64 | SDESIM_F32 (state, coefficients and filtration values) */
  int32_t compat;         /* SDESIM_COMPAT_REFERENCE | SDESIM_COMPAT_TEXTBOOK */
  int32_t outputs;        /* bit mask of SDESIM_OUT_* */
  int32_t paths_per_lane; /* 0 = default */
  int32_t block_size;     /* 0 = default (256) */
  uint64_t seed;          /* path s uses seed + s (src/sim/mod.rs:56,65) */
  uint64_t scramble_seed; /* XOR masks = words of ChaCha8Rng::seed_from_u64(scramble_seed) */
  int64_t sobol_first;    /* point index of path 0; < 0 = default (5 for shift, 0 for XOR). In XOR
                             mode a multiple of 64 lets each wave read its shared table words with
                             scalar loads (the fast form); any other value is also exact */
} sdesim_options;

/* Fill `o` with the defaults: Euler, pseudo, f64, reference compat, filtration output. */
void sdesim_default_options(sdesim_options* o);

/* Replaces proc::util::parse_equations (src/proc/util.rs:52-66) and ProcessUniverse::new
 * (src/proc/mod.rs:70-89). Parse errors return SDESIM_EVALUE with the reference's message. */
int sdesim_parse_equations(const char* const* equations, int32_t n_equations, sdesim_universe** out);
void sdesim_free_universe(sdesim_universe* u);

/* ProcessUniverse accessors (src/proc/mod.rs:61-68). */
int32_t sdesim_num_processes(const sdesim_universe* u);
int32_t sdesim_num_increments(const sdesim_universe* u); /* stochastic_registry.len() */
const char* sdesim_process_name(const sdesim_universe* u, int32_t p);
int32_t sdesim_process_is_algebraic(const sdesim_universe* u, int32_t p);
int32_t sdesim_process_index(const sdesim_universe* u, const char* name); /* registry lookup; -1 if absent */

/* Plan a run of sim::simulate (src/sim/mod.rs:20-92) for a fixed time grid, initial row and
 * path count: generates and compiles the fused kernel (hipRTC, cached by source text) and
 * uploads the constant inputs. `x0` holds one initial value per process (the caller resolves
 * names as ScenarioFiltration::new does, src/filtration.rs:42-50). `payoff` (nullable) is an
 * expression over process names and t evaluated on the terminal row. */
int sdesim_plan_create(const sdesim_universe* u, const double* times, int32_t n_times, const double* x0,
                       int64_t scenarios, const sdesim_options* opt, const char* payoff, sdesim_plan** out);
/* Frees the plan. A plan whose run was issued on a capturing stream (HIP graph capture) keeps its
 * device buffers (time grid, initial row, tables, partials) until the process ends, since the
 * captured graph may still launch with them. */
void sdesim_plan_free(sdesim_plan* plan);

/* Bytes of the [T][P][S] filtration buffer the plan writes (4 or 8 bytes per value). */
int64_t sdesim_plan_filtration_bytes(const sdesim_plan* plan);

/* Run the plan on `stream`: the kernel writes the step-major filtration [T][P][S] into
 * d_filtration (device, may be NULL when SDESIM_OUT_FILTRATION is off) and, with
 * SDESIM_OUT_TERMINAL, reduces terminal moments into the plan. `d_injected` ([N*m][S] f64,
 * device) is read only with SDESIM_INJECTED. Asynchronous. */
int sdesim_plan_run(sdesim_plan* plan, void* d_filtration, const double* d_injected, void* stream);

/* Terminal moments of the last run: for each process p (and the payoff last, if any)
 * out[2q] = mean, out[2q+1] = sample variance (ddof 1). Synchronises the stream. */
int sdesim_plan_moments(sdesim_plan* plan, double* h_out, void* stream);

/* Allocation-free form of sdesim_plan_run for a plan with SDESIM_OUT_TERMINAL. No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The kernel's per-block partials
 * go to a device array of SDESIM_SCRATCH_BYTES inside the library (one per device, kept until the
 * process ends) instead of the plan's buffer, and the terminal moments, laid out as
 * sdesim_plan_moments returns them, are written to d_moments (device memory, 2 * (processes + 1
 * if the plan has a payoff) doubles). The call allocates nothing and copies nothing to the host;
 * it is asynchronous on `stream`. A plan whose partials (blocks x 2 x quantities x 8 bytes)
 * exceed SDESIM_SCRATCH_BYTES gets SDESIM_EVALUE (use sdesim_plan_run). Every call on a device
 * shares that device's array, which keeps the last call's contents: calls that overlap on one
 * device (other streams, other threads) race, and the caller must order them. */
#define SDESIM_SCRATCH_BYTES (4u << 20)
int sdesim_plan_run_scratch(sdesim_plan* plan, void* d_filtration, const double* d_injected, double* d_moments,
                            void* stream);

/* The reference's long layout (src/filtration.rs:81-116, rows ordered scenario -> time ->
 * process, src/sim/mod.rs:88-91): transpose a [T][P][S] device filtration (f32 or f64) into
 * the value column [S][T][P] as f64 on the device. */
int sdesim_filtration_to_long(const void* d_filtration, int32_t precision, int64_t scenarios, int32_t n_times,
                              int32_t n_processes, double* d_long, void* stream);

/* One-call equivalent of the PyO3 simulate (src/py_binding.rs:8-56) minus the DataFrame
 * assembly: validates, plans, runs and writes the long value column [S][T][P] (f64) to host
 * memory. Initial values are given by name as in the reference. */
int sdesim_simulate_long(const sdesim_universe* u, const double* times, int32_t n_times,
                         const char* const* init_names, const double* init_values, int32_t n_init,
                         int64_t scenarios, const sdesim_options* opt, double* h_values, void* stream);

/* Generated HIP source of a plan's kernel (for inspection); valid while the plan lives. */
const char* sdesim_plan_source(const sdesim_plan* plan);
/* The identity of the plan's compiled kernel: a 64-bit FNV-1a hash of its generated source, the
 * device headers, the compile options and the hipRTC version (the key of the on-disk kernel
 * cache), so measurements of a kernel can be tied to the code they measured. 0 for a plan with
 * no kernel. No reference counterpart. */
uint64_t sdesim_plan_code_hash(const sdesim_plan* plan);

/* Host-only helpers (no device needed): the kernel source a plan would generate for a time grid
 * of n_times points, and a gfx950 compile of a source through hipRTC. `needed` receives the
 * buffer size including the terminating NUL. */
int sdesim_generate_source(const sdesim_universe* u, int32_t n_times, const sdesim_options* opt, const char* payoff,
                           char* buf, int64_t buf_len, int64_t* needed);
int sdesim_compile_source(const char* source, int64_t* code_bytes);

/* Compiled run-time kernels are kept per process (by device and source) and on disk in the cache
 * directory (SDESIM_CACHE_DIR, else $HOME/.cache/sdesim), so a new process loads a kernel it has
 * compiled before instead of compiling it again. No reference counterpart (the reference
 * compiles nothing at run time). sdesim_kernel_cache_reset forgets the per-process table: the
 * next plan of a given source loads it from disk or compiles it. Modules already loaded stay
 * loaded until the process ends, so existing plans keep working. */
void sdesim_kernel_cache_reset(void);

/* Host evaluation of one coefficient expression with the kernels' semantics (the value
 * Function::eval, src/func.rs:32-42, returns for that namespace). An undefined name returns
 * SDESIM_EPANIC, as the reference's unwrap does. */
int sdesim_eval_expression(const char* expr, const char* const* names, const double* values, int32_t n,
                           double* out);

/* Sobol points in Gray-code order from the same sliced tables the kernels read: out[i][d] for
 * point index first + i, dims dimensions (the reference's SobolEngine, src/rng/sobol.rs:10-26,
 * before its f64 conversion x * 2^-32). */
int sdesim_sobol_points(int32_t dims, uint64_t first, int64_t n, uint32_t* out);

/* ---------------------------------------------------------------------------------------
 * Batches: many independent 1-D instances of uneven size in one persistent launch.
 * No reference counterpart (the reference runs one universe per simulate call,
 * src/sim/mod.rs:20-92); each instance is what simulate would compute for its built-in
 * model's DSL form on the uniform grid t_n = n * horizon / steps with seed `seed`:
 *   kind 0 GBM  dX = (p0 * X) * dt + (p1 * X) * dW1
 *   kind 1 OU   dX = (p0 * (p1 - X)) * dt + (p2) * dW1
 *   kind 2 CIR  dX = (p0 * (p1 - max(X, 0))) * dt + (p2 * max(X, 0)^0.5) * dW1
 * Euler-Maruyama, pseudo-random (ChaCha8, path s seeded seed + s), terminal mean and variance.
 * --------------------------------------------------------------------------------------- */
typedef struct sdesim_batch sdesim_batch;

typedef struct sdesim_instance {
  int32_t kind;    /* 0 GBM, 1 OU, 2 CIR */
  int32_t steps;   /* N >= 1 */
  int64_t paths;   /* S >= 1 */
  double x0;
  double horizon;  /* T > 0 */
  double p0, p1, p2;
  uint64_t seed;
} sdesim_instance;

/* keep_terminal != 0 keeps every path's terminal value (instance-major) for sdesim_batch_terminal. */
int sdesim_batch_create(const sdesim_instance* instances, int32_t n, int32_t precision, int32_t keep_terminal,
                        sdesim_batch** out);
void sdesim_batch_free(sdesim_batch* b);
/* paths_per_lane in {1, 2, 4}, block_size in {64, 128, 256}; drops a captured graph. */
int sdesim_batch_set_config(sdesim_batch* b, int32_t paths_per_lane, int32_t block_size);
/* Time every configuration on this batch (or reuse a remembered choice for this device and
 * workload shape) and apply the fastest. */
int sdesim_batch_autotune(sdesim_batch* b, void* stream, int32_t* paths_per_lane, int32_t* block_size, double* best_ms);
/* Eager run: counter reset, persistent kernel, per-instance finalize. Asynchronous. */
int sdesim_batch_run(sdesim_batch* b, void* stream);
/* Capture the three launches into a hipGraph (on a stream the batch owns) / replay it on `stream`. */
int sdesim_batch_capture(sdesim_batch* b);
int sdesim_batch_replay(sdesim_batch* b, void* stream);
/* out[2i] = mean, out[2i+1] = sample variance of instance i's terminal values. Synchronises. */
int sdesim_batch_moments(sdesim_batch* b, double* h_out, void* stream);
/* Terminal values (f32 or f64 by precision), instance-major. Synchronises. */
int sdesim_batch_terminal(sdesim_batch* b, void* h_out, void* stream);
int sdesim_batch_info(const sdesim_batch* b, int32_t* paths_per_lane, int32_t* block_size, int32_t* grid,
                      int32_t* n_items, int64_t* total_paths);

/* ---------------------------------------------------------------------------------------
 * Correlated GBM basket: price and pathwise adjoint. No reference counterpart (the reference
 * computes no sensitivities). Asset i follows the Euler step of the DSL form
 *   dX_i = (mu_i * X_i) * dt + sum_k (sigma_i * L_ik * X_i) * dW_k
 * on the grid dt = horizon / steps with pseudo-random draws of ChaCha8(seed + s); payoff on
 * B = sum_i w_i X_{N,i}: width * softplus((B - strike) / width), or max(B - strike, 0) if
 * width == 0. The adjoint checkpoints the state every 19 steps in HBM and recomputes each
 * segment during the reverse sweep.
 * --------------------------------------------------------------------------------------- */
typedef struct sdesim_basket sdesim_basket;

typedef struct sdesim_basket_spec {
  int32_t assets;     /* 1..8 */
  int32_t steps;
  int64_t paths;
  double horizon;
  double strike;
  double width;
  uint64_t seed;      /* path s uses seed + s */
  int32_t precision;  /* SDESIM_F32 | SDESIM_F64 (path arithmetic; path sums are f64) */
  int32_t storage;    /* adjoint state storage: SDESIM_STORAGE_AUTO, _TAPE or _CHECKPOINTS */
} sdesim_basket_spec;

/* Adjoint state storage (basket and forward-rate model). TAPE keeps every state in device memory
 * ((steps + 1) * state size * paths values in the path precision, one write and one read each;
 * one step's states must stay below 2^31 bytes). CHECKPOINTS keeps every 19th (basket) or 16th
 * (forward-rate model) state and recomputes the rest during the reverse sweep (one more forward
 * pass of arithmetic, for that factor less memory). AUTO takes the tape when it fits in half of
 * the free device memory at creation, otherwise checkpoints. */
#define SDESIM_STORAGE_AUTO 0
#define SDESIM_STORAGE_TAPE 1
#define SDESIM_STORAGE_CHECKPOINTS 2

int sdesim_basket_create(const sdesim_basket_spec* spec, sdesim_basket** out);
void sdesim_basket_free(sdesim_basket* b);
/* The storage the handle uses: SDESIM_STORAGE_TAPE or SDESIM_STORAGE_CHECKPOINTS. */
int32_t sdesim_basket_storage(const sdesim_basket* b);
/* chol: lower Cholesky factor of the correlation, row-major [assets][assets]. out[0] = mean
 * payoff, out[1] = its sample variance; with adjoint != 0 also the path-averaged derivatives
 * d/dx0[assets], d/dmu[assets], d/dsigma[assets], d/dL[assets (assets + 1) / 2] (row-major lower
 * triangle). Synchronises the stream. */
int sdesim_basket_eval(sdesim_basket* b, const double* x0, const double* mu, const double* sigma,
                       const double* weight, const double* chol, int32_t adjoint, double* out, void* stream);

/* ---------------------------------------------------------------------------------------
 * Forward-rate (LIBOR market) model: receiver swaption price and the 40 curve sensitivities.
 * No reference counterpart. 40 forwards on tenor `tenor`, 3 factors (loadings [40][3], unit
 * rows), constant volatilities; spot-measure log-Euler with the drift as a prefix sum over
 * forwards; `steps` = steps_per_period * expiry_index steps to expiry T_e; uniforms from the
 * Sobol sequence (dimension steps * 3, point index = path) with a digital XOR shift drawn from
 * scramble_seed. Payoff max(K A - (1 - P_last), 0) / numeraire (swap over [T_e, T_40]).
 * --------------------------------------------------------------------------------------- */
typedef struct sdesim_lmm sdesim_lmm;

typedef struct sdesim_lmm_spec {
  int32_t forwards;          /* 40 */
  int32_t factors;           /* 3 */
  int32_t steps;
  int32_t steps_per_period;
  int32_t expiry_index;      /* e: the swaption expires at T_e = e * tenor */
  int32_t precision;         /* SDESIM_F32 | SDESIM_F64 */
  int64_t paths;             /* <= 2^22 */
  double tenor;
  double strike;
  uint64_t scramble_seed;
  int32_t storage;           /* adjoint state storage: SDESIM_STORAGE_AUTO, _TAPE or _CHECKPOINTS
                              * (checkpoints every 16 steps, each segment recomputed once) */
  int32_t pad_;
} sdesim_lmm_spec;

int sdesim_lmm_create(const sdesim_lmm_spec* spec, sdesim_lmm** out);
void sdesim_lmm_free(sdesim_lmm* m);
/* The storage the handle uses: SDESIM_STORAGE_TAPE or SDESIM_STORAGE_CHECKPOINTS. */
int32_t sdesim_lmm_storage(const sdesim_lmm* m);
/* out[0] = price, out[1] = sample variance of the discounted payoff, and with adjoint != 0
 * out[2 + i] = d price / d L_i(0), i = 0..39. Synchronises the stream. */
int sdesim_lmm_eval(sdesim_lmm* m, const double* L0, const double* vol, const double* loadings, int32_t adjoint,
                    double* out, void* stream);

const char* sdesim_last_error(void);
int32_t sdesim_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SDESIM_H */

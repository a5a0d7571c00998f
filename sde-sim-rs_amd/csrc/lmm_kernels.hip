// lmm_kernels.hip -- multi-factor forward-rate (LIBOR market) model: swaption price and pathwise
// curve sensitivities. No reference counterpart (the reference has no built-in models and no
// sensitivities); the stepping is log-Euler with uniforms from the same scrambled Sobol stream
// as the generic kernel (src/rng/sobol.rs points, XOR digital shift) and the A&S inverse normal
// (src/proc/increment.rs:160-179).
//
// Forwards L_i on [T_i, T_{i+1}], T_i = i * tau, i = 0 .. NF-1, NF = 40; F = 3 factors with
// loadings B (NF x F, unit rows); volatilities sig_i. Spot measure: at step n (t_n = n dt) the
// alive forwards are i >= eta = n / steps_per_period + 1, and for them
//   ln L_i += (sig_i sum_f B_if C_f(i) - sig_i^2 / 2) dt + sig_i sqrt(dt) sum_f B_if eps_f,
//   C_f(i) = sum_{j=eta..i} tau sig_j L_j / (1 + tau L_j) B_jf     (a prefix sum over j),
// eps_f = Psi(u_{n F + f}). Dead forwards keep their fixing value.
// Payoff at expiry T_e (e = expiry index): receiver swaption on the swap over [T_e, T_NF] with
// fixed rate K: V = max(K A - (1 - P_{NF-1}), 0) / N, P_j = prod_{k=e..j} 1 / (1 + tau L_k),
// A = tau sum_{j=e..NF-1} P_j, numeraire N = prod_{k<e} (1 + tau L_k).
//
// Adjoint: the reverse sweep needs every state L_n. Two storage forms:
//  * state tape (MODE 2): the forward pass writes L_0 .. L_N to HBM ([n][i][path], 27 GB for
//    161 states x 40 forwards x 2^20 paths in f32) and a second kernel reads them back;
//  * checkpoints (MODE 1): the forward pass stores L every SEG steps ([slot][i][path]); the
//    reverse sweep takes the segments last to first, recomputes a segment's states once from its
//    checkpoint into the segment tape ([k][i][path], rewritten per segment) and then steps back
//    through them.
// For one step, with a_j = tau sig_j L_j / (1 + tau L_j) and
// E_i = L'_i / L_i (read as the ratio of the stored states, so the reverse step needs no exp):
//   Cbar_f(i) = Lbar'_i L'_i dt sig_i B_if,   G_f(j) = sum_{i >= j} Cbar_f(i)  (a suffix sum),
//   Lbar_j = Lbar'_j E_j + (sum_f B_jf G_f(j)) tau sig_j / (1 + tau L_j)^2.
#include <hip/hip_runtime.h>

#include "device/sdesim_device.hpp"
#include "adjoint.hpp"

namespace {

constexpr int NF = 40;
constexpr int F = 3;
constexpr int SEG = 16;  // checkpoint interval (~sqrt of the 160-step workload)
constexpr int NSL = 3;  // Sobol table slices: point indices below 2^22

// exp of an exponent given in the parameter rows' units: base 2 for f32 (the rows carry the
// log2(e) factor, so this is one v_exp_f32), natural for f64.
template <typename R>
__device__ __forceinline__ R lmm_exp_row(R x) {
  if constexpr (sizeof(R) == 4) return __builtin_amdgcn_exp2f(x);
  else return exp(x);
}

// Where a lane's Sobol words live: its slice-0 entry (byte offset within a step's [64][F] block)
// and the wave's rows of slices 1-2 (the wave's 64 point indices share those Gray-code bits).
struct LmmSobol {
  sde_u32 voff0;
  sde_const_u32* urow[NSL - 1];
};

// The F uniforms of step n (digital XOR shift folded into the tables: the XOR of the slices is the
// complement of the scrambled point), as standard normals: one vector load of the lane's F
// slice-0 words and scalar loads of the wave's words of slices 1-2.
template <typename R>
__device__ __forceinline__ void lmm_noise(__amdgpu_buffer_rsrc_t rsrc, const LmmSobol& sb, int n, R (&eps)[F]) {
  sde_u32 w[F];
  sde_load_words<F>(rsrc, (int)sb.voff0, n * (64 * F * 4), w);
#pragma unroll
  for (int f = 0; f < F; ++f) {
    SdeUnifX u;
    u.yc = w[f] ^ sb.urow[0][n * F + f] ^ sb.urow[1][n * F + f];
    u.tail = nullptr;
    u.sm = u.sm2 = 0x80000000u;
    eps[f] = sde_psi<R>(u);
  }
}

// Model parameters: the workgroup stages them once from the device buffer ([40] volatilities
// then [40][3] loadings) into LDS, one 8-value row per forward with the step's constant products
// precomputed, so a forward's parameters are two broadcast LDS reads. Each step reads them at
// index i + off, where off = step * a.zero is 0 at run time but not known to the compiler: the
// reads repeat per step (issued ahead of use within the step) instead of 320 loop-invariant
// values being hoisted into registers for the whole kernel.
template <typename R>
struct LmmParRow {
  R ts;           // tau sigma_i
  R b0, b1, b2;   // loadings B_i0..B_i2
  R q0, q1, q2;   // sigma_i B_i0..B_i2
  R h;            // -sigma_i^2 dt / 2 (f32: times log2 e)
};

template <typename R>
struct LmmPar {
  const LmmParRow<R>* rows;  // LDS, already advanced by the step's offset
  __device__ __forceinline__ LmmParRow<R> row(int i) const { return rows[i]; }
};

// The step's view of the parameter rows: rows + off, with the offset moved into a vector register
// the compiler cannot see through, so the LDS address is formed once per step and each forward's
// reads use it with constant offsets (with the offset in a scalar register the compiler copies
// the address into a vector register again in every forward's branch).
template <typename R>
__device__ __forceinline__ LmmPar<R> lmm_par(const LmmParRow<R>* rows, int off) {
  asm volatile("" : "+v"(off));
  return LmmPar<R>{rows + off};
}

template <typename R>
__device__ __forceinline__ R lmm_recip(R x) {
  if constexpr (sizeof(R) == 4) return __builtin_amdgcn_rcpf(x);  // v_rcp_f32, 1 ulp
  else return R(1) / x;
}

// One log-Euler step of all alive forwards (forwards below `eta` are dead and unchanged). The
// exponent (sigma_i (B_i.C) - sigma_i^2 / 2) dt + sigma_i sqrt(dt) (B_i.eps) is formed as
// q_i.(dt C + sqrt(dt) eps) + h_i with q_i = sigma_i B_i: the step's sqrt(dt) eps is shared by all
// forwards. dtu and sdtu are dt and sqrt(dt) in the exponent's unit (log2 e for f32).
template <typename R>
__device__ __forceinline__ void lmm_step(const LmmPar<R>& tab, R (&L)[NF], const R (&eps)[F], int eta, R dtu, R sdtu,
                                         R tau) {
  R C[F] = {R(0), R(0), R(0)};
  const R se0 = sdtu * eps[0], se1 = sdtu * eps[1], se2 = sdtu * eps[2];
#pragma unroll
  for (int i = 0; i < NF; ++i) {
    if (i >= eta) {
      const LmmParRow<R> pr = tab.row(i);
      const R ai = pr.ts * L[i] * lmm_recip<R>(R(1) + tau * L[i]);
      C[0] = C[0] + ai * pr.b0;
      C[1] = C[1] + ai * pr.b1;
      C[2] = C[2] + ai * pr.b2;
      const R x = pr.q0 * (dtu * C[0] + se0) + pr.q1 * (dtu * C[1] + se1) + pr.q2 * (dtu * C[2] + se2) + pr.h;
      L[i] = L[i] * lmm_exp_row<R>(x);
    }
  }
}

// Adjoint of one forward's part of lmm_step, forwards taken from the last down: Li and Lpi are
// L_i before and after the step, Lbi the adjoint of Lp_i on entry and of L_i on exit, G the
// running sums over the forwards above. Forwards are positive, so E_i = Lp_i / L_i.
template <typename R>
__device__ __forceinline__ void lmm_forward_adjoint(const LmmParRow<R>& pr, R Li, R Lpi, R& Lbi, R (&G)[F], R dt,
                                                    R tau) {
  const R b0 = pr.b0, b1 = pr.b1, b2 = pr.b2;
  const R db = Lbi * Lpi * dt;
  G[0] = G[0] + db * pr.q0;
  G[1] = G[1] + db * pr.q1;
  G[2] = G[2] + db * pr.q2;
  const R ab = b0 * G[0] + b1 * G[1] + b2 * G[2];
  const R rop = lmm_recip<R>(R(1) + tau * Li);
  Lbi = Lbi * (Lpi * lmm_recip<R>(Li)) + ab * pr.ts * (rop * rop);
}

// Adjoint of lmm_step: L is the state before the step and Lp the state after it; on entry Lb is
// the adjoint of Lp, on exit the adjoint of L.
template <typename R>
__device__ __forceinline__ void lmm_step_adjoint(const LmmPar<R>& tab, const R (&L)[NF], const R (&Lp)[NF],
                                                 R (&Lb)[NF], int eta, R dt, R tau) {
  R G[F] = {R(0), R(0), R(0)};
#pragma unroll
  for (int i = NF - 1; i >= 0; --i) {
    if (i >= eta) lmm_forward_adjoint<R>(tab.row(i), L[i], Lp[i], Lb[i], G, dt, tau);
  }
}

// Payoff at the final forwards L (discounted by the spot numeraire) and, with GRAD, its
// derivatives dV/dL_k in Lb: the numeraire part for k < e, the swap part for k >= e (only when
// in the money).
template <typename R, bool GRAD>
__device__ __forceinline__ R lmm_payoff(const R (&L)[NF], int e, R tau, R K, R (&Lb)[NF]) {
  R numer = R(1);
#pragma unroll
  for (int k = 0; k < NF; ++k)
    if (k < e) numer = numer * (R(1) + tau * L[k]);
  R P = R(1), A = R(0);
  R Pj[NF];
#pragma unroll
  for (int j = 0; j < NF; ++j) {
    Pj[j] = R(0);
    if (j >= e) {
      P = P / (R(1) + tau * L[j]);
      Pj[j] = P;
      A = A + tau * P;
    }
  }
  const R U = K * A - (R(1) - P);
  const R V = U > R(0) ? U / numer : R(0);
  if (GRAD) {
    R tail = R(0);  // sum_{j >= k} P_j, built from the end
#pragma unroll
    for (int k = NF - 1; k >= 0; --k) {
      if (k >= e) tail = tail + Pj[k];
      const R op = R(1) + tau * L[k];
      if (k < e) {
        Lb[k] = -V * tau / op;
      } else {
        Lb[k] = U > R(0) ? (-K * tau * tau * tail - tau * P) / (op * numer) : R(0);
      }
    }
  }
  return V;
}

// Model parameters staged into LDS by the first NF threads of the workgroup.
template <typename R>
__device__ __forceinline__ void lmm_stage_params(const LmmArgs& a, LmmParRow<R>* par_rows) {
  if (threadIdx.x < NF) {
    const R dt = (R)a.dt, tau = (R)a.tau;
    const R* par = sizeof(R) == 4 ? (const R*)(const void*)a.parf : (const R*)(const void*)a.pard;
    const int i = (int)threadIdx.x;
    const R sig = par[i];
    const R un = sizeof(R) == 4 ? R(1.4426950408889634) : R(1);  // exponent unit: log2 e for f32
    const R b0 = par[NF + i * F], b1 = par[NF + i * F + 1], b2 = par[NF + i * F + 2];
    par_rows[i] = LmmParRow<R>{tau * sig, b0, b1, b2, sig * b0, sig * b1, sig * b2, R(-0.5) * sig * sig * dt * un};
  }
  __syncthreads();
}

// MODE 0: price; 1: price and adjoint with checkpoints, the reverse sweep recomputing each
// segment into the segment tape; 2: the forward pass of the tape form, which writes L_0 .. L_N to
// the state tape and leaves the payoff, the reverse sweep and every partial to
// sdesim_lmm_reverse. The f32 checkpoint adjoint is held to 3 waves per SIMD (168 VGPRs): the
// extra wave hides the tape and LDS latency of its reverse sweep.
template <typename R, int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MODE == 1 && sizeof(R) == 4 ? 3 : 1)))
sdesim_lmm(const LmmArgs a) {
  constexpr int NQ = 2 + NF;
  const long long S = a.paths;
  const int N = a.steps;
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = s < S;
  const long long sid = live ? s : 0;
  const sde_u32 voff = live ? (sde_u32)(sid * (long long)sizeof(R)) : 0x80000000u;
  const R dt = (R)a.dt, tau = (R)a.tau, K = (R)a.strike;
  const R un = sizeof(R) == 4 ? R(1.4426950408889634) : R(1);  // exponent unit: log2 e for f32
  const R dtu = (R)a.dt * un, sdtu = (R)a.sdt * un;
  const int e = a.expiry_index;

  __shared__ LmmParRow<R> par_rows[NF];
  lmm_stage_params<R>(a, par_rows);

  LmmSobol sb;
  {
    // Point index = path index; blocks of 256 make every wave one aligned run of 64 indices.
    // Lane 0 has the wave's smallest index: if it is past the last path so is every lane of the
    // wave, whose values are then never used.
    const sde_u64 g = (sde_u64)sid ^ ((sde_u64)sid >> 1);
    sb.voff0 = (sde_u32)(g & 63u) * (sde_u32)(F * 4);
#pragma unroll
    for (int sl = 1; sl < NSL; ++sl) {
      const sde_u32 e = __builtin_amdgcn_readfirstlane((sde_u32)(g >> (6 + 8 * (sl - 1))) & 255u);
      sb.urow[sl - 1] = (sde_const_u32*)a.sobol_up + ((size_t)(sl - 1) * 256u + e) * (size_t)a.sobol_dims;
    }
  }
  __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc((void*)a.sobol_tab, (short)0, a.sobol_tab_words * 4, 0x00020000);

  R L[NF];
#pragma unroll
  for (int i = 0; i < NF; ++i) L[i] = (R)a.L0[i];
  R* ck = (R*)a.checkpoints;
  for (int n = 0; n < N; ++n) {
    if (MODE == 2) sde_tape_store<R, NF>(a.tape, n, S, voff, L);
    if (MODE == 1 && (n % SEG) == 0 && live) {
      const long long slot = n / SEG;
#pragma unroll
      for (int i = 0; i < NF; ++i) ck[(slot * NF + i) * S + sid] = L[i];
    }
    R eps[F];
    lmm_noise<R>(rsrc, sb, n, eps);
    lmm_step<R>(lmm_par<R>(par_rows, n * a.zero), L, eps, n / a.steps_per_period + 1, dtu, sdtu, tau);
  }
  if (MODE == 2) {
    sde_tape_store<R, NF>(a.tape, N, S, voff, L);
    return;
  }

  R Lb[NF];
  const R V = lmm_payoff<R, MODE == 1>(L, e, tau, K, Lb);
  if (MODE == 1) {
    // Lp holds the state after the step being reversed: L_N first, then each state read back.
    R Lp[NF];
#pragma unroll
    for (int i = 0; i < NF; ++i) Lp[i] = L[i];
    R* tape = (R*)a.tape;
    const int nseg = (N + SEG - 1) / SEG;
    for (int seg = nseg - 1; seg >= 0; --seg) {
      const int n0 = seg * SEG;
      const int len = N - n0 < SEG ? N - n0 : SEG;
      R Xc[NF];
#pragma unroll
      for (int i = 0; i < NF; ++i) Xc[i] = ck[((long long)seg * NF + i) * S + sid];
      // States before steps n0 .. n0 + len - 2 go to the tape; the last one stays in Xc.
      for (int k = 0; k + 1 < len; ++k) {
        if (live) {
#pragma unroll
          for (int i = 0; i < NF; ++i) tape[((long long)k * NF + i) * S + sid] = Xc[i];
        }
        R eps[F];
        lmm_noise<R>(rsrc, sb, n0 + k, eps);
        lmm_step<R>(lmm_par<R>(par_rows, (n0 + k) * a.zero), Xc, eps, (n0 + k) / a.steps_per_period + 1, dtu, sdtu,
                    tau);
      }
      for (int k = len - 1; k >= 0; --k) {
        if (k + 1 < len) {
#pragma unroll
          for (int i = 0; i < NF; ++i) Xc[i] = tape[((long long)k * NF + i) * S + sid];
        }
        lmm_step_adjoint<R>(lmm_par<R>(par_rows, (n0 + k) * a.zero), Xc, Lp, Lb, (n0 + k) / a.steps_per_period + 1,
                            dt, tau);
#pragma unroll
        for (int i = 0; i < NF; ++i) Lp[i] = Xc[i];
      }
    }
  }
  sde_reduce_store_chunked<NQ>(
      [&](int q) {
        return !live ? 0.0 : q == 0 ? (double)V : q == 1 ? (double)V * (double)V : MODE == 1 ? (double)Lb[q - 2] : 0.0;
      },
      a.partials + (long long)blockIdx.x * NQ);
}

// Reverse sweep of the tape form. L_N from the tape gives the payoff; then, for each step from
// the last, the adjoint runs over the forwards from the last down, and as soon as forward i of
// the state is used its register is refilled with forward i of the previous step's state (the
// request after step 0 re-reads step 0 and is unused). The loads are issued in the order in
// which the next step consumes them, so each has the rest of the step to arrive. The f32 kernel is
// held to 3 waves per SIMD (168 VGPRs; it needs 169 unconstrained).
template <typename R>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 3 : 1)))
sdesim_lmm_reverse(const LmmArgs a) {
  constexpr int NQ = 2 + NF;
  const long long S = a.paths;
  const int N = a.steps;
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = s < S;
  const sde_u32 voff = live ? (sde_u32)(s * (long long)sizeof(R)) : 0x80000000u;
  const R dt = (R)a.dt, tau = (R)a.tau, K = (R)a.strike;

  __shared__ LmmParRow<R> par_rows[NF];
  lmm_stage_params<R>(a, par_rows);

  R L[NF], Lb[NF], Lp[NF];
  sde_tape_load<R, NF>(a.tape, N, S, voff, Lp);
  const R V = lmm_payoff<R, true>(Lp, a.expiry_index, tau, K, Lb);
  sde_tape_load<R, NF>(a.tape, N - 1, S, voff, L);
  for (int n = N - 1; n >= 0; --n) {
    const LmmPar<R> tab = lmm_par<R>(par_rows, n * a.zero);
    const int eta = n / a.steps_per_period + 1;
    const __amdgpu_buffer_rsrc_t rs = sde_tape_rsrc<R, NF>(a.tape, n > 0 ? n - 1 : 0, S);
    R G[F] = {R(0), R(0), R(0)};
#pragma unroll
    for (int i = NF - 1; i >= 0; --i) {
      if (i >= eta) lmm_forward_adjoint<R>(tab.row(i), L[i], Lp[i], Lb[i], G, dt, tau);
      Lp[i] = L[i];
      L[i] = sde_tape_get<R>(rs, voff, i, S);
    }
  }
  sde_reduce_store_chunked<NQ>(
      [&](int q) { return !live ? 0.0 : q == 0 ? (double)V : q == 1 ? (double)V * (double)V : (double)Lb[q - 2]; },
      a.partials + (long long)blockIdx.x * NQ);
}

}  // namespace

// mode: 0 price, 1 adjoint with checkpoints and the segment tape, 2 adjoint with the state tape
// (a.tape holds [steps + 1][40][paths]; a forward and a reverse kernel).
hipError_t sdesim_launch_lmm(int f32, int mode, int grid, const LmmArgs& a, hipStream_t stream) {
#define SDE_LMM(R, MODE)                                                                                  \
  hipLaunchKernelGGL((sdesim_lmm<R, MODE>), dim3(grid), dim3(256), 0, stream, a);                            \
  if (MODE == 2) hipLaunchKernelGGL((sdesim_lmm_reverse<R>), dim3(grid), dim3(256), 0, stream, a);          \
  break;
  switch ((f32 ? 10 : 0) + mode) {
    case 12: SDE_LMM(float, 2)
    case 11: SDE_LMM(float, 1)
    case 10: SDE_LMM(float, 0)
    case 2: SDE_LMM(double, 2)
    case 1: SDE_LMM(double, 1)
    case 0: SDE_LMM(double, 0)
    default: return hipErrorInvalidValue;
  }
#undef SDE_LMM
  return hipGetLastError();
}

// basket_kernels.hip -- correlated GBM basket: price and pathwise adjoint (reverse sweep).
//
// Model: A assets, X_{n+1,i} = X_{n,i} + X_{n,i} (mu_i dt + sigma_i sqrt(dt) Z_{n,i}),
// Z_n = L eps_n with L the lower Cholesky factor of the correlation and eps_{n,k} = Psi(u_{n,k})
// (A&S inverse normal, src/proc/increment.rs:160-179) of the pseudo-random draw n*A + k of
// ChaCha8Rng::seed_from_u64(seed + s) -- the Euler step of the reference (src/sim/euler.rs:5-37)
// applied to the DSL form dX_i = (mu_i * X_i) * dt + sum_k (sigma_i * L_ik * X_i) * dW_k.
// Payoff on the weighted basket B = sum_i w_i X_{N,i}: width * softplus((B - K) / width), or the
// plain call max(B - K, 0) when width == 0.
//
// Adjoint (no reference counterpart): pathwise derivatives of the payoff with respect to X0, mu,
// sigma and L, per path, reduced over paths in f64. The reverse sweep needs every state X_n. Two
// storage forms:
//  * tape (MODE 2): the forward pass writes every state to HBM ([step][asset][path], 4 or 8 bytes
//    each; 49 GB for 8 assets x 365 steps x 2^22 paths in f32) and the reverse sweep reads them
//    back, one step ahead of its arithmetic;
//  * checkpoints (MODE 1): the forward pass stores the state every SEG steps ([slot][asset][path]);
//    the reverse sweep walks the segments backwards, recomputes each segment's states from its
//    checkpoint into registers and then steps back through it (one more forward pass of
//    arithmetic, for SEG times less memory).
// A reverse step reads the step's noise off the two states instead of regenerating it:
//   g_n = X_{n+1} / X_n = 1 + mu dt + sigma sqrt(dt) Z_n,  d_n = (X_{n+1} - X_n) - X_n mu dt
// (X_{n+1} - X_n is exact, the two being within a factor 2; forming 1 + mu dt first would round
// mu dt ~ 2e-4 to the spacing of 1 and bias every recovered Z_n)
//   Xbar_n = Xbar_{n+1} g_n,  mubar += Xbar_{n+1} X_n dt,  sigmabar += Xbar_{n+1} d_n / sigma
//   Zbar_n = Xbar_{n+1} X_n sigma sqrt(dt),  Z_n = d_n / (X_n sigma sqrt(dt)),
//   eps_n = L^-1 Z_n (forward substitution),  Lbar_ik += Zbar_{n,i} eps_{n,k}  (k <= i)
#include <hip/hip_runtime.h>

#include "device/sdesim_device.hpp"
#include "adjoint.hpp"

namespace {

template <int D>
struct BasketConst {
  static constexpr int NL = D * (D + 1) / 2;   // lower-triangular entries of L
  static constexpr int NQ = 2 + 3 * D + NL;    // reduced quantities: payoff, payoff^2, gradients
};

// The words (low, high) of the D f64 draws of step n: draws n*D .. n*D + D - 1 of the path's
// stream (draw j is words 2j and 2j + 1). PADDED (D = 8 holding A < 8 assets, the rest padding
// with zero weight): step n uses draws n*A .. n*A + A - 1, as the DSL form with A Wiener
// increments does; they lie in ChaCha blocks (n*A) / 8 and the next one, and the wave-uniform
// word offset (n*A) mod 8 selects them (the padding assets take the following draws, unused).
template <int D, bool PADDED>
__device__ __forceinline__ void basket_draws(const sde_u32 (&key)[8], int n, int A, sde_u32 (&lo)[D],
                                             sde_u32 (&hi)[D]) {
  if constexpr (PADDED) {
    static_assert(D == 8, "padded baskets run the 8-asset kernels");
    const long long j0 = (long long)n * A;
    sde_u32 b0[16], b1[16];
    sde_chacha_block<8>(key, (sde_u64)(j0 >> 3), b0);
    sde_chacha_block<8>(key, (sde_u64)(j0 >> 3) + 1, b1);
    const int w0 = (int)(j0 & 7);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int w = w0 + k;  // < 15
      sde_u32 l = b0[0], h = b0[1];
#pragma unroll
      for (int q = 1; q < 16; ++q) {
        l = (w == q) ? (q < 8 ? b0[2 * q] : b1[2 * q - 16]) : l;
        h = (w == q) ? (q < 8 ? b0[2 * q + 1] : b1[2 * q - 15]) : h;
      }
      lo[k] = l;
      hi[k] = h;
    }
  } else if constexpr (D == 8) {
    sde_u32 blk[16];
    sde_chacha_block<8>(key, (sde_u64)n, blk);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      lo[k] = blk[2 * k];
      hi[k] = blk[2 * k + 1];
    }
  } else {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const long long j = (long long)n * D + k;
      sde_u32 blk[16];
      sde_chacha_block<8>(key, (sde_u64)(j >> 3), blk);
      const int w = (int)(j & 7);
      sde_u32 l = blk[0], h = blk[1];
#pragma unroll
      for (int q = 1; q < 8; ++q) {
        l = (w == q) ? blk[2 * q] : l;
        h = (w == q) ? blk[2 * q + 1] : h;
      }
      lo[k] = l;
      hi[k] = h;
    }
  }
}

template <typename R, int D>
struct BasketStep {
  R eps[D], Z[D];
};

template <typename R, int D, bool PADDED>
__device__ __forceinline__ void basket_noise(const sde_u32 (&key)[8], int n, int A, const R (&L)[D][D],
                                             BasketStep<R, D>& st) {
  sde_u32 lo[D], hi[D];
  basket_draws<D, PADDED>(key, n, A, lo, hi);
#pragma unroll
  for (int k = 0; k < D; ++k) st.eps[k] = sde_psi_words<R>(lo[k], hi[k]);
#pragma unroll
  for (int i = 0; i < D; ++i) {
    R z = R(0);
#pragma unroll
    for (int k = 0; k <= i; ++k) z = z + L[i][k] * st.eps[k];
    st.Z[i] = z;
  }
}

template <typename R>
__device__ __forceinline__ R basket_recip(R x) {
  if constexpr (sizeof(R) == 4) return __builtin_amdgcn_rcpf(x);  // v_rcp_f32, 1 ulp
  else return R(1) / x;
}

// One reverse step: x = X_n, xp = X_{n+1} (replaced by x on return), with the adjoint state
// Xb = Xbar_{n+1} (replaced by Xbar_n) and the gradient sums gm, gs (sigma's, times sigma), gL.
template <typename R, int D>
__device__ __forceinline__ void basket_reverse_step(const R (&x)[D], R (&xp)[D], R (&Xb)[D], R (&gm)[D], R (&gs)[D],
                                                    R (&gL)[D * (D + 1) / 2], const R (&L)[D][D], const R (&rLd)[D],
                                                    const R (&mdt)[D], const R (&rsd)[D], const R (&sig)[D], R sdt,
                                                    R dt) {
  R Zb[D], Z[D], eps[D];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    const R xb = Xb[i], xi = x[i], xpi = xp[i];
    const R rx = basket_recip<R>(xi);
    const R d = (xpi - xi) - xi * mdt[i];
    gm[i] = gm[i] + xb * xi * dt;
    gs[i] = gs[i] + xb * d;
    Zb[i] = xb * xi * sig[i] * sdt;
    Z[i] = d * rx * rsd[i];
    Xb[i] = xb * (xpi * rx);
    xp[i] = xi;
  }
#pragma unroll
  for (int i = 0; i < D; ++i) {
    R r = Z[i];
#pragma unroll
    for (int kk = 0; kk < i; ++kk) r = r - L[i][kk] * eps[kk];
    eps[i] = r * rLd[i];
  }
#pragma unroll
  for (int i = 0; i < D; ++i) {
#pragma unroll
    for (int kk = 0; kk <= i; ++kk) {
      const int q = i * (i + 1) / 2 + kk;
      gL[q] = gL[q] + Zb[i] * eps[kk];
    }
  }
}

template <typename R>
__device__ __forceinline__ R basket_softplus(R x) {
  // log(1 + e^x) without overflow: max(x, 0) + log1p(exp(-|x|))
  return (x > R(0) ? x : R(0)) + (R)log1p(exp(-(double)fabs((double)x)));
}

// Payoff of the weighted basket at the final state and its derivative with respect to B.
template <typename R, int D>
__device__ __forceinline__ void basket_payoff(const R (&X)[D], const R (&w)[D], R K, R width, R& pay, R& dpay) {
  R B = R(0);
#pragma unroll
  for (int i = 0; i < D; ++i) B = B + w[i] * X[i];
  if (width > R(0)) {
    const R x = (B - K) / width;
    pay = width * basket_softplus<R>(x);
    dpay = (R)(1.0 / (1.0 + exp(-(double)x)));
  } else {
    pay = B > K ? B - K : R(0);
    dpay = B > K ? R(1) : R(0);
  }
}

// The reverse sweep's parameters (wave-uniform) and per-path adjoint state.
template <typename R, int D>
struct BasketReverse {
  static constexpr int NL = BasketConst<D>::NL;
  R L[D][D], sig[D], mdt[D], rsd[D], rLd[D];  // L, sigma, mu dt, 1 / (sigma sqrt(dt)), 1 / L_ii
  R dt, sdt;
  R Xb[D], gm[D], gs[D], gL[NL];
  R Xp[D];  // the state after the step being reversed: X_N first, then each state stepped back to

  __device__ __forceinline__ BasketReverse(const BasketArgs& a, const R (&XN)[D], R dpay) {
    dt = (R)a.dt;
    sdt = (R)a.sdt;
#pragma unroll
    for (int i = 0; i < D; ++i) {
#pragma unroll
      for (int k = 0; k < D; ++k) L[i][k] = (R)a.chol[i * 8 + k];
      sig[i] = (R)a.sigma[i];
      mdt[i] = (R)a.mu[i] * dt;
      rsd[i] = R(1) / (sig[i] * sdt);
      rLd[i] = R(1) / L[i][i];
      Xb[i] = dpay * (R)a.weight[i];
      gm[i] = R(0);
      gs[i] = R(0);
      Xp[i] = XN[i];
    }
#pragma unroll
    for (int q = 0; q < NL; ++q) gL[q] = R(0);
  }
  __device__ __forceinline__ void step(const R (&x)[D]) {
    basket_reverse_step<R, D>(x, Xp, Xb, gm, gs, gL, L, rLd, mdt, rsd, sig, sdt, dt);
  }
  // Gradient q of this path, q < 3 D + NL: d/dX0, d/dmu, d/dsigma, d/dL (q a constant after
  // unrolling).
  __device__ __forceinline__ double grad(int q) const {
    if (q < D) return (double)Xb[q];
    if (q < 2 * D) return (double)gm[q - D];
    if (q < 3 * D) return (double)(gs[q - 2 * D] / sig[q - 2 * D]);
    return (double)gL[q - 3 * D];
  }
};

// MODE 0: price; 1: price and adjoint with checkpoints and recomputation; 2: the forward pass of
// the tape form, which writes X_0 .. X_N to the tape and leaves the payoff, the reverse sweep and
// every partial to sdesim_basket_reverse.
template <typename R, int D, int SEG, int MODE, bool PADDED>
__global__ void __launch_bounds__(256) sdesim_basket(const BasketArgs a) {
  constexpr bool TAPE = MODE == 2;
  constexpr int NQ = BasketConst<D>::NQ;
  const long long S = a.paths;
  const int N = a.steps;
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = s < S;
  const long long sid = live ? s : 0;
  const sde_u32 voff = live ? (sde_u32)(sid * (long long)sizeof(R)) : 0x80000000u;

  R mu[D], sig[D], L[D][D];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    mu[i] = (R)a.mu[i];
    sig[i] = (R)a.sigma[i];
#pragma unroll
    for (int k = 0; k < D; ++k) L[i][k] = (R)a.chol[i * 8 + k];
  }
  const R dt = (R)a.dt, sdt = (R)a.sdt;

  sde_u32 key[8];
  sde_seed_from_u64(a.seed + (sde_u64)sid, key);
  R X[D];
#pragma unroll
  for (int i = 0; i < D; ++i) X[i] = (R)a.x0[i];

  R* ck = (R*)a.states;
  // Forward pass.
  for (int n = 0; n < N; ++n) {
    if (TAPE) sde_tape_store<R, D>(a.states, n, S, voff, X);
    if (MODE == 1 && (n % SEG) == 0 && live) {
      const long long slot = n / SEG;
#pragma unroll
      for (int i = 0; i < D; ++i) ck[(slot * D + i) * S + sid] = X[i];
    }
    BasketStep<R, D> st;
    basket_noise<R, D, PADDED>(key, n, a.assets, L, st);
#pragma unroll
    for (int i = 0; i < D; ++i) X[i] = X[i] + X[i] * (mu[i] * dt + sig[i] * sdt * st.Z[i]);
  }
  if (TAPE) {
    sde_tape_store<R, D>(a.states, N, S, voff, X);
    return;
  }
  R w[D];
#pragma unroll
  for (int i = 0; i < D; ++i) w[i] = (R)a.weight[i];
  R pay, dpay;
  basket_payoff<R, D>(X, w, (R)a.strike, (R)a.width, pay, dpay);

  if (MODE == 0) {
    sde_reduce_store_chunked<NQ>([&](int q) { return !live || q > 1 ? 0.0 : q == 0 ? (double)pay : (double)pay * (double)pay; },
                            a.partials + (long long)blockIdx.x * NQ);
  }
  if (MODE == 1) {
    BasketReverse<R, D> rv(a, X, dpay);
    for (int seg = (N + SEG - 1) / SEG - 1; seg >= 0; --seg) {
      const int n0 = seg * SEG;
      R Xs[SEG][D];
      R Xc[D];
#pragma unroll
      for (int i = 0; i < D; ++i) Xc[i] = ck[((long long)seg * D + i) * S + sid];
      // Recompute the segment's states.
#pragma unroll
      for (int k = 0; k < SEG; ++k) {
#pragma unroll
        for (int i = 0; i < D; ++i) Xs[k][i] = Xc[i];
        if (n0 + k < N && k + 1 < SEG) {
          BasketStep<R, D> st;
          basket_noise<R, D, PADDED>(key, n0 + k, a.assets, L, st);
#pragma unroll
          for (int i = 0; i < D; ++i) Xc[i] = Xc[i] + Xc[i] * (mu[i] * dt + sig[i] * sdt * st.Z[i]);
        }
      }
      // Step back through it.
#pragma unroll
      for (int k = SEG - 1; k >= 0; --k) {
        if (n0 + k < N) rv.step(Xs[k]);
      }
    }
    sde_reduce_store_chunked<NQ>(
        [&](int q) { return !live ? 0.0 : q == 0 ? (double)pay : q == 1 ? (double)pay * (double)pay : rv.grad(q - 2); },
        a.partials + (long long)blockIdx.x * NQ);
  }
}

// Reverse sweep of the tape form: X_N from the tape gives the payoff, then the states X_{N-1} ..
// X_0 are read back in two alternating register buffers, each requested one reverse step before
// its use (the requests past step 0 are clamped to step 0 and unused).
template <typename R, int D>
__global__ void __launch_bounds__(256) sdesim_basket_reverse(const BasketArgs a) {
  constexpr int NQ = BasketConst<D>::NQ;
  const long long S = a.paths;
  const int N = a.steps;
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = s < S;
  const sde_u32 voff = live ? (sde_u32)(s * (long long)sizeof(R)) : 0x80000000u;

  R X[D], w[D];
  sde_tape_load<R, D>(a.states, N, S, voff, X);
#pragma unroll
  for (int i = 0; i < D; ++i) w[i] = (R)a.weight[i];
  R pay, dpay;
  basket_payoff<R, D>(X, w, (R)a.strike, (R)a.width, pay, dpay);

  BasketReverse<R, D> rv(a, X, dpay);
  R xa[D], xb[D];
  sde_tape_load<R, D>(a.states, N - 1, S, voff, xa);
  int n = N - 1;
  for (; n >= 1; n -= 2) {
    sde_tape_load<R, D>(a.states, n - 1, S, voff, xb);
    rv.step(xa);  // step n
    sde_tape_load<R, D>(a.states, n >= 2 ? n - 2 : 0, S, voff, xa);
    rv.step(xb);  // step n - 1
  }
  if (n == 0) rv.step(xa);
  sde_reduce_store_chunked<NQ>(
      [&](int q) { return !live ? 0.0 : q == 0 ? (double)pay : q == 1 ? (double)pay * (double)pay : rv.grad(q - 2); },
      a.partials + (long long)blockIdx.x * NQ);
}

// Sum of per-block partials per quantity (one block per quantity, Neumaier-compensated).
__global__ void __launch_bounds__(256) sdesim_sum_partials(const double* __restrict__ partials, int nblocks, int nq,
                                                           double* __restrict__ out) {
  const int q = blockIdx.x;
  double s = 0.0, c = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
    const double x = partials[(long long)b * nq + q];
    const double t = s + x;
    c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  __shared__ double sh[256], shc[256];
  sh[threadIdx.x] = s;
  shc[threadIdx.x] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    double S = 0.0, C = 0.0;
    for (int i = 0; i < (int)blockDim.x; ++i) {
      for (double x : {sh[i], shc[i]}) {
        const double t = S + x;
        C += (fabs(S) >= fabs(x)) ? (S - t) + x : (x - t) + S;
        S = t;
      }
    }
    out[q] = S + C;
  }
}

}  // namespace

// mode: 0 price, 1 adjoint with checkpoints (a.states holds [slots][D][paths], one per 19 steps),
// 2 adjoint with the tape (a.states holds [steps + 1][D][paths]; a forward and a reverse kernel).
hipError_t sdesim_launch_basket(int f32, int assets, int mode, int grid, const BasketArgs& a, int nq,
                                double* sums, hipStream_t stream) {
  // assets 1 runs the one-asset kernels, 8 the eight-asset ones, and 2-7 the eight-asset ones
  // padded (their own draw stride).
#define SDE_BASKET(R, D, MODE, PAD)                                                                        \
  hipLaunchKernelGGL((sdesim_basket<R, D, 19, MODE, PAD>), dim3(grid), dim3(256), 0, stream, a);              \
  if (MODE == 2) hipLaunchKernelGGL((sdesim_basket_reverse<R, D>), dim3(grid), dim3(256), 0, stream, a);     \
  break;
  bool ok = true;
  const int shape = assets == 1 ? 1 : assets == 8 ? 8 : (assets >= 2 && assets < 8) ? 9 : 0;
  switch ((f32 ? 100 : 0) + shape * 3 + mode) {
    case 100 + 27 + 2: SDE_BASKET(float, 8, 2, true)
    case 100 + 27 + 1: SDE_BASKET(float, 8, 1, true)
    case 100 + 27 + 0: SDE_BASKET(float, 8, 0, true)
    case 100 + 24 + 2: SDE_BASKET(float, 8, 2, false)
    case 100 + 24 + 1: SDE_BASKET(float, 8, 1, false)
    case 100 + 24 + 0: SDE_BASKET(float, 8, 0, false)
    case 100 + 3 + 2: SDE_BASKET(float, 1, 2, false)
    case 100 + 3 + 1: SDE_BASKET(float, 1, 1, false)
    case 100 + 3 + 0: SDE_BASKET(float, 1, 0, false)
    case 27 + 2: SDE_BASKET(double, 8, 2, true)
    case 27 + 1: SDE_BASKET(double, 8, 1, true)
    case 27 + 0: SDE_BASKET(double, 8, 0, true)
    case 24 + 2: SDE_BASKET(double, 8, 2, false)
    case 24 + 1: SDE_BASKET(double, 8, 1, false)
    case 24 + 0: SDE_BASKET(double, 8, 0, false)
    case 3 + 2: SDE_BASKET(double, 1, 2, false)
    case 3 + 1: SDE_BASKET(double, 1, 1, false)
    case 3 + 0: SDE_BASKET(double, 1, 0, false)
    default: ok = false;
  }
#undef SDE_BASKET
  if (!ok) return hipErrorInvalidValue;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(sdesim_sum_partials, dim3(nq), dim3(256), 0, stream, a.partials, grid, nq, sums);
  return hipGetLastError();
}

hipError_t sdesim_launch_sum_partials(const double* partials, int nblocks, int nq, double* out, hipStream_t stream) {
  hipLaunchKernelGGL(sdesim_sum_partials, dim3(nq), dim3(256), 0, stream, partials, nblocks, nq, out);
  return hipGetLastError();
}

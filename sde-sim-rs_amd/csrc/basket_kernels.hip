// basket_kernels.hip -- correlated GBM basket: price and pathwise adjoint (reverse sweep).
//
// Model: D assets, X_{n+1,i} = X_{n,i} + X_{n,i} (mu_i dt + sigma_i sqrt(dt) Z_{n,i}),
// Z_n = L eps_n with L the lower Cholesky factor of the correlation and eps_{n,k} = Psi(u_{n,k})
// (A&S inverse normal, src/proc/increment.rs:160-179) of the pseudo-random draw n*D + k of
// ChaCha8Rng::seed_from_u64(seed + s) -- the Euler step of the reference (src/sim/euler.rs:5-37)
// applied to the DSL form dX_i = (mu_i * X_i) * dt + sum_k (sigma_i * L_ik * X_i) * dW_k.
// Payoff on the weighted basket B = sum_i w_i X_{N,i}: width * softplus((B - K) / width), or the
// plain call max(B - K, 0) when width == 0.
//
// Adjoint (no reference counterpart): pathwise derivatives of the payoff with respect to X0, mu,
// sigma and L, per path, reduced over paths in f64. The forward pass stores the state every SEG
// steps (checkpoints in HBM, [slot][asset][path]); the reverse sweep walks the segments
// backwards, recomputes each segment's states from its checkpoint and then steps back through
// it. A reverse step reads the step's noise off the two states instead of regenerating it:
//   g_n = X_{n+1} / X_n = 1 + mu dt + sigma sqrt(dt) Z_n,  d_n = (X_{n+1} - X_n) - X_n mu dt
// (X_{n+1} - X_n is exact, the two being within a factor 2; forming 1 + mu dt first would round
// mu dt ~ 2e-4 to the spacing of 1 and bias every recovered Z_n)
//   Xbar_n = Xbar_{n+1} g_n,  mubar += Xbar_{n+1} X_n dt,  sigmabar += Xbar_{n+1} d_n / sigma
//   Zbar_n = Xbar_{n+1} X_n sigma sqrt(dt),  Z_n = d_n / (X_n sigma sqrt(dt)),
//   eps_n = L^-1 Z_n (forward substitution),  Lbar_ik += Zbar_{n,i} eps_{n,k}  (k <= i)
#include <hip/hip_runtime.h>

#include "device/sdesim_device.hpp"

namespace {

template <int D>
struct BasketConst {
  static constexpr int NL = D * (D + 1) / 2;   // lower-triangular entries of L
  static constexpr int NQ = 2 + 3 * D + NL;    // reduced quantities: payoff, payoff^2, gradients
};

// The D f64 draws of step n: draws n*D .. n*D + D - 1 of the path's stream.
template <int D>
__device__ __forceinline__ void basket_draws(const sde_u32 (&key)[8], int n, double (&u)[D]) {
  if constexpr (D == 8) {
    sde_u32 blk[16];
    sde_chacha_block<8>(key, (sde_u64)n, blk);
#pragma unroll
    for (int k = 0; k < 8; ++k) u[k] = sde_u01_from_words(blk[2 * k], blk[2 * k + 1]);
  } else {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const long long j = (long long)n * D + k;
      sde_u32 blk[16];
      sde_chacha_block<8>(key, (sde_u64)(j >> 3), blk);
      const int w = (int)(j & 7);
      sde_u32 lo = blk[0], hi = blk[1];
#pragma unroll
      for (int q = 1; q < 8; ++q) {
        lo = (w == q) ? blk[2 * q] : lo;
        hi = (w == q) ? blk[2 * q + 1] : hi;
      }
      u[k] = sde_u01_from_words(lo, hi);
    }
  }
}

template <typename R, int D>
struct BasketStep {
  R eps[D], Z[D];
};

template <typename R, int D>
__device__ __forceinline__ void basket_noise(const sde_u32 (&key)[8], int n, const R (&L)[D][D],
                                             BasketStep<R, D>& st) {
  double u[D];
  basket_draws<D>(key, n, u);
#pragma unroll
  for (int k = 0; k < D; ++k) st.eps[k] = sde_psi<R>(u[k]);
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

template <typename R>
__device__ __forceinline__ R basket_softplus(R x) {
  // log(1 + e^x) without overflow: max(x, 0) + log1p(exp(-|x|))
  return (x > R(0) ? x : R(0)) + (R)log1p(exp(-(double)fabs((double)x)));
}

template <typename R, int D, int SEG, bool ADJ>
__global__ void __launch_bounds__(256) sdesim_basket(const BasketArgs a) {
  constexpr int NL = BasketConst<D>::NL;
  constexpr int NQ = BasketConst<D>::NQ;
  const long long S = a.paths;
  const int N = a.steps;
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = s < S;
  const long long sid = live ? s : 0;

  R mu[D], sig[D], w[D], L[D][D];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    mu[i] = (R)a.mu[i];
    sig[i] = (R)a.sigma[i];
    w[i] = (R)a.weight[i];
#pragma unroll
    for (int k = 0; k < D; ++k) L[i][k] = (R)a.chol[i * 8 + k];
  }
  const R dt = (R)a.dt, sdt = (R)a.sdt;
  const R K = (R)a.strike, width = (R)a.width;

  sde_u32 key[8];
  sde_seed_from_u64(a.seed + (sde_u64)sid, key);
  R X[D];
#pragma unroll
  for (int i = 0; i < D; ++i) X[i] = (R)a.x0[i];

  R* ck = (R*)a.checkpoints;
  // Forward pass.
  for (int n = 0; n < N; ++n) {
    if (ADJ && (n % SEG) == 0 && live) {
      const long long slot = n / SEG;
#pragma unroll
      for (int i = 0; i < D; ++i) ck[(slot * D + i) * S + sid] = X[i];
    }
    BasketStep<R, D> st;
    basket_noise<R, D>(key, n, L, st);
#pragma unroll
    for (int i = 0; i < D; ++i) X[i] = X[i] + X[i] * (mu[i] * dt + sig[i] * sdt * st.Z[i]);
  }
  R B = R(0);
#pragma unroll
  for (int i = 0; i < D; ++i) B = B + w[i] * X[i];
  R pay, dpay;
  if (width > R(0)) {
    const R x = (B - K) / width;
    pay = width * basket_softplus<R>(x);
    dpay = (R)(1.0 / (1.0 + exp(-(double)x)));
  } else {
    pay = B > K ? B - K : R(0);
    dpay = B > K ? R(1) : R(0);
  }

  double acc[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
  if (live) {
    acc[0] = (double)pay;
    acc[1] = (double)pay * (double)pay;
  }

  if (ADJ) {
    R Xb[D], gm[D], gs[D], gL[NL];
#pragma unroll
    for (int i = 0; i < D; ++i) { Xb[i] = dpay * w[i]; gm[i] = R(0); gs[i] = R(0); }
#pragma unroll
    for (int q = 0; q < NL; ++q) gL[q] = R(0);
    // Per-asset constants of the reverse step: mu dt, 1 / (sigma sqrt(dt)), 1 / L_ii.
    R mdt[D], rsd[D], rLd[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
      mdt[i] = mu[i] * dt;
      rsd[i] = R(1) / (sig[i] * sdt);
      rLd[i] = R(1) / L[i][i];
    }
    R Xp[D];  // the state after the step being reversed: X_N first, then each state stepped back to
#pragma unroll
    for (int i = 0; i < D; ++i) Xp[i] = X[i];
    const int nseg = (N + SEG - 1) / SEG;
    for (int seg = nseg - 1; seg >= 0; --seg) {
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
          basket_noise<R, D>(key, n0 + k, L, st);
#pragma unroll
          for (int i = 0; i < D; ++i) Xc[i] = Xc[i] + Xc[i] * (mu[i] * dt + sig[i] * sdt * st.Z[i]);
        }
      }
      // Step back through it.
#pragma unroll
      for (int k = SEG - 1; k >= 0; --k) {
        if (n0 + k < N) {
          R Zb[D], Z[D], eps[D];
#pragma unroll
          for (int i = 0; i < D; ++i) {
            const R xb = Xb[i], x = Xs[k][i], xp = Xp[i];
            const R rx = basket_recip<R>(x);
            const R d = (xp - x) - x * mdt[i];
            gm[i] = gm[i] + xb * x * dt;
            gs[i] = gs[i] + xb * d;
            Zb[i] = xb * x * sig[i] * sdt;
            Z[i] = d * rx * rsd[i];
            Xb[i] = xb * (xp * rx);
            Xp[i] = x;
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
      }
    }
#pragma unroll
    for (int i = 0; i < D; ++i) gs[i] = gs[i] / sig[i];
    if (live) {
#pragma unroll
      for (int i = 0; i < D; ++i) {
        acc[2 + i] = (double)Xb[i];
        acc[2 + D + i] = (double)gm[i];
        acc[2 + 2 * D + i] = (double)gs[i];
      }
#pragma unroll
      for (int q = 0; q < NL; ++q) acc[2 + 3 * D + q] = (double)gL[q];
    }
  }
  sde_block_reduce_store<NQ>(acc, a.partials + (long long)blockIdx.x * NQ);
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

hipError_t sdesim_launch_basket(int f32, int assets, int adjoint, int grid, const BasketArgs& a, int nq,
                                double* sums, hipStream_t stream) {
#define SDE_BASKET(R, D, ADJ)                                                                              \
  hipLaunchKernelGGL((sdesim_basket<R, D, 19, ADJ>), dim3(grid), dim3(256), 0, stream, a);                    \
  break;
  bool ok = true;
  switch ((f32 ? 100 : 0) + assets * 2 + (adjoint ? 1 : 0)) {
    case 100 + 16 + 1: SDE_BASKET(float, 8, true)
    case 100 + 16 + 0: SDE_BASKET(float, 8, false)
    case 100 + 2 + 1: SDE_BASKET(float, 1, true)
    case 100 + 2 + 0: SDE_BASKET(float, 1, false)
    case 16 + 1: SDE_BASKET(double, 8, true)
    case 16 + 0: SDE_BASKET(double, 8, false)
    case 2 + 1: SDE_BASKET(double, 1, true)
    case 2 + 0: SDE_BASKET(double, 1, false)
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

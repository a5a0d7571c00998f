// sdesim_device.hpp -- device code shared by the ahead-of-time library kernels and the kernels
// compiled at run time (hipRTC) from the equation DSL. gfx950 only; wave64.
//
// Contents
//   * ChaCha8 block function, rand_core seed_from_u64 key expansion and rand's f64 conversion:
//     the counter-addressed form of ChaCha8Rng::seed_from_u64(s + seed).random::<f64>()
//     (src/rng/pseudo.rs:14-31, src/sim/mod.rs:65).
//   * Sobol coordinates from byte-sliced XOR tables of the Joe-Kuo direction numbers, addressed
//     by the Gray code of the point index (the sequence the reference's SobolEngine walks,
//     src/rng/sobol.rs:15-25), with the reference's per-path f64 random shift
//     (src/rng/sobol.rs:62-79) or a digital XOR shift shared by all paths.
//   * A&S 26.2.23 inverse normal (src/proc/increment.rs:160-179) in f64, and in f32 with an f64
//     evaluation for |z| > 5; sequential inverse Poisson (src/proc/increment.rs:181-200).
//   * sde_path_kernel, the fused path simulator (arguments in sdesim_args.hpp): one path
//     per lane (or PPL paths per lane), state in registers, uniforms generated in registers,
//     rows of the filtration written step-major / path-contiguous, terminal moments reduced
//     through wave shuffles and LDS into one f64 partial per block.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

#include "sdesim_args.hpp"

// ------------------------------------------------------------------------------------------
// ChaCha8
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ sde_u32 sde_rotl(sde_u32 v, int c) { return (v << c) | (v >> (32 - c)); }

#define SDE_QR(a, b, c, d)                  \
  a += b; d = sde_rotl(d ^ a, 16);          \
  c += d; b = sde_rotl(b ^ c, 12);          \
  a += b; d = sde_rotl(d ^ a, 8);           \
  c += d; b = sde_rotl(b ^ c, 7);

// One ChaCha block with `rounds` rounds: state = constants, key[8], 64-bit block counter,
// 64-bit stream id 0 (rand_chacha's layout). Writes 16 output words.
template <int ROUNDS>
__device__ __forceinline__ void sde_chacha_block(const sde_u32 (&key)[8], sde_u64 counter, sde_u32 (&out)[16]) {
  const sde_u32 c0 = 0x61707865u, c1 = 0x3320646eu, c2 = 0x79622d32u, c3 = 0x6b206574u;
  const sde_u32 n0 = (sde_u32)counter, n1 = (sde_u32)(counter >> 32);
  sde_u32 x0 = c0, x1 = c1, x2 = c2, x3 = c3;
  sde_u32 x4 = key[0], x5 = key[1], x6 = key[2], x7 = key[3];
  sde_u32 x8 = key[4], x9 = key[5], x10 = key[6], x11 = key[7];
  sde_u32 x12 = n0, x13 = n1, x14 = 0u, x15 = 0u;
#pragma unroll
  for (int r = 0; r < ROUNDS; r += 2) {
    SDE_QR(x0, x4, x8, x12) SDE_QR(x1, x5, x9, x13) SDE_QR(x2, x6, x10, x14) SDE_QR(x3, x7, x11, x15)
    SDE_QR(x0, x5, x10, x15) SDE_QR(x1, x6, x11, x12) SDE_QR(x2, x7, x8, x13) SDE_QR(x3, x4, x9, x14)
  }
  out[0] = x0 + c0; out[1] = x1 + c1; out[2] = x2 + c2; out[3] = x3 + c3;
  out[4] = x4 + key[0]; out[5] = x5 + key[1]; out[6] = x6 + key[2]; out[7] = x7 + key[3];
  out[8] = x8 + key[4]; out[9] = x9 + key[5]; out[10] = x10 + key[6]; out[11] = x11 + key[7];
  out[12] = x12 + n0; out[13] = x13 + n1; out[14] = x14; out[15] = x15;
}

// rand_core's seed_from_u64: eight PCG32 outputs become the 256-bit key.
__device__ __forceinline__ void sde_seed_from_u64(sde_u64 state, sde_u32 (&key)[8]) {
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    state = state * 6364136223846793005ULL + 11634580027462260723ULL;
    const sde_u32 xs = (sde_u32)(((state >> 18) ^ state) >> 27);
    const sde_u32 rot = (sde_u32)(state >> 59);
    key[i] = (xs >> rot) | (xs << ((32u - rot) & 31u));
  }
}

// rand's StandardUniform f64 from two consecutive stream words (low word first).
__device__ __forceinline__ double sde_u01_from_words(sde_u32 lo, sde_u32 hi) {
  const sde_u64 w = ((sde_u64)hi << 32) | (sde_u64)lo;
  return (double)(w >> 11) * 0x1.0p-53;
}

// ------------------------------------------------------------------------------------------
// Sobol
// ------------------------------------------------------------------------------------------
// Coordinate `d` of the point whose Gray code is `g`: XOR of one entry per table slice
// (slice 0 covers Gray bits 0-5, slice k >= 1 covers bits 6+8(k-1) .. 13+8(k-1)).
__device__ __forceinline__ sde_u32 sde_sobol_coord(const sde_u32* __restrict__ tab, int nslices, int d, sde_u64 g) {
  const int stride = 64 + 256 * (nslices - 1);
  const sde_u32* t = tab + (long long)d * stride;
  sde_u32 x = t[g & 63u];
#pragma unroll
  for (int k = 1; k < 5; ++k) {
    if (k < nslices) x ^= t[64 + 256 * (k - 1) + ((g >> (6 + 8 * (k - 1))) & 255u)];
  }
  return x;
}

// ------------------------------------------------------------------------------------------
// Inverse CDFs
// ------------------------------------------------------------------------------------------
// A&S 26.2.23, exactly the reference's f64 formula.
__device__ __forceinline__ double sde_inv_normal(double p) {
  const double q = (p < 0.5) ? p : 1.0 - p;
  const double t = sqrt(-2.0 * log(q));
  const double x = t - ((0.010328 * t + 0.802853) * t + 2.515517) /
                           (((0.001308 * t + 0.189269) * t + 1.432788) * t + 1.0);
  return (p < 0.5) ? -x : x;
}

// The same rational approximation in f32. q = min(p, 1 - p) is formed in f64 (exact for a
// 53-bit uniform) so the tails keep their resolution; q < 2.87e-7 (|z| > 5) takes the f64 path.
__device__ __forceinline__ float sde_inv_normal_f32(double p) {
  const double qd = (p < 0.5) ? p : 1.0 - p;
  float x;
  if (qd < 2.87e-7) {
    const double t = sqrt(-2.0 * log(qd));
    x = (float)(t - ((0.010328 * t + 0.802853) * t + 2.515517) /
                        (((0.001308 * t + 0.189269) * t + 1.432788) * t + 1.0));
  } else {
    const float q = (float)qd;
    // -2 ln q = -2 ln(2) log2(q); v_log_f32 is a base-2 logarithm.
    const float t = __builtin_sqrtf(-1.3862943611198906f * __log2f(q));
    const float num = (0.010328f * t + 0.802853f) * t + 2.515517f;
    const float den = ((0.001308f * t + 0.189269f) * t + 1.432788f) * t + 1.0f;
    x = t - num * __builtin_amdgcn_rcpf(den);
  }
  return (p < 0.5) ? -x : x;
}

template <typename R> __device__ __forceinline__ R sde_psi(double p);
template <> __device__ __forceinline__ double sde_psi<double>(double p) { return sde_inv_normal(p); }
template <> __device__ __forceinline__ float sde_psi<float>(double p) { return sde_inv_normal_f32(p); }

// Sequential inverse Poisson CDF capped at k = 200 (f64, as the reference).
__device__ __forceinline__ double sde_inv_poisson(double u, double lambda) {
  if (!(lambda > 0.0)) return 0.0;
  double p = exp(-lambda);
  double f = p;
  int k = 0;
  while (u > f && k < 200) {
    k += 1;
    p *= lambda / (double)k;
    f += p;
  }
  return (double)k;
}

// ------------------------------------------------------------------------------------------
// Expression helpers (fasteval semantics; see expr.hpp)
// ------------------------------------------------------------------------------------------
// Truthy = not within 8 * DBL_EPSILON of zero (NaN is truthy).
__device__ __forceinline__ bool sde_truthy(double a) { return !(fabs(a) <= 8.0 * 2.220446049250313e-16); }
__device__ __forceinline__ bool sde_truthy(float a) { return !((double)fabsf(a) <= 8.0 * 2.220446049250313e-16); }
template <typename R> __device__ __forceinline__ R sde_not(R a) { return sde_truthy(a) ? R(0) : R(1); }
template <typename R> __device__ __forceinline__ R sde_and(R a, R b) { return sde_truthy(a) ? b : a; }
template <typename R> __device__ __forceinline__ R sde_or(R a, R b) { return sde_truthy(a) ? a : b; }
template <typename R> __device__ __forceinline__ R sde_eq(R a, R b) { return sde_truthy(a - b) ? R(0) : R(1); }
template <typename R> __device__ __forceinline__ R sde_ne(R a, R b) { return sde_truthy(a - b) ? R(1) : R(0); }
template <typename R> __device__ __forceinline__ R sde_sq(R a) { return a * a; }

__device__ __forceinline__ double sde_sqrt(double a) { return sqrt(a); }
__device__ __forceinline__ float sde_sqrt(float a) { return __builtin_sqrtf(a); }
__device__ __forceinline__ double sde_pow(double a, double b) { return pow(a, b); }
__device__ __forceinline__ float sde_pow(float a, float b) { return powf(a, b); }
__device__ __forceinline__ double sde_fmod(double a, double b) { return fmod(a, b); }
__device__ __forceinline__ float sde_fmod(float a, float b) { return fmodf(a, b); }
__device__ __forceinline__ double sde_fmin(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ float sde_fmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ double sde_fmax(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float sde_fmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double sde_log(double a) { return log(a); }
__device__ __forceinline__ float sde_log(float a) { return logf(a); }
__device__ __forceinline__ double sde_sign(double a) { return a != a ? a : copysign(1.0, a); }
__device__ __forceinline__ float sde_sign(float a) { return a != a ? a : copysignf(1.0f, a); }

#define SDE_MATH1(NAME, D, F)                                                    \
  __device__ __forceinline__ double sde_##NAME(double a) { return D(a); }        \
  __device__ __forceinline__ float sde_##NAME(float a) { return F(a); }
SDE_MATH1(log10, log10, log10f)
SDE_MATH1(round, round, roundf)
SDE_MATH1(trunc, trunc, truncf)
SDE_MATH1(ceil, ceil, ceilf)
SDE_MATH1(floor, floor, floorf)
SDE_MATH1(fabs, fabs, fabsf)
SDE_MATH1(sin, sin, sinf)
SDE_MATH1(cos, cos, cosf)
SDE_MATH1(tan, tan, tanf)
SDE_MATH1(asin, asin, asinf)
SDE_MATH1(acos, acos, acosf)
SDE_MATH1(atan, atan, atanf)
SDE_MATH1(sinh, sinh, sinhf)
SDE_MATH1(cosh, cosh, coshf)
SDE_MATH1(tanh, tanh, tanhf)
SDE_MATH1(asinh, asinh, asinhf)
SDE_MATH1(acosh, acosh, acoshf)
SDE_MATH1(atanh, atanh, atanhf)

// ------------------------------------------------------------------------------------------
// Block reduction of per-lane f64 pairs into one partial per block (deterministic order).
// ------------------------------------------------------------------------------------------
template <int NV>
__device__ __forceinline__ void sde_block_reduce_store(double (&v)[NV], double* __restrict__ out) {
#pragma unroll
  for (int i = 0; i < NV; ++i) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_xor(v[i], off, 64);
  }
  __shared__ double sde_red[16][NV];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) sde_red[wave][i] = v[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += sde_red[w][i];
      out[i] = s;
    }
  }
}

// ------------------------------------------------------------------------------------------
// The fused path kernel
// ------------------------------------------------------------------------------------------
// A model (emitted from the DSL, or written by hand for the built-ins) provides:
//   typedef real;  P (processes), M (stochastic increments), NK (carried registers, >= 1),
//   HAS_PAYOFF;
//   prologue(X, C, K, t0)                       -- state of the cache and carried values at t0
//   step(X, C, K, tc, tn, dt, sdt, u)           -- one step of one path; X becomes row n+1
//   payoff(X, tT)                               -- terminal payoff (if HAS_PAYOFF)
// Steps are processed in chunks of U = 8 / gcd(M, 8) so that the pseudo-random words of a
// chunk are whole ChaCha blocks held in registers with compile-time word positions.
template <int M>
struct SdeChunk {
  static constexpr int g8 = (M % 8 == 0) ? 8 : (M % 4 == 0) ? 4 : (M % 2 == 0) ? 2 : 1;
  static constexpr int U = (M == 0) ? 1 : 8 / g8;
  static constexpr int BLOCKS = (M == 0) ? 1 : (U * M) / 8;
};

template <class Mdl, int RNG, int OUT, int PPL>
__device__ __forceinline__ void sde_path_kernel(const SimArgs& a) {
  typedef typename Mdl::real real;
  constexpr int P = Mdl::P;
  constexpr int M = Mdl::M;
  constexpr int MM = M > 0 ? M : 1;
  constexpr int NK = Mdl::NK;
  constexpr bool USE_CHACHA = (RNG == SDE_RNG_PSEUDO || RNG == SDE_RNG_SOBOL_SHIFT) && M > 0;
  constexpr int U = USE_CHACHA ? SdeChunk<M>::U : 1;
  constexpr int NB = SdeChunk<M>::BLOCKS;
  constexpr bool F32 = sizeof(real) == 4;
  constexpr int KQ = P + (Mdl::HAS_PAYOFF ? 1 : 0);

  const long long S = a.S;
  const int N = a.N;
  const long long first = (long long)blockIdx.x * ((long long)blockDim.x * PPL) + threadIdx.x;

  real X[PPL][P], C[PPL][P], K[PPL][NK];
  sde_u32 key[PPL][8];
  sde_u64 gray[PPL];
  long long sid[PPL];
  bool live[PPL];

#pragma unroll
  for (int j = 0; j < PPL; ++j) {
    const long long s = first + (long long)j * blockDim.x;
    live[j] = s < S;
    sid[j] = live[j] ? s : 0;
#pragma unroll
    for (int p = 0; p < P; ++p) { X[j][p] = (real)a.x0[p]; C[j][p] = X[j][p]; }
#pragma unroll
    for (int k = 0; k < NK; ++k) K[j][k] = real(0);
    Mdl::prologue(X[j], C[j], K[j], F32 ? (real)a.times_f[0] : (real)a.times[0]);
    if (USE_CHACHA) sde_seed_from_u64(a.seed + (sde_u64)sid[j], key[j]);
    if (RNG == SDE_RNG_SOBOL_SHIFT || RNG == SDE_RNG_SOBOL_XOR) {
      const sde_u64 idx = (sde_u64)(a.sobol_first + sid[j]);
      gray[j] = idx ^ (idx >> 1);
    }
    if ((OUT & SDE_OUT_FILTRATION) && live[j]) {
      real* out = (real*)a.filtration;
#pragma unroll
      for (int p = 0; p < P; ++p) __builtin_nontemporal_store(X[j][p], out + (long long)p * S + sid[j]);
    }
  }

  for (int n0 = 0; n0 < N; n0 += U) {
    sde_u32 W[PPL][USE_CHACHA ? NB * 16 : 1];
    if (USE_CHACHA) {
#pragma unroll
      for (int j = 0; j < PPL; ++j) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          sde_u32 blk[16];
          sde_chacha_block<8>(key[j], (sde_u64)n0 * M / 8 + b, blk);
#pragma unroll
          for (int w = 0; w < 16; ++w) W[j][USE_CHACHA ? 16 * b + w : 0] = blk[w];
        }
      }
    }
#pragma unroll
    for (int i = 0; i < U; ++i) {
      const int n = n0 + i;
      if (U == 1 || n < N) {
        real tc, tn, dt, sdt;
        if (F32) {
          tc = (real)a.times_f[n]; tn = (real)a.times_f[n + 1]; dt = (real)a.dts_f[n]; sdt = (real)a.sdts_f[n];
        } else {
          tc = (real)a.times[n]; tn = (real)a.times[n + 1]; dt = (real)a.dts[n]; sdt = (real)a.sdts[n];
        }
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
          double u[MM];
#pragma unroll
          for (int k = 0; k < MM; ++k) {
            if (M == 0) { u[k] = 0.5; continue; }
            const int d = n * M + k;
            if (RNG == SDE_RNG_PSEUDO) {
              const int w = 2 * (i * M + k);
              u[k] = sde_u01_from_words(W[j][USE_CHACHA ? w : 0], W[j][USE_CHACHA ? w + 1 : 0]);
            } else if (RNG == SDE_RNG_SOBOL_SHIFT) {
              const int w = 2 * (i * M + k);
              const double sh = sde_u01_from_words(W[j][USE_CHACHA ? w : 0], W[j][USE_CHACHA ? w + 1 : 0]);
              const double x = (double)sde_sobol_coord(a.sobol_tab, a.nslices, d, gray[j]) * 0x1.0p-32;
              const double y = x + sh;
              u[k] = y - trunc(y);
            } else if (RNG == SDE_RNG_SOBOL_XOR) {
              const sde_u32 y = sde_sobol_coord(a.sobol_tab, a.nslices, d, gray[j]) ^ a.xor_mask[d];
              u[k] = ((double)y + 0.5) * 0x1.0p-32;
            } else if (RNG == SDE_RNG_INJECTED) {
              u[k] = a.injected[(long long)d * S + sid[j]];
            } else {
              u[k] = 0.5;
            }
          }
          Mdl::step(X[j], C[j], K[j], tc, tn, dt, sdt, u);
        }
        if (OUT & SDE_OUT_FILTRATION) {
          real* out = (real*)a.filtration + (long long)(n + 1) * P * S;
#pragma unroll
          for (int j = 0; j < PPL; ++j) {
            if (live[j]) {
#pragma unroll
              for (int p = 0; p < P; ++p) __builtin_nontemporal_store(X[j][p], out + (long long)p * S + sid[j]);
            }
          }
        }
      }
    }
  }

  if (OUT & SDE_OUT_TERMINAL) {
    double acc[2 * KQ];
#pragma unroll
    for (int q = 0; q < 2 * KQ; ++q) acc[q] = 0.0;
    const real tT = F32 ? (real)a.times_f[N] : (real)a.times[N];
#pragma unroll
    for (int j = 0; j < PPL; ++j) {
      if (live[j]) {
#pragma unroll
        for (int p = 0; p < P; ++p) {
          const double dv = (double)X[j][p] - a.x0[p];
          acc[2 * p] += dv;
          acc[2 * p + 1] += dv * dv;
        }
        if (Mdl::HAS_PAYOFF) {
          const double v = (double)Mdl::payoff(X[j], tT);
          acc[2 * (KQ - 1)] += v;
          acc[2 * (KQ - 1) + 1] += v * v;
        }
      }
    }
    sde_block_reduce_store<2 * KQ>(acc, a.partials + (long long)blockIdx.x * (2 * KQ));
  }
}

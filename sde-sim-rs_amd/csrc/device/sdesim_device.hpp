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
// The table row of a dimension holds one 64-entry slice for Gray-code bits 0-5 and one
// 256-entry slice per further byte (bits 6+8(k-1) .. 13+8(k-1)); coordinate d of the point with
// Gray code g is the XOR of one entry per slice. sde_sobol_offsets gives a path's entry index
// within each slice row (fixed for the whole path).
template <int NSL>
__device__ __forceinline__ void sde_sobol_offsets(sde_u64 g, sde_u32 (&off)[NSL]) {
  off[0] = (sde_u32)(g & 63u);
#pragma unroll
  for (int k = 1; k < NSL; ++k) off[k] = 64u + 256u * (k - 1) + (sde_u32)((g >> (6 + 8 * (k - 1))) & 255u);
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

// f64 evaluation of |z| for a small q = min(p, 1 - p). The empty volatile asm makes the block
// non-speculatable, so callers keep it behind a branch that waves skip when no lane needs it
// (probability ~6e-7 per draw) instead of if-converting it into every step. It is a real call,
// not inlined: the f64 log/sqrt code and its constants then live in the callee (17 VGPRs, no
// stack) instead of pinning registers for the whole step loop of the caller.
// `scale` multiplies the result before it is rounded to f32.
__device__ __attribute__((noinline)) float sde_inv_normal_tail(double q, double scale = 1.0) {
  asm volatile("");
  const double t = sqrt(-2.0 * log(q));
  return (float)(scale * (t - ((0.010328 * t + 0.802853) * t + 2.515517) /
                                  (((0.001308 * t + 0.189269) * t + 1.432788) * t + 1.0)));
}

// The same rational approximation in f32. q = min(p, 1 - p) is formed in f64 (exact for a
// 53-bit uniform) so the tails keep their resolution; q < 2.87e-7 (|z| > 5) takes the f64 path.
__device__ __forceinline__ float sde_inv_normal_f32(double p) {
  const double qd = (p < 0.5) ? p : 1.0 - p;
  float x;
  if (qd < 2.87e-7) {
    x = sde_inv_normal_tail(qd);
  } else {
    const float q = (float)qd;
    // -2 ln q = -2 ln(2) log2(q); v_log_f32 is a base-2 logarithm (q >= 2.87e-7 is normal).
    const float t = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(q));
    const float num = (0.010328f * t + 0.802853f) * t + 2.515517f;
    const float den = ((0.001308f * t + 0.189269f) * t + 1.432788f) * t + 1.0f;
    x = t - num * __builtin_amdgcn_rcpf(den);
  }
  return (p < 0.5) ? -x : x;
}

// The same f32 inverse normal of the rand f64 uniform u = (w >> 11) 2^-53 of the 64-bit word
// w = hi:lo (sde_u01_from_words), formed from the words without f64 arithmetic on the fast path.
// With m the sign extension of bit 31 of hi (set exactly when u >= 1/2),
//   q' = (hi ^ m) 2^-32 + (lo ^ m) 2^-64
// is q = min(u, 1 - u) up to 2^-53 (the 11 low bits that the f64 conversion drops and, for
// u >= 1/2, the complement's offset of one unit of 2^-53): below 2^-31 relative for
// q >= 2.87e-7, inside the rounding of the two f32 conversions. q' < 2.87e-7 (|z| > 5) takes the
// f64 path with the exact q.
__device__ __forceinline__ float sde_inv_normal_f32_words(sde_u32 lo, sde_u32 hi) {
  const sde_u32 m = (sde_u32)((int)hi >> 31);
  const float q = __builtin_fmaf((float)(hi ^ m), 0x1.0p-32f, (float)(lo ^ m) * 0x1.0p-64f);
  float x;
  if (q < 2.87e-7f) {
    const double p = sde_u01_from_words(lo, hi);
    x = sde_inv_normal_tail(p < 0.5 ? p : 1.0 - p);
  } else {
    const float t = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(q));
    const float num = (0.010328f * t + 0.802853f) * t + 2.515517f;
    const float den = ((0.001308f * t + 0.189269f) * t + 1.432788f) * t + 1.0f;
    x = t - num * __builtin_amdgcn_rcpf(den);
  }
  // Negate when u < 1/2 (bit 31 of hi clear): flip the sign bit with the complement of hi's.
  return __builtin_bit_cast(float, __builtin_bit_cast(sde_u32, x) ^ (~hi & 0x80000000u));
}

// Psi of the pseudo-random draw held as its two ChaCha words (lo, hi).
template <typename R> __device__ __forceinline__ R sde_psi_words(sde_u32 lo, sde_u32 hi);
template <> __device__ __forceinline__ double sde_psi_words<double>(sde_u32 lo, sde_u32 hi) {
  return sde_inv_normal(sde_u01_from_words(lo, hi));
}
template <> __device__ __forceinline__ float sde_psi_words<float>(sde_u32 lo, sde_u32 hi) {
  return sde_inv_normal_f32_words(lo, hi);
}

// A pseudo-random uniform kept as its two ChaCha words: u = (hi:lo >> 11) 2^-53. u > 1/2 exactly
// when hi:lo >= 2^63 + 2^11, i.e. hi > 2^31, or hi = 2^31 and lo >= 2^11.
struct SdeUnifW {
  sde_u32 lo, hi;
};
__device__ __forceinline__ double sde_uval(SdeUnifW u) { return sde_u01_from_words(u.lo, u.hi); }
__device__ __forceinline__ bool sde_gt_half(SdeUnifW u) {
  return u.hi > 0x80000000u || (u.hi == 0x80000000u && u.lo >= 2048u);
}

// A uniform from the digital-XOR-scrambled Sobol stream, kept as the complement yc = ~y of its
// 32-bit integer y; its value is (y + 0.5) * 2^-32, strictly inside (0, 1). Bit 31 of yc is set
// exactly when u < 1/2, which is the sign bit of Psi(u) and of sign(u - 1/2) sqrt(dt).
struct SdeUnifX {
  sde_u32 yc;
  const float* tail;  // the f32 increment's table of the 1233 tail values (|z| > 5), or null
  sde_u32 sm, sm2;    // 0x80000000, the sign-bit mask of the f32 selects, in two vector registers
};

__device__ __forceinline__ double sde_uval(double u) { return u; }
__device__ __forceinline__ double sde_uval(SdeUnifX u) { return ((double)(~u.yc) + 0.5) * 0x1.0p-32; }
__device__ __forceinline__ bool sde_gt_half(double u) { return u > 0.5; }
__device__ __forceinline__ bool sde_gt_half(SdeUnifX u) { return (int)u.yc >= 0; }

// q = min(u, 1 - u) of u = (y + 0.5) 2^-32 from yc = ~y, in f32. q is (qi + 0.5) 2^-32 with
// qi = min(y, ~y) = yc ^ (yc >> 31, arithmetic). The word is not folded first: as a signed
// integer, yc + 0.5 is qi + 0.5 when bit 31 is clear (qi = yc) and -(qi + 0.5) when it is set
// (yc - 2^32 = -(~yc + 1)), so q = |(float)(int)yc 2^-32 + 2^-33|, with the absolute value on the
// consumer's source modifier. That is exact for qi < 2^23. Above, the 24-bit conversion rounds
// qi on one half and qi + 1 on the other: q is within one f32 rounding of its value, with a
// bias of half a unit of 2^-32 whose sign differs between the halves (the FP32 call payoff of
// the 2^20-path Heston run moves by 3e-8 relative against FP64). q < 1233 2^-32 holds exactly
// when qi < 1233 (|z| > 5).
__device__ __forceinline__ float sde_q_f32(sde_u32 yc) {
  return __builtin_fabsf(__builtin_fmaf((float)(int)yc, 0x1.0p-32f, 0x1.0p-33f));
}
__device__ __forceinline__ bool sde_q_in_tail(float q) { return q < 1233.0f * 0x1.0p-32f; }
__device__ __forceinline__ sde_u32 sde_q_index(sde_u32 yc) { return yc ^ (sde_u32)((int)yc >> 31); }

// f32 inverse normal of (y + 0.5) 2^-32 from yc = ~y; the tail qi < 1233 (|z| > 5) takes the
// f64 path.
__device__ __forceinline__ float sde_inv_normal_f32_x(sde_u32 yc) {
  const float q = sde_q_f32(yc);
  float x;
  if (sde_q_in_tail(q)) {
    x = sde_inv_normal_tail(((double)sde_q_index(yc) + 0.5) * 0x1.0p-32);
  } else {
    const float t = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(q));
    const float num = (0.010328f * t + 0.802853f) * t + 2.515517f;
    const float den = ((0.001308f * t + 0.189269f) * t + 1.432788f) * t + 1.0f;
    x = t - num * __builtin_amdgcn_rcpf(den);
  }
  // Negate when u < 1/2 (bit 31 of yc set): flip the sign bit with yc's.
  return __builtin_bit_cast(float, __builtin_bit_cast(sde_u32, x) ^ (yc & 0x80000000u));
}

// Wiener increments sqrt(dt) * Psi(u) (src/proc/increment.rs:75-79) and the RK probe's
// sign(u - 1/2) * sqrt(dt) (src/sim/runge_kutta.rs:22-27).
//
// f32 with XOR uniforms: with c = sqrt(2 ln 2) and t = c t', t' = sqrt(-log2 q) (the negation
// rides on the square root's source modifier), the A&S value is z = c x' with
//   x' = t' - num'(t') / den'(t'),  num'(t') = num(c t') / c,  den'(t') = den(c t'),
// and c and the sign of z go into the multiplier: one bit-select on bit 31 of the uniform's word
// (set when u < 1/2) picks +-c sqrt(dt).
#define SDE_PSI_C 1.1774100225154747  // sqrt(2 ln 2)

// v (>= 0) with the sign of u - 1/2 (bit 31 of yc): one 3-input bit select on the uniform's word
// (v_and_or_b32 or v_bitop3_b32). The mask m = 0x80000000 comes from a vector register that the
// compiler cannot see through (VOP3 takes no literal on gfx9), so a wave-uniform v (sqrt(dt) in
// a scalar register) is the select's only scalar operand. The Wiener increment and the RK sign
// of the same uniform use different mask registers (SdeUnifX::sm and sm2): with one, the shared
// (yc & m) would be formed once and each select would become an OR plus that AND.
__device__ __forceinline__ sde_u32 sde_sign_mask() {
  sde_u32 m;
  asm volatile("v_mov_b32 %0, 0x80000000" : "=v"(m));
  return m;
}
__device__ __forceinline__ float sde_sign_select(sde_u32 yc, float v, sde_u32 m) {
  return __builtin_bit_cast(float, (m & yc) | (~m & __builtin_bit_cast(sde_u32, v)));
}

// The tail (qi < 1233, |z| > 5) takes only 1233 values, tabulated by the host in f64 with the
// same formula and rounded to f32: tail[qi] = A&S((qi + 0.5) 2^-32) / c. The branch is rare
// (~6e-7 per draw) and holds one load, so no f64 code (nor its hoisted constants) sits in the
// step loop.
__device__ __forceinline__ float sde_inv_normal_f32_x_scaled(sde_u32 yc, const float* tail) {
  constexpr double c = SDE_PSI_C;
  constexpr float A0 = (float)(2.515517 / c), A1 = 0.802853f, A2 = (float)(0.010328 * c);
  constexpr float B1 = (float)(1.432788 * c), B2 = (float)(0.189269 * c * c), B3 = (float)(0.001308 * c * c * c);
  // The rational is evaluated for every lane (q >= 2^-33 keeps it finite) and the tail lanes then
  // replace it: the fast path stays out of the branch, so the draws of a step share one basic
  // block and can be interleaved.
  const float q = sde_q_f32(yc);
  const float t = __builtin_amdgcn_sqrtf(-__builtin_amdgcn_logf(q));
  const float num = (A2 * t + A1) * t + A0;
  const float den = ((B3 * t + B2) * t + B1) * t + 1.0f;
  float x = t - num * __builtin_amdgcn_rcpf(den);
  if (sde_q_in_tail(q)) {
    asm volatile("");
    x = tail[sde_q_index(yc)];
  }
  return x;
}

template <typename R> __device__ __forceinline__ R sde_psi(double p);
template <> __device__ __forceinline__ double sde_psi<double>(double p) { return sde_inv_normal(p); }
template <> __device__ __forceinline__ float sde_psi<float>(double p) { return sde_inv_normal_f32(p); }
template <typename R> __device__ __forceinline__ R sde_psi(SdeUnifX u);
template <> __device__ __forceinline__ double sde_psi<double>(SdeUnifX u) { return sde_inv_normal(sde_uval(u)); }
template <> __device__ __forceinline__ float sde_psi<float>(SdeUnifX u) { return sde_inv_normal_f32_x(u.yc); }

__device__ __forceinline__ double sde_dw(double u, double sdt) { return sdt * sde_inv_normal(u); }
__device__ __forceinline__ float sde_dw(double u, float sdt) { return sdt * sde_inv_normal_f32(u); }
__device__ __forceinline__ double sde_dw(SdeUnifX u, double sdt) { return sdt * sde_inv_normal(sde_uval(u)); }
__device__ __forceinline__ double sde_dw(SdeUnifW u, double sdt) { return sdt * sde_inv_normal(sde_uval(u)); }
__device__ __forceinline__ float sde_dw(SdeUnifW u, float sdt) { return sdt * sde_inv_normal_f32_words(u.lo, u.hi); }
__device__ __forceinline__ float sde_dw(SdeUnifX u, float sdt) {
  return sde_sign_select(u.yc, sdt * (float)SDE_PSI_C, u.sm) * sde_inv_normal_f32_x_scaled(u.yc, u.tail);
}

template <typename R> __device__ __forceinline__ R sde_signed(double u, R sdt) { return (u > 0.5 ? R(1) : R(-1)) * sdt; }
__device__ __forceinline__ double sde_signed(SdeUnifX u, double sdt) { return (sde_gt_half(u) ? 1.0 : -1.0) * sdt; }
template <typename R> __device__ __forceinline__ R sde_signed(SdeUnifW u, R sdt) { return (sde_gt_half(u) ? R(1) : R(-1)) * sdt; }
__device__ __forceinline__ float sde_signed(SdeUnifX u, float sdt) { return sde_sign_select(u.yc, sdt, u.sm2); }

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
// An addition the compiler may not reassociate (defined outside any reassociating scope): the
// f32 step adds its increment to the state through it, so the state's rounding sequence does
// not depend on how far the step loop is unrolled.
template <typename R> __device__ __forceinline__ R sde_add_strict(R a, R b) { return a + b; }

__device__ __forceinline__ double sde_sqrt(double a) { return sqrt(a); }
__device__ __forceinline__ float sde_sqrt(float a) { return __builtin_amdgcn_sqrtf(a); }  // v_sqrt_f32, 1 ulp
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
// Pseudo-random steps are processed in chunks of U = 8 / gcd(M, 8) so that the words of a chunk
// are whole ChaCha blocks held in registers with compile-time word positions. Sobol and
// injected uniforms, and the time grid, are fetched one step ahead so that their memory latency
// overlaps the current step's arithmetic.
template <int M>
struct SdeChunk {
  static constexpr int g8 = (M % 8 == 0) ? 8 : (M % 4 == 0) ? 4 : (M % 2 == 0) ? 2 : 1;
  static constexpr int U = (M == 0) ? 1 : 8 / g8;
  static constexpr int BLOCKS = (M == 0) ? 1 : (U * M) / 8;
};

typedef sde_u32 SdeU32x2 __attribute__((ext_vector_type(2)));

// The uniform a step consumes: pseudo-random draws (and, in FP32, Sobol points with the random
// shift) as two 32-bit words of a 53-bit fraction, XOR-scrambled Sobol points as their word,
// otherwise an f64 value.
template <int RNG, typename R> struct SdeUnifOf { typedef double type; };
template <typename R> struct SdeUnifOf<SDE_RNG_SOBOL_XOR, R> { typedef SdeUnifX type; };
template <typename R> struct SdeUnifOf<SDE_RNG_PSEUDO, R> { typedef SdeUnifW type; };
template <> struct SdeUnifOf<SDE_RNG_SOBOL_SHIFT, float> { typedef SdeUnifW type; };

template <typename R>
struct SdeGrid {
  R tc, tn, dt, sdt;
};

// Time-grid values of step n. UG: the grid is uniform in the kernel's precision, so dt and
// sqrt(dt) come from the argument block and only the time points the model reads are loaded.
template <typename R, bool UG>
__device__ __forceinline__ SdeGrid<R> sde_grid(const SimArgs& a, int n) {
  SdeGrid<R> g;
  if (sizeof(R) == 4) {
    g.tc = (R)a.times_f[n]; g.tn = (R)a.times_f[n + 1];
    g.dt = UG ? (R)a.dtf_u : (R)a.dts_f[n];
    g.sdt = UG ? (R)a.sdtf_u : (R)a.sdts_f[n];
  } else {
    g.tc = (R)a.times[n]; g.tn = (R)a.times[n + 1];
    g.dt = UG ? (R)a.dt_u : (R)a.dts[n];
    g.sdt = UG ? (R)a.sdt_u : (R)a.sdts[n];
  }
  return g;
}

#ifndef SDE_SOBOL_DEPTH
#define SDE_SOBOL_DEPTH 1
#endif

typedef const __attribute__((address_space(4))) sde_u32 sde_const_u32;
typedef sde_u32 SdeU32x3 __attribute__((ext_vector_type(3)));
typedef sde_u32 SdeU32x4 __attribute__((ext_vector_type(4)));

// W consecutive words at (voff + soff) of a buffer resource, in loads of up to 4 words.
template <int W>
__device__ __forceinline__ void sde_load_words(__amdgpu_buffer_rsrc_t r, int voff, int soff, sde_u32* out) {
  if constexpr (W >= 4) {
    const SdeU32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
    out[0] = v.x; out[1] = v.y; out[2] = v.z; out[3] = v.w;
    sde_load_words<W - 4>(r, voff + 16, soff, out + 4);
  } else if constexpr (W == 3) {
    const SdeU32x3 v = __builtin_amdgcn_raw_buffer_load_b96(r, voff, soff, 0);
    out[0] = v.x; out[1] = v.y; out[2] = v.z;
  } else if constexpr (W == 2) {
    const SdeU32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
    out[0] = v.x; out[1] = v.y;
  } else if constexpr (W == 1) {
    out[0] = (sde_u32)__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
  }
}

// AL (XOR mode): the Sobol index of path 0 is a multiple of 64 and a block holds whole waves,
// so the 64 paths of a wave are one aligned run of 64 consecutive point indices; their Gray
// codes agree above bit 5 and every table slice but the first gives the same entry to the
// whole wave. The host then lays the tables out for that access: slice 0 as [N][64][M] (a
// lane's M words of a step are one vector load) and slices 1.. as [slice][entry][N * M] (a
// wave's words of a step are consecutive, read with scalar loads).
template <class Mdl, int RNG, int OUT, int PPL, int NSL, bool UG, bool AL = false>
__device__ __forceinline__ void sde_path_kernel(const SimArgs& a) {
  typedef typename Mdl::real real;
  typedef typename SdeUnifOf<RNG, real>::type unif;
  constexpr int P = Mdl::P;
  constexpr int M = Mdl::M;
  constexpr int MM = M > 0 ? M : 1;
  constexpr int NK = Mdl::NK;
  constexpr bool USE_CHACHA = (RNG == SDE_RNG_PSEUDO || RNG == SDE_RNG_SOBOL_SHIFT) && M > 0;
  constexpr bool USE_SOBOL = (RNG == SDE_RNG_SOBOL_SHIFT || RNG == SDE_RNG_SOBOL_XOR) && M > 0;
  constexpr bool USE_INJ = RNG == SDE_RNG_INJECTED && M > 0;
  // Table-driven uniforms are requested DEPTH steps ahead (DEPTH buffers, so the loop is
  // unrolled by DEPTH and each buffer index is a constant).
  constexpr int DEPTH = USE_SOBOL ? SDE_SOBOL_DEPTH : 1;
  constexpr int U = USE_CHACHA ? SdeChunk<M>::U : DEPTH;
  constexpr int NB = SdeChunk<M>::BLOCKS;
  constexpr int KQ = P + (Mdl::HAS_PAYOFF ? 1 : 0);
  constexpr bool FILT = (OUT & SDE_OUT_FILTRATION) != 0;

  const long long S = a.S;
  const int N = a.N;
  const long long first = (long long)blockIdx.x * ((long long)blockDim.x * PPL) + threadIdx.x;
  const sde_u32 stride = 64u + 256u * (NSL - 1);

  real X[PPL][P], C[PPL][P], K[PPL][NK];
  sde_u32 key[PPL][8];
  sde_u32 soff[PPL][NSL];  // byte offsets of this path's entries within a dimension's table row
                           // (with AL: soff[j][0] is the byte offset of its words in a step block)
  sde_const_u32* urow[PPL][NSL];  // with AL: the wave's row of slice sl >= 1 in sobol_up
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
    Mdl::prologue(X[j], C[j], K[j], sizeof(real) == 4 ? (real)a.times_f[0] : (real)a.times[0]);
    if (USE_CHACHA) sde_seed_from_u64(a.seed + (sde_u64)sid[j], key[j]);
    if (USE_SOBOL) {
      const sde_u64 idx = (sde_u64)(a.sobol_first + sid[j]);
      const sde_u64 g = idx ^ (idx >> 1);
      if (AL) {
        soff[j][0] = (sde_u32)(g & 63u) * (sde_u32)(MM * 4);
#pragma unroll
        for (int sl = 1; sl < NSL; ++sl) {
          // Lane 0 has the wave's smallest path index: if it is past the last path, so is every
          // lane of the wave, whose values are then never used.
          const sde_u32 e = __builtin_amdgcn_readfirstlane((sde_u32)(g >> (6 + 8 * (sl - 1))) & 255u);
          urow[j][sl] = (sde_const_u32*)a.sobol_up + ((size_t)(sl - 1) * 256u + e) * (size_t)a.sobol_dims;
        }
      } else {
        sde_sobol_offsets<NSL>(g, soff[j]);
#pragma unroll
        for (int sl = 0; sl < NSL; ++sl) soff[j][sl] *= 4u;
      }
    }
  }
  // Filtration rows are written through a buffer resource rebased on each row (scalar address
  // arithmetic only): the path's byte offset is the 32-bit vector offset and the process's
  // offset within the row the scalar offset; a row of P * S values must stay below 2 GiB. The
  // stores carry the non-temporal cache policy (aux = 2): the filtration streams past L2, which
  // keeps the Sobol tables resident (with the default policy the kernel runs ~1.5x slower).
  // Lanes past the last path store at vector offset 2^31, which is beyond the row resource's
  // size whether or not the scalar offset is counted, so the hardware drops those stores; the
  // stores are therefore unconditional and leave no exec-masked branch behind that would make
  // the next wait for table loads also wait for this step's stores.
  const long long row_bytes = (long long)P * S * (long long)sizeof(real);
  sde_u32 voff[PPL];
#pragma unroll
  for (int j = 0; j < PPL; ++j) voff[j] = live[j] ? (sde_u32)(sid[j] * (long long)sizeof(real)) : 0x80000000u;
  auto store_row = [&](int r) {
    char* base = (char*)a.filtration + (long long)r * row_bytes;
    __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)row_bytes, 0x00020000);
#pragma unroll
    for (int j = 0; j < PPL; ++j) {
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const int soff_p = (int)((long long)p * S * (long long)sizeof(real));
        if constexpr (sizeof(real) == 4)
          __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(sde_u32, (float)X[j][p]), rs, (int)voff[j], soff_p, 2);
        else
          __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(SdeU32x2, (double)X[j][p]), rs, (int)voff[j],
                                                soff_p, 2);
      }
    }
  };
  if (FILT) store_row(0);

  // Sobol table entries are read through a buffer resource: the uniform row offset goes in the
  // scalar offset and the path's fixed entry offsets in the 32-bit vector offset, so a gather
  // costs no address arithmetic (and reads outside the table would return 0, not fault).
  __amdgpu_buffer_rsrc_t tab_rsrc = __builtin_amdgcn_make_buffer_rsrc(
      (void*)a.sobol_tab, (short)0, USE_SOBOL ? a.sobol_tab_words * 4 : 0, 0x00020000);
  // With AL the wave-uniform words are read with scalar loads (through the constant address
  // space; indices stay below N * M), which leaves one vector load per step and keeps these
  // reads off the vector memory counter that the filtration stores also occupy.

  // Software pipeline for table-driven uniforms: the words of step n+DEPTH are requested before
  // step n is computed, and those of step n+1 are combined into `nx` after it.
  sde_u32 sraw[DEPTH][PPL][MM][NSL];
  sde_u32 nx[PPL][MM];
  double ni[PPL][MM];
  auto request = [&](int n, int b) {
    if (USE_SOBOL && AL) {
#pragma unroll
      for (int j = 0; j < PPL; ++j) {
        sde_u32 w[MM];
        sde_load_words<MM>(tab_rsrc, (int)soff[j][0], n * (64 * MM * 4), w);
#pragma unroll
        for (int k = 0; k < MM; ++k) {
          sraw[b][j][k][0] = w[k];
#pragma unroll
          for (int sl = 1; sl < NSL; ++sl) sraw[b][j][k][sl] = urow[j][sl][n * M + k];
        }
      }
      return;
    }
#pragma unroll
    for (int k = 0; k < MM; ++k) {
      const int d = n * M + k;
      if (USE_SOBOL) {
        const int base = (int)((sde_u32)d * stride * 4u);
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
#pragma unroll
          for (int sl = 0; sl < NSL; ++sl)
            sraw[b][j][k][sl] = (sde_u32)__builtin_amdgcn_raw_buffer_load_b32(tab_rsrc, (int)soff[j][sl], base, 0);
        }
      }
      if (USE_INJ) {
#pragma unroll
        for (int j = 0; j < PPL; ++j) ni[j][k] = a.injected[(long long)d * S + sid[j]];
      }
    }
  };
  // XOR mode needs no mask here: the host folds the complement of dimension d's mask into the
  // 64 slice-0 entries of its table row, so the XOR of the slices is already the complement of
  // the scrambled point (and no per-step vector load of the mask forces a wait on this step's
  // outstanding filtration stores).
  auto combine = [&](int b) {
#pragma unroll
    for (int k = 0; k < MM; ++k) {
      if (USE_SOBOL) {
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
          sde_u32 x = 0u;
#pragma unroll
          for (int sl = 0; sl < NSL; ++sl) x ^= sraw[b][j][k][sl];
          nx[j][k] = x;
        }
      }
    }
  };
  if (N > 0) {
#pragma unroll
    for (int b = 0; b < DEPTH; ++b) request(b < N ? b : N - 1, b);
    combine(0);
  }
  SdeGrid<real> g{};
  if (N > 0) g = sde_grid<real, UG>(a, 0);
  const bool sel_masks = RNG == SDE_RNG_SOBOL_XOR && sizeof(real) == 4;
  const sde_u32 smask = sel_masks ? sde_sign_mask() : 0x80000000u;
  const sde_u32 smask2 = sel_masks ? sde_sign_mask() : 0x80000000u;

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
        unif u[PPL][MM];
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
#pragma unroll
          for (int k = 0; k < MM; ++k) {
            double v = 0.5;
            if (RNG == SDE_RNG_SOBOL_SHIFT && M > 0) {
              const int w = 2 * (i * M + k);
              const double sh = sde_u01_from_words(W[j][USE_CHACHA ? w : 0], W[j][USE_CHACHA ? w + 1 : 0]);
              const double y = (double)nx[j][k] * 0x1.0p-32 + sh;
              v = y - trunc(y);
            } else if (USE_INJ) {
              v = ni[j][k];
            }
            if constexpr (RNG == SDE_RNG_SOBOL_XOR) {
              u[j][k].yc = nx[j][k];
              u[j][k].tail = a.tail_tab;
              u[j][k].sm = smask;
              u[j][k].sm2 = smask2;
            } else if constexpr (RNG == SDE_RNG_PSEUDO) {
              const int w = 2 * (i * M + k);
              u[j][k].lo = W[j][USE_CHACHA ? w : 0];
              u[j][k].hi = W[j][USE_CHACHA ? w + 1 : 0];
            } else if constexpr (RNG == SDE_RNG_SOBOL_SHIFT && sizeof(real) == 4) {
              // frac(y 2^-32 + shift) as an exact 53-bit fraction: (y 2^21 + (word >> 11)) mod
              // 2^53, held as the 64-bit word (fraction << 11). The f64 path rounds the sum to
              // 53 bits first when it exceeds 1, a difference of at most 2^-53 in the uniform.
              const int w = 2 * (i * M + k);
              const sde_u64 sh = (((sde_u64)W[j][USE_CHACHA ? w + 1 : 0] << 32) | W[j][USE_CHACHA ? w : 0]) >> 11;
              const sde_u64 f = (((sde_u64)nx[j][k] << 21) + sh) << 11;
              u[j][k].lo = (sde_u32)f;
              u[j][k].hi = (sde_u32)(f >> 32);
            } else {
              u[j][k] = v;
            }
          }
        }
        // The grid and table words ahead are requested unconditionally (near the end the last
        // step's are requested again, in bounds and unused), so no branch separates the loads
        // from the stores below and the wait before `combine` covers the loads only.
        const SdeGrid<real> gc = g;
        g = sde_grid<real, UG>(a, n + 1 < N ? n + 1 : n);
        request(n + DEPTH < N ? n + DEPTH : N - 1, i % DEPTH);
#pragma unroll
        for (int j = 0; j < PPL; ++j) Mdl::step(X[j], C[j], K[j], gc.tc, gc.tn, gc.dt, gc.sdt, u[j]);
        if (FILT) store_row(n + 1);
        combine((i + 1) % DEPTH);
      }
    }
  }

  if (OUT & SDE_OUT_TERMINAL) {
    double acc[2 * KQ];
#pragma unroll
    for (int q = 0; q < 2 * KQ; ++q) acc[q] = 0.0;
    const real tT = sizeof(real) == 4 ? (real)a.times_f[N] : (real)a.times[N];
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

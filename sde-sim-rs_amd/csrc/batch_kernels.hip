// batch_kernels.hip -- many independent 1-D SDE instances of uneven size in one launch.
//
// Built-in models with run-time parameters (their DSL forms, compiled by the generic path, give
// bit-identical f64 paths):
//   GBM  dX = (mu * X) * dt + (sigma * X) * dW1
//   OU   dX = (kappa * (theta - X)) * dt + (sigma) * dW1
//   CIR  dX = (kappa * (theta - max(X, 0))) * dt + (sigma * max(X, 0)^0.5) * dW1
// Euler-Maruyama (src/sim/euler.rs:5-37) on a uniform grid dt = T / N, pseudo-random uniforms
// from ChaCha8Rng::seed_from_u64(seed_i + s) exactly as the generic kernel draws them.
//
// Scheduling: a persistent grid of waves pulls work items (instance, 64 * PPL paths) from an
// atomic counter. Items are queued longest-first (LPT on steps x paths), each wave reduces its
// item's terminal values to an f64 (sum, sum of squares) pair written to the item's fixed slot,
// and a finalize kernel reduces each instance's slots with Neumaier summation, so results do
// not depend on which wave ran which item. The counter is reset by a kernel (no memset), so the
// three launches can be captured into one hipGraph and replayed.
#include <hip/hip_runtime.h>

#include "device/sdesim_device.hpp"

namespace {

// One Euler step of a built-in model. FP64 adds the terms to the state one at a time, in the
// reference's order (src/sim/euler.rs:20-27: the drift term, then the Wiener term), which keeps it
// bitwise equal to the DSL form's kernel. FP32 sums the step's terms first and adds the sum to the
// state once, as the generated FP32 kernels do: added alone, a drift of a few ulps of x per step
// (GBM with mu dt ~ 3e-7) is rounded to whole ulps at every step in the same direction, which
// biased a 1000-step GBM mean by 1.5e-5.
template <typename R>
__device__ __forceinline__ R sde_euler_builtin(int kind, R x, R p0, R p1, R p2, R dt, R sdt, R z) {
  R v0, v1;
  if (kind == 0) {
    v0 = p0 * x;
    v1 = p1 * x;
  } else if (kind == 1) {
    v0 = p0 * (p1 + (-x));
    v1 = p2;
  } else {
    v0 = p0 * (p1 + (-sde_fmax(x, R(0))));
    v1 = p2 * sde_sqrt(sde_fmax(x, R(0)));
  }
  if constexpr (sizeof(R) == 4) {
    return x + (v0 * dt + v1 * (sdt * z));
  } else {
    R a0 = x;
    a0 = a0 + v0 * dt;
    a0 = a0 + v1 * (sdt * z);
    return a0;
  }
}

template <typename R, int PPL>
__global__ void __launch_bounds__(256) sdesim_batch(const SdeInstance* __restrict__ inst,
                                                     const SdeItem* __restrict__ items, int n_items,
                                                     unsigned int* __restrict__ counter,
                                                     double* __restrict__ partials, R* __restrict__ terminal,
                                                     const long long* __restrict__ term_offset) {
  const int lane = threadIdx.x & 63;
  // A wave can never process more than n_items items, which bounds the loop whatever happens.
  for (int iter = 0; iter < n_items; ++iter) {
    // Claim the next item without a divergent branch: every lane adds (lane == 0) and lane 0's
    // result, the counter before this wave's claim, is made wave-uniform by readfirstlane (all
    // lanes are active at the loop head). The exit test is then a scalar branch of the whole wave.
    const unsigned int claim = atomicAdd(counter, lane == 0 ? 1u : 0u);
    const int k = (int)__builtin_amdgcn_readfirstlane(claim);
    if (k >= n_items) break;
    const SdeItem it = items[k];
    const SdeInstance I = inst[it.instance];
    const R p0 = (R)I.p0, p1 = (R)I.p1, p2 = (R)I.p2;
    const R dt = sizeof(R) == 4 ? (R)I.dtf : (R)I.dt;
    const R sdt = sizeof(R) == 4 ? (R)I.sdtf : (R)I.sdt;
    const int N = I.steps;
    const int kind = I.kind;
    R x[PPL];
    sde_u32 key[PPL][8];
    long long sid[PPL];
    bool live[PPL];
#pragma unroll
    for (int j = 0; j < PPL; ++j) {
      const long long s = (long long)it.chunk * 64 * PPL + j * 64 + lane;
      live[j] = s < I.paths;
      sid[j] = live[j] ? s : 0;
      x[j] = (R)I.x0;
      sde_seed_from_u64(I.seed + (sde_u64)sid[j], key[j]);
    }
    for (int n0 = 0; n0 < N; n0 += 8) {
      sde_u32 W[PPL][16];
#pragma unroll
      for (int j = 0; j < PPL; ++j) {
        sde_u32 blk[16];
        sde_chacha_block<8>(key[j], (sde_u64)(n0 / 8), blk);
#pragma unroll
        for (int w = 0; w < 16; ++w) W[j][w] = blk[w];
      }
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        if (n0 + i < N) {
#pragma unroll
          for (int j = 0; j < PPL; ++j) {
            const R z = sde_psi_words<R>(W[j][2 * i], W[j][2 * i + 1]);
            x[j] = sde_euler_builtin<R>(kind, x[j], p0, p1, p2, dt, sdt, z);
          }
        }
      }
    }
    double acc[2] = {0.0, 0.0};
#pragma unroll
    for (int j = 0; j < PPL; ++j) {
      if (live[j]) {
        const double d = (double)x[j] - I.x0;
        acc[0] += d;
        acc[1] += d * d;
        if (terminal) terminal[term_offset[it.instance] + sid[j]] = x[j];
      }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) acc[i] += __shfl_xor(acc[i], off, 64);
    }
    if (lane == 0) {
      partials[2 * (long long)it.slot] = acc[0];
      partials[2 * (long long)it.slot + 1] = acc[1];
    }
  }
}

__global__ void sdesim_batch_reset(unsigned int* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] = 0u;
}

struct Neumaier {
  double s = 0.0, c = 0.0;
  __device__ void add(double x) {
    const double t = s + x;
    c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
    s = t;
  }
};

// One thread per instance: its item slots are [slot0[i], slot0[i+1]).
__global__ void sdesim_batch_finalize(const SdeInstance* __restrict__ inst, const int* __restrict__ slot0, int n,
                                      const double* __restrict__ partials, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Neumaier a, b;
  for (int s = slot0[i]; s < slot0[i + 1]; ++s) {
    a.add(partials[2 * (long long)s]);
    b.add(partials[2 * (long long)s + 1]);
  }
  const double S = (double)inst[i].paths;
  const double s1 = a.s + a.c, s2 = b.s + b.c;
  const double m1 = s1 / S;
  out[2 * i] = inst[i].x0 + m1;
  out[2 * i + 1] = S > 1.0 ? (s2 - s1 * m1) / (S - 1.0) : 0.0;
}

}  // namespace

// Host launchers (batch runtime in runtime_batch.cpp).
hipError_t sdesim_launch_batch(int f32, int ppl, int grid, int block, const SdeInstance* inst, const SdeItem* items,
                               int n_items, unsigned int* counter, double* partials, void* terminal,
                               const long long* term_offset, hipStream_t stream) {
  hipLaunchKernelGGL(sdesim_batch_reset, dim3(1), dim3(64), 0, stream, counter);
#define SDE_BATCH_CASE(R, P)                                                                            \
  if (ppl == P) {                                                                                        \
    hipLaunchKernelGGL((sdesim_batch<R, P>), dim3(grid), dim3(block), 0, stream, inst, items, n_items,     \
                       counter, partials, (R*)terminal, term_offset);                                   \
    return hipGetLastError();                                                                            \
  }
  if (f32) {
    SDE_BATCH_CASE(float, 1)
    SDE_BATCH_CASE(float, 2)
    SDE_BATCH_CASE(float, 4)
  } else {
    SDE_BATCH_CASE(double, 1)
    SDE_BATCH_CASE(double, 2)
    SDE_BATCH_CASE(double, 4)
  }
#undef SDE_BATCH_CASE
  return hipErrorInvalidValue;
}

hipError_t sdesim_launch_batch_finalize(const SdeInstance* inst, const int* slot0, int n, const double* partials,
                                        double* out, hipStream_t stream) {
  hipLaunchKernelGGL(sdesim_batch_finalize, dim3((n + 255) / 256), dim3(256), 0, stream, inst, slot0, n, partials,
                     out);
  return hipGetLastError();
}

// Resident blocks per CU for a (precision, PPL, block) variant: sizes the persistent grid.
hipError_t sdesim_batch_occupancy(int f32, int ppl, int block, int* blocks_per_cu) {
#define SDE_OCC_CASE(R, P)                                                                                   \
  if (ppl == P) return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, sdesim_batch<R, P>, block, 0);
  if (f32) {
    SDE_OCC_CASE(float, 1)
    SDE_OCC_CASE(float, 2)
    SDE_OCC_CASE(float, 4)
  } else {
    SDE_OCC_CASE(double, 1)
    SDE_OCC_CASE(double, 2)
    SDE_OCC_CASE(double, 4)
  }
#undef SDE_OCC_CASE
  return hipErrorInvalidValue;
}

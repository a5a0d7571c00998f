// adjoint.hpp -- pieces shared by the adjoint kernels (basket_kernels.hip, lmm_kernels.hip): the
// state tape of the reverse sweeps and a chunked per-block reduction of the per-path results.
//
// Tape layout [step][row][path]: row r of step n holds one state variable of every path. The ROWS rows
// of a step sit under one buffer resource of ROWS * S * sizeof(R) bytes (below 2^31, checked by
// the host); the path's byte offset is the vector offset and the row's the scalar offset. Lanes
// past the last path pass vector offset 2^31, beyond the resource: their stores are dropped and
// their loads return 0, so no exec-masked branch surrounds the accesses. Both directions use the
// non-temporal policy: each value is written once and read once.
#pragma once

#include "device/sdesim_device.hpp"

// n is wave-uniform; readfirstlane keeps the resource in scalar registers even where the index
// was formed by a vector instruction (a clamped subtraction, say), which would otherwise make
// the compiler wrap every access in a loop over distinct resources.
template <typename R, int ROWS>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t sde_tape_rsrc(void* tape, int n, long long S) {
  n = __builtin_amdgcn_readfirstlane(n);
  const long long step_bytes = (long long)ROWS * S * (long long)sizeof(R);
  return __builtin_amdgcn_make_buffer_rsrc((void*)((char*)tape + (long long)n * step_bytes), (short)0, (int)step_bytes,
                                           0x00020000);
}

template <typename R>
__device__ __forceinline__ void sde_tape_put(__amdgpu_buffer_rsrc_t rs, sde_u32 voff, int row, long long S, R v) {
  const int soff = (int)((long long)row * S * (long long)sizeof(R));
  if constexpr (sizeof(R) == 4)
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(sde_u32, (float)v), rs, (int)voff, soff, 2);
  else
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(SdeU32x2, (double)v), rs, (int)voff, soff, 2);
}

template <typename R>
__device__ __forceinline__ R sde_tape_get(__amdgpu_buffer_rsrc_t rs, sde_u32 voff, int row, long long S) {
  const int soff = (int)((long long)row * S * (long long)sizeof(R));
  if constexpr (sizeof(R) == 4)
    return __builtin_bit_cast(float, (sde_u32)__builtin_amdgcn_raw_buffer_load_b32(rs, (int)voff, soff, 2));
  else
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, (int)voff, soff, 2));
}

template <typename R, int ROWS>
__device__ __forceinline__ void sde_tape_store(void* tape, int n, long long S, sde_u32 voff, const R (&X)[ROWS]) {
  const __amdgpu_buffer_rsrc_t rs = sde_tape_rsrc<R, ROWS>(tape, n, S);
#pragma unroll
  for (int i = 0; i < ROWS; ++i) sde_tape_put<R>(rs, voff, i, S, X[i]);
}

template <typename R, int ROWS>
__device__ __forceinline__ void sde_tape_load(void* tape, int n, long long S, sde_u32 voff, R (&X)[ROWS]) {
  const __amdgpu_buffer_rsrc_t rs = sde_tape_rsrc<R, ROWS>(tape, n, S);
#pragma unroll
  for (int i = 0; i < ROWS; ++i) X[i] = sde_tape_get<R>(rs, voff, i, S);
}

// One partial per block of each of the NQ per-path values get(q): a butterfly over the wave's
// lanes, then the waves in order (the summation order of sde_block_reduce_store), eight
// quantities at a time so that only eight f64 values per lane are live.
template <int NQ, class G>
__device__ __forceinline__ void sde_reduce_store_chunked(const G& get, double* __restrict__ out) {
  __shared__ double red[16][8];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int q0 = 0; q0 < NQ; q0 += 8) {
    double v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = q0 + i < NQ ? get(q0 + i) : 0.0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_xor(v[i], off, 64);
    }
    if (lane == 0) {
#pragma unroll
      for (int i = 0; i < 8; ++i) red[wave][i] = v[i];
    }
    __syncthreads();
    if ((int)threadIdx.x < 8 && q0 + (int)threadIdx.x < NQ) {
      double t = 0.0;
      for (int w = 0; w < nw; ++w) t += red[w][threadIdx.x];
      out[q0 + threadIdx.x] = t;
    }
    __syncthreads();
  }
}

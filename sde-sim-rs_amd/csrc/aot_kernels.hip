// aot_kernels.hip -- kernels compiled ahead of time into libsdesim.so (gfx950).
//   * sdesim_finalize: cross-block reduction of the per-block (sum, sum of squares) partials of
//     the fused kernel into terminal mean and sample variance; Neumaier-compensated f64.
//   * sdesim_to_long: [T*P][S] step-major filtration (f32 or f64) -> [S][T*P] f64, the value
//     column of the reference's long layout (src/filtration.rs:81-116), through LDS tiles so that
//     both the read and the write are coalesced.
//   * scratch: the device array behind the allocation-free terminal run (sdesim_plan_run_scratch).
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

#include "../../include/sdesim.h"

// Per-block terminal partials of sdesim_plan_run_scratch, which may not allocate: a fixed array in
// the library's code object, SDESIM_SCRATCH_BYTES per device (each device that runs the library
// holds its own copy until the process ends). Host code never passes `scratch` itself to a kernel:
// the name is a host-side stand-in; sdesim_scratch_address gives this device's copy.
static __device__ __align__(256) double scratch[SDESIM_SCRATCH_BYTES / sizeof(double)];

namespace {

struct Neumaier {
  double s = 0.0, c = 0.0;
  __device__ void add(double x) {
    const double t = s + x;
    c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  __device__ double value() const { return s + c; }
};

// One block per reduced quantity q; 256 threads.
__global__ void __launch_bounds__(256) sdesim_finalize(const double* __restrict__ partials, int nblocks, int kq,
                                                       long long S, const double* __restrict__ shifts,
                                                       double* __restrict__ out) {
  const int q = blockIdx.x;
  Neumaier s1, s2;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
    s1.add(partials[(long long)b * 2 * kq + 2 * q]);
    s2.add(partials[(long long)b * 2 * kq + 2 * q + 1]);
  }
  __shared__ double a1[256], c1[256], a2[256], c2[256];
  a1[threadIdx.x] = s1.s; c1[threadIdx.x] = s1.c;
  a2[threadIdx.x] = s2.s; c2[threadIdx.x] = s2.c;
  __syncthreads();
  if (threadIdx.x == 0) {
    Neumaier t1, t2;
    for (int i = 0; i < (int)blockDim.x; ++i) { t1.add(a1[i]); t1.add(c1[i]); t2.add(a2[i]); t2.add(c2[i]); }
    const double n = (double)S;
    const double m1 = t1.value() / n;
    out[2 * q] = shifts[q] + m1;
    out[2 * q + 1] = S > 1 ? (t2.value() - t1.value() * m1) / (n - 1.0) : 0.0;
  }
}

// Long layout: [rows][S] (rows = T * P) to [S][rows] f64. A block moves a tile of 256 rows x W
// scenarios through LDS (W = 32 for f32 input, 16 for f64: 128 bytes of one row per 8 lanes):
//  * reads: each lane loads 16 bytes (4 f32 or 2 f64 consecutive scenarios), a whole 128-byte
//    line of a row per 8 lanes, with the non-temporal policy (each input byte is read once);
//  * writes: for each scenario, the tile's 256 rows as f64 (2 KB contiguous), each lane storing
//    two adjacent rows in one 16-byte non-temporal store (the output is written once and not
//    read back by this kernel);
//  * the LDS row pitch is W + 1 elements.
// Tiles are numbered with the row tile fastest (tile b covers row tile b % ny of scenario group
// b / ny), so the blocks resident together fill whole output rows of a few scenario groups instead
// of 2 KB pieces of every scenario. One block per tile: blocks that loop over several tiles were
// 7 % slower.
// Among the layouts measured on 2050 x 2^20 values (64 x 128 up to 512 x 16 tiles, 8- or 16-byte
// stores, default or non-temporal policy, scenario- or row-tile-fastest order, one tile per block
// or a block walking all row tiles of its scenarios) this one moved the most bytes per second.
// Tiles that are not whole (the last rows or scenarios), or data whose rows are not 16-byte aligned
// (S not a multiple of the vector width, an odd row count), take element-wise guarded accesses.
template <typename T>
__global__ void __launch_bounds__(256) sdesim_to_long(const T* __restrict__ in, long long S, int rows, int ny,
                                                      double* __restrict__ out) {
  constexpr int VEC = 16 / (int)sizeof(T);
  constexpr int W = 8 * VEC;
  constexpr int RT = 256;
  __shared__ T tile[RT][W + 1];
  const int t = threadIdx.x;
  const long long b = blockIdx.x;
  const long long s0 = (b / ny) * W;
  const int r0 = (int)(b % ny) * RT;
  const bool whole = r0 + RT <= rows && s0 + W <= S && S % VEC == 0;
#pragma unroll
  for (int pass = 0; pass < RT / 32; ++pass) {
    const int r = pass * 32 + (t >> 3);  // 32 rows per pass, 8 lanes per row
    const int c = (t & 7) * VEC;
    const int row = r0 + r;
    const long long s = s0 + c;
    T v[VEC];
    if (whole) {
      typedef T TV __attribute__((ext_vector_type(VEC)));
      const TV x = __builtin_nontemporal_load((const TV*)(in + (long long)row * S + s));
#pragma unroll
      for (int e = 0; e < VEC; ++e) v[e] = x[e];
    } else {
#pragma unroll
      for (int e = 0; e < VEC; ++e) v[e] = (row < rows && s + e < S) ? in[(long long)row * S + s + e] : T(0);
    }
#pragma unroll
    for (int e = 0; e < VEC; ++e) tile[r][c + e] = v[e];
  }
  __syncthreads();
  typedef double D2 __attribute__((ext_vector_type(2)));
  const int wave = t >> 6, lane = t & 63;
  const bool pairs = whole && rows % 2 == 0;
  for (int c = wave; c < W; c += 4) {
    const long long s = s0 + c;
#pragma unroll
    for (int k = 0; k < RT / 128; ++k) {
      const int r = k * 128 + 2 * lane;
      const int row = r0 + r;
      if (pairs) {
        const D2 v = {(double)tile[r][c], (double)tile[r + 1][c]};
        __builtin_nontemporal_store(v, (D2*)(out + s * rows + row));
      } else {
#pragma unroll
        for (int e = 0; e < 2; ++e)
          if (row + e < rows && s < S) __builtin_nontemporal_store((double)tile[r + e][c], out + s * rows + row + e);
      }
    }
  }
}

}  // namespace

// This device's address of `scratch` (and its size), looked up once per device with that device
// current and kept.
hipError_t sdesim_scratch_address(double** ptr, size_t* bytes) {
  static std::mutex mu;
  static std::map<int, double*> addr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(mu);
  auto it = addr.find(dev);
  if (it == addr.end()) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    it = addr.emplace(dev, (double*)p).first;
  }
  *ptr = it->second;
  *bytes = SDESIM_SCRATCH_BYTES;
  return hipSuccess;
}

hipError_t sdesim_launch_finalize(const double* partials, int nblocks, int kq, long long S, const double* shifts,
                                  double* out, hipStream_t stream) {
  hipLaunchKernelGGL(sdesim_finalize, dim3(kq), dim3(256), 0, stream, partials, nblocks, kq, S, shifts, out);
  return hipGetLastError();
}

hipError_t sdesim_launch_to_long(const void* filtration, int f32, long long S, int rows, double* out,
                                 hipStream_t stream) {
  const long long W = f32 ? 32 : 16;  // scenarios per tile (256 rows per tile)
  const int ny = (rows + 255) / 256;
  const long long ntiles = ((S + W - 1) / W) * ny;
  // A dispatch holds fewer than 2^32 work-items. 2^24 tiles of 256 x W would be 550 GB of input.
  if (ntiles >= (1LL << 24)) return hipErrorInvalidValue;
  const dim3 grid((unsigned)ntiles);
  if (f32)
    hipLaunchKernelGGL(sdesim_to_long<float>, grid, dim3(256), 0, stream, (const float*)filtration, S, rows, ny, out);
  else
    hipLaunchKernelGGL(sdesim_to_long<double>, grid, dim3(256), 0, stream, (const double*)filtration, S, rows, ny, out);
  return hipGetLastError();
}

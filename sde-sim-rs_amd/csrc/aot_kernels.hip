// aot_kernels.hip -- kernels compiled ahead of time into libsdesim.so (gfx950).
//   * sdesim_finalize: cross-block reduction of the per-block (sum, sum of squares) partials of
//     the fused kernel into terminal mean and sample variance; Neumaier-compensated f64.
//   * sdesim_to_long: [T*P][S] step-major filtration (f32 or f64) -> [S][T*P] f64, the value
//     column of the reference's long layout (src/filtration.rs:81-116), through LDS tiles so that
//     both the read and the write are coalesced.
#include <hip/hip_runtime.h>

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

// Tile of 64 rows (time-process pairs) x 64 scenarios; 256 threads = 64 x 4.
template <typename T>
__global__ void __launch_bounds__(256) sdesim_to_long(const T* __restrict__ in, long long S, int rows,
                                                      double* __restrict__ out) {
  __shared__ double tile[64][65];
  const long long s0 = (long long)blockIdx.x * 64;
  const int r0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int r = ty; r < 64; r += 4) {
    const int row = r0 + r;
    const long long s = s0 + tx;
    if (row < rows && s < S) tile[r][tx] = (double)in[(long long)row * S + s];
  }
  __syncthreads();
  for (int c = ty; c < 64; c += 4) {
    const long long s = s0 + c;
    const int row = r0 + tx;
    if (row < rows && s < S) out[s * rows + row] = tile[tx][c];
  }
}

}  // namespace

hipError_t sdesim_launch_finalize(const double* partials, int nblocks, int kq, long long S, const double* shifts,
                                  double* out, hipStream_t stream) {
  hipLaunchKernelGGL(sdesim_finalize, dim3(kq), dim3(256), 0, stream, partials, nblocks, kq, S, shifts, out);
  return hipGetLastError();
}

hipError_t sdesim_launch_to_long(const void* filtration, int f32, long long S, int rows, double* out,
                                 hipStream_t stream) {
  dim3 grid((unsigned)((S + 63) / 64), (unsigned)((rows + 63) / 64));
  if (grid.y > 65535u) return hipErrorInvalidValue;
  if (f32)
    hipLaunchKernelGGL(sdesim_to_long<float>, grid, dim3(256), 0, stream, (const float*)filtration, S, rows, out);
  else
    hipLaunchKernelGGL(sdesim_to_long<double>, grid, dim3(256), 0, stream, (const double*)filtration, S, rows, out);
  return hipGetLastError();
}

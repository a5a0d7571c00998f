// runtime_lmm.cpp -- host side of the forward-rate model kernels (include/sdesim.h: sdesim_lmm_*).
// No reference counterpart: the reference has no built-in models and computes no sensitivities.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/sdesim.h"
#include "device/sdesim_args.hpp"

hipError_t sdesim_launch_lmm(int f32, int mode, int grid, const LmmArgs& a, hipStream_t stream);
hipError_t sdesim_launch_sum_partials(const double* partials, int nblocks, int nq, double* out, hipStream_t stream);
int sdesim_set_error(int code, const std::string& msg);
void sdesim_log(const char* fmt, ...);
void sdesim_host_sobol_tables(int dims, int nslices, std::vector<unsigned int>& out);
void sdesim_chacha8_words(unsigned long long seed, size_t n, std::vector<unsigned int>& out);
extern "C" const int sdesim_jk_dims;  // dimensions of the embedded Joe-Kuo table

namespace {
constexpr int kNF = 40, kF = 3, kSeg = 16, kNQ = 2 + kNF;

struct LFail {
  int code;
  std::string msg;
};
void lk(hipError_t e, const char* what) {
  if (e != hipSuccess) throw LFail{SDESIM_ERUNTIME, std::string(what) + ": " + hipGetErrorString(e)};
}
template <class F>
int guard(F&& f) {
  try {
    f();
    return SDESIM_OK;
  } catch (const LFail& e) {
    return sdesim_set_error(e.code, e.msg);
  } catch (const std::exception& e) {
    return sdesim_set_error(SDESIM_ERUNTIME, e.what());
  }
}
}  // namespace

struct sdesim_lmm {
  sdesim_lmm_spec spec{};
  long long grid = 0;
  unsigned int* tab = nullptr;    // slice 0, [steps][64][3], mask complements folded in
  unsigned int* tabup = nullptr;  // slices 1-2, [2][256][steps * 3]
  int tab_words = 0;
  int storage = SDESIM_STORAGE_CHECKPOINTS;  // resolved at creation: TAPE or CHECKPOINTS
  void* ckpt = nullptr;                      // checkpoints (CHECKPOINTS)
  void* tape = nullptr;                      // the segment tape (CHECKPOINTS) or the state tape (TAPE)
  double* pard = nullptr;
  float* parf = nullptr;
  double* partials = nullptr;
  double* sums = nullptr;
  ~sdesim_lmm() {
    for (void* p : {(void*)tab, (void*)tabup, ckpt, tape, (void*)pard, (void*)parf, (void*)partials, (void*)sums})
      if (p) (void)hipFree(p);
  }
};

extern "C" {

int sdesim_lmm_create(const sdesim_lmm_spec* spec, sdesim_lmm** out) {
  return guard([&] {
    if (!spec || !out) throw LFail{SDESIM_EVALUE, "null argument"};
    const sdesim_lmm_spec& s = *spec;
    if (s.forwards != kNF || s.factors != kF) throw LFail{SDESIM_EVALUE, "compiled for 40 forwards and 3 factors"};
    if (s.steps < 1 || s.steps_per_period < 1) throw LFail{SDESIM_EVALUE, "steps must be >= 1"};
    if (s.expiry_index < 1 || s.expiry_index >= kNF) throw LFail{SDESIM_EVALUE, "expiry_index out of range"};
    if ((long long)s.steps_per_period * s.expiry_index != s.steps)
      throw LFail{SDESIM_EVALUE, "steps must equal steps_per_period * expiry_index"};
    if (s.paths < 1 || s.paths > (1LL << 22)) throw LFail{SDESIM_EVALUE, "paths must be in [1, 2^22]"};
    if (!(s.tenor > 0.0)) throw LFail{SDESIM_EVALUE, "tenor must be > 0"};
    if ((long long)s.steps * kF > sdesim_jk_dims)
      throw LFail{SDESIM_EVALUE, "Sobol dimension (steps x factors) exceeds " + std::to_string(sdesim_jk_dims)};
    if (s.storage < SDESIM_STORAGE_AUTO || s.storage > SDESIM_STORAGE_CHECKPOINTS)
      throw LFail{SDESIM_EVALUE, "storage must be SDESIM_STORAGE_AUTO, _TAPE or _CHECKPOINTS"};
    std::unique_ptr<sdesim_lmm> m(new sdesim_lmm());
    m->spec = s;
    m->grid = (s.paths + 255) / 256;
    // Sobol tables laid out for the wave-aligned read (paths 0 .. S-1: each wave's 64 point
    // indices share their Gray-code bits above bit 5): slice 0 as [steps][64][3] with the
    // complement of each dimension's XOR mask folded in (one vector load per lane and step), and
    // slices 1-2 as [2][256][dims] (consecutive words of a step for the wave's entry, read with
    // scalar loads). Paths stay below 2^22, so three slices cover the Gray codes.
    const int dims = s.steps * kF;
    const size_t stride = 64 + 256 * 2;
    std::vector<unsigned int> tab, masks;
    sdesim_host_sobol_tables(dims, 3, tab);
    sdesim_chacha8_words(s.scramble_seed, (size_t)dims, masks);
    std::vector<unsigned int> t0((size_t)s.steps * 64 * kF), up((size_t)2 * 256 * dims);
    for (int n = 0; n < s.steps; ++n)
      for (int e = 0; e < 64; ++e)
        for (int f = 0; f < kF; ++f) {
          const int d = n * kF + f;
          t0[((size_t)n * 64 + e) * kF + f] = tab[(size_t)d * stride + e] ^ ~masks[d];
        }
    for (int sl = 1; sl < 3; ++sl)
      for (int e = 0; e < 256; ++e)
        for (int d = 0; d < dims; ++d)
          up[((size_t)(sl - 1) * 256 + e) * dims + d] = tab[(size_t)d * stride + 64 + 256 * (sl - 1) + e];
    lk(hipMalloc(&m->tab, t0.size() * sizeof(unsigned int)), "hipMalloc(sobol tables)");
    lk(hipMemcpy(m->tab, t0.data(), t0.size() * sizeof(unsigned int), hipMemcpyHostToDevice), "hipMemcpy");
    lk(hipMalloc(&m->tabup, up.size() * sizeof(unsigned int)), "hipMalloc(sobol tables)");
    lk(hipMemcpy(m->tabup, up.data(), up.size() * sizeof(unsigned int), hipMemcpyHostToDevice), "hipMemcpy");
    m->tab_words = (int)t0.size();
    const size_t rb = s.precision == SDESIM_F32 ? 4 : 8;
    const size_t slots = (size_t)(s.steps + kSeg - 1) / kSeg;
    // One step of the state tape is 40 * paths * rb <= 1.34 GB (paths <= 2^22): one buffer
    // resource always covers it.
    const size_t tape = (size_t)kNF * (size_t)s.paths * rb * ((size_t)s.steps + 1);  // L_0 .. L_N
    size_t free_b = 0, total_b = 0;
    lk(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    if (s.storage == SDESIM_STORAGE_TAPE) {
      if (tape + (64u << 20) > free_b) throw LFail{SDESIM_ERUNTIME, "the tape does not fit in free device memory"};
      m->storage = SDESIM_STORAGE_TAPE;
    } else if (s.storage == SDESIM_STORAGE_AUTO && tape + (64u << 20) <= free_b / 2) {
      m->storage = SDESIM_STORAGE_TAPE;
    }
    if (m->storage == SDESIM_STORAGE_TAPE) {
      lk(hipMalloc(&m->tape, tape), "hipMalloc(state tape)");
    } else {
      lk(hipMalloc(&m->ckpt, slots * kNF * (size_t)s.paths * rb), "hipMalloc(checkpoints)");
      lk(hipMalloc(&m->tape, (size_t)kSeg * kNF * (size_t)s.paths * rb), "hipMalloc(segment tape)");
    }
    lk(hipMalloc(&m->pard, sizeof(double) * kNF * (1 + kF)), "hipMalloc(parameters)");
    lk(hipMalloc(&m->parf, sizeof(float) * kNF * (1 + kF)), "hipMalloc(parameters)");
    lk(hipMalloc(&m->partials, sizeof(double) * kNQ * m->grid), "hipMalloc(partials)");
    lk(hipMalloc(&m->sums, sizeof(double) * kNQ), "hipMalloc(sums)");
    sdesim_log("forward-rate model: %lld paths, %d steps, adjoint storage %s", (long long)s.paths, s.steps,
               m->storage == SDESIM_STORAGE_TAPE ? "tape" : "checkpoints");
    *out = m.release();
  });
}

void sdesim_lmm_free(sdesim_lmm* m) { delete m; }

int32_t sdesim_lmm_storage(const sdesim_lmm* m) { return m ? m->storage : -1; }

int sdesim_lmm_eval(sdesim_lmm* m, const double* L0, const double* vol, const double* loadings, int32_t adjoint,
                    double* out, void* stream) {
  return guard([&] {
    if (!m || !L0 || !vol || !loadings || !out) throw LFail{SDESIM_EVALUE, "null argument"};
    const sdesim_lmm_spec& s = m->spec;
    LmmArgs a;
    std::memset(&a, 0, sizeof a);
    std::vector<double> pd(kNF * (1 + kF));
    std::vector<float> pf(kNF * (1 + kF));
    for (int i = 0; i < kNF; ++i) {
      a.L0[i] = L0[i];
      pd[i] = vol[i];
      for (int f = 0; f < kF; ++f) pd[kNF + i * kF + f] = loadings[i * kF + f];
    }
    for (size_t q = 0; q < pd.size(); ++q) pf[q] = (float)pd[q];
    hipStream_t st = (hipStream_t)stream;
    // Ordered before the launch on the same stream; the host vectors outlive the synchronous
    // wait at the end of this call.
    lk(hipMemcpyAsync(m->pard, pd.data(), sizeof(double) * pd.size(), hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    lk(hipMemcpyAsync(m->parf, pf.data(), sizeof(float) * pf.size(), hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    a.pard = m->pard;
    a.parf = m->parf;
    const double T = s.tenor * s.expiry_index;
    a.dt = T / s.steps;
    a.sdt = std::sqrt(a.dt);
    a.tau = s.tenor;
    a.strike = s.strike;
    a.paths = s.paths;
    a.sobol_tab = m->tab;
    a.sobol_tab_words = m->tab_words;
    a.sobol_up = m->tabup;
    a.sobol_dims = s.steps * kF;
    a.checkpoints = m->ckpt;
    a.tape = m->tape;
    a.partials = m->partials;
    a.steps = s.steps;
    a.steps_per_period = s.steps_per_period;
    a.expiry_index = s.expiry_index;
    const int mode = !adjoint ? 0 : m->storage == SDESIM_STORAGE_TAPE ? 2 : 1;
    lk(sdesim_launch_lmm(s.precision == SDESIM_F32 ? 1 : 0, mode, (int)m->grid, a, st), "lmm launch");
    lk(sdesim_launch_sum_partials(m->partials, (int)m->grid, kNQ, m->sums, st), "lmm reduce");
    std::vector<double> sums(kNQ);
    lk(hipMemcpyAsync(sums.data(), m->sums, sizeof(double) * kNQ, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    lk(hipStreamSynchronize(st), "hipStreamSynchronize");
    const double n = (double)s.paths;
    out[0] = sums[0] / n;
    out[1] = s.paths > 1 ? (sums[1] - sums[0] * out[0]) / (n - 1.0) : 0.0;
    if (adjoint)
      for (int i = 0; i < kNF; ++i) out[2 + i] = sums[2 + i] / n;
  });
}

}  // extern "C"

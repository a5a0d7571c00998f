// runtime_basket.cpp -- host side of the correlated-basket price / adjoint kernels
// (include/sdesim.h: sdesim_basket_*). No reference counterpart: the reference computes no
// sensitivities; a basket in the reference is a universe of correlated GBM equations whose
// Cholesky factors are written into the coefficients.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/sdesim.h"
#include "device/sdesim_args.hpp"

hipError_t sdesim_launch_basket(int f32, int assets, int mode, int grid, const BasketArgs& a, int nq,
                                double* sums, hipStream_t stream);
int sdesim_set_error(int code, const std::string& msg);
void sdesim_log(const char* fmt, ...);

namespace {
constexpr int kSeg = 19;  // checkpoint interval of the compiled reverse sweep

struct BFail {
  int code;
  std::string msg;
};
void bk(hipError_t e, const char* what) {
  if (e != hipSuccess) throw BFail{SDESIM_ERUNTIME, std::string(what) + ": " + hipGetErrorString(e)};
}
template <class F>
int guard(F&& f) {
  try {
    f();
    return SDESIM_OK;
  } catch (const BFail& e) {
    return sdesim_set_error(e.code, e.msg);
  } catch (const std::exception& e) {
    return sdesim_set_error(SDESIM_ERUNTIME, e.what());
  }
}
}  // namespace

struct sdesim_basket {
  sdesim_basket_spec spec{};
  int D = 8;          // compiled width (1 or 8)
  int nq = 0;
  long long grid = 0;
  int storage = SDESIM_STORAGE_CHECKPOINTS;  // resolved at creation: TAPE or CHECKPOINTS
  void* states = nullptr;                   // the tape or the checkpoints
  double* partials = nullptr;
  double* sums = nullptr;
  ~sdesim_basket() {
    if (states) (void)hipFree(states);
    if (partials) (void)hipFree(partials);
    if (sums) (void)hipFree(sums);
  }
};

extern "C" {

int sdesim_basket_create(const sdesim_basket_spec* spec, sdesim_basket** out) {
  return guard([&] {
    if (!spec || !out) throw BFail{SDESIM_EVALUE, "null argument"};
    const sdesim_basket_spec& s = *spec;
    if (s.assets < 1 || s.assets > 8) throw BFail{SDESIM_EVALUE, "assets must be between 1 and 8"};
    if (s.steps < 1) throw BFail{SDESIM_EVALUE, "steps must be >= 1"};
    if (s.paths < 1) throw BFail{SDESIM_EVALUE, "paths must be >= 1"};
    if (!(s.horizon > 0.0)) throw BFail{SDESIM_EVALUE, "horizon must be > 0"};
    if (!(s.width >= 0.0)) throw BFail{SDESIM_EVALUE, "width must be >= 0"};
    if (s.storage < SDESIM_STORAGE_AUTO || s.storage > SDESIM_STORAGE_CHECKPOINTS)
      throw BFail{SDESIM_EVALUE, "storage must be SDESIM_STORAGE_AUTO, _TAPE or _CHECKPOINTS"};
    std::unique_ptr<sdesim_basket> b(new sdesim_basket());
    b->spec = s;
    b->D = s.assets == 1 ? 1 : 8;
    b->nq = 2 + 3 * b->D + b->D * (b->D + 1) / 2;
    b->grid = (s.paths + 255) / 256;
    if (b->grid > 0x7fffffffLL) throw BFail{SDESIM_EVALUE, "too many paths"};
    const size_t rb = s.precision == SDESIM_F32 ? 4 : 8;
    const long long slots = (s.steps + kSeg - 1) / kSeg;
    const size_t ck = (size_t)slots * b->D * (size_t)s.paths * rb;
    const size_t step_bytes = (size_t)b->D * (size_t)s.paths * rb;  // one step of the tape
    const size_t tape = step_bytes * ((size_t)s.steps + 1);  // X_0 .. X_N
    size_t free_b = 0, total_b = 0;
    bk(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    const bool tape_addressable = step_bytes < (1ull << 31);  // one buffer resource per step
    if (s.storage == SDESIM_STORAGE_TAPE) {
      if (!tape_addressable) throw BFail{SDESIM_EVALUE, "the tape needs assets * paths * size < 2^31 bytes per step"};
      if (tape + (64u << 20) > free_b) throw BFail{SDESIM_ERUNTIME, "the tape does not fit in free device memory"};
      b->storage = SDESIM_STORAGE_TAPE;
    } else if (s.storage == SDESIM_STORAGE_AUTO && tape_addressable && tape + (64u << 20) <= free_b / 2) {
      b->storage = SDESIM_STORAGE_TAPE;
    }
    if (b->storage == SDESIM_STORAGE_TAPE) {
      bk(hipMalloc(&b->states, tape), "hipMalloc(tape)");
    } else {
      if (ck + (64u << 20) > free_b) throw BFail{SDESIM_ERUNTIME, "checkpoints do not fit in free device memory"};
      bk(hipMalloc(&b->states, ck), "hipMalloc(checkpoints)");
    }
    bk(hipMalloc(&b->partials, sizeof(double) * b->nq * b->grid), "hipMalloc(partials)");
    bk(hipMalloc(&b->sums, sizeof(double) * b->nq), "hipMalloc(sums)");
    sdesim_log("basket: %d assets, %lld paths, %d steps, adjoint storage %s", s.assets, (long long)s.paths, s.steps,
               b->storage == SDESIM_STORAGE_TAPE ? "tape" : "checkpoints");
    *out = b.release();
  });
}

void sdesim_basket_free(sdesim_basket* b) { delete b; }

int32_t sdesim_basket_storage(const sdesim_basket* b) { return b ? b->storage : -1; }

int sdesim_basket_eval(sdesim_basket* b, const double* x0, const double* mu, const double* sigma,
                       const double* weight, const double* chol, int32_t adjoint, double* out, void* stream) {
  return guard([&] {
    if (!b || !x0 || !mu || !sigma || !weight || !chol || !out) throw BFail{SDESIM_EVALUE, "null argument"};
    const sdesim_basket_spec& s = b->spec;
    const int A = s.assets, D = b->D;
    BasketArgs a;
    std::memset(&a, 0, sizeof a);
    for (int i = 0; i < 8; ++i) a.chol[i * 8 + i] = 1.0;  // padding assets: identity rows, zero weight
    for (int i = 0; i < A; ++i) {
      a.x0[i] = x0[i];
      a.mu[i] = mu[i];
      a.sigma[i] = sigma[i];
      a.weight[i] = weight[i];
      for (int k = 0; k < A; ++k) a.chol[i * 8 + k] = k <= i ? chol[i * A + k] : 0.0;
    }
    a.dt = s.horizon / s.steps;
    a.sdt = std::sqrt(a.dt);
    a.strike = s.strike;
    a.width = s.width;
    a.seed = s.seed;
    a.paths = s.paths;
    a.steps = s.steps;
    a.states = b->states;
    a.partials = b->partials;
    a.assets = A;
    hipStream_t st = (hipStream_t)stream;
    const int mode = !adjoint ? 0 : b->storage == SDESIM_STORAGE_TAPE ? 2 : 1;
    bk(sdesim_launch_basket(s.precision == SDESIM_F32 ? 1 : 0, A, mode, (int)b->grid, a, b->nq, b->sums, st),
       "basket launch");
    std::vector<double> sums(b->nq);
    bk(hipMemcpyAsync(sums.data(), b->sums, sizeof(double) * b->nq, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    bk(hipStreamSynchronize(st), "hipStreamSynchronize");
    const double n = (double)s.paths;
    const double m1 = sums[0] / n;
    out[0] = m1;
    out[1] = s.paths > 1 ? (sums[1] - sums[0] * m1) / (n - 1.0) : 0.0;
    if (adjoint) {
      for (int i = 0; i < A; ++i) {
        out[2 + i] = sums[2 + i] / n;
        out[2 + A + i] = sums[2 + D + i] / n;
        out[2 + 2 * A + i] = sums[2 + 2 * D + i] / n;
      }
      int q = 2 + 3 * A;
      for (int i = 0; i < A; ++i)
        for (int k = 0; k <= i; ++k) out[q++] = sums[2 + 3 * D + i * (i + 1) / 2 + k] / n;
    }
  });
}

}  // extern "C"

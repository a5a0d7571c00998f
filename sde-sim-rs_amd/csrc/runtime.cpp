// runtime.cpp -- host runtime and C ABI (include/sdesim.h).
//
// Replaces sim::simulate's orchestration (src/sim/mod.rs:20-92): instead of a rayon loop over
// scenarios that clones the universe, builds an RNG per scenario and steps it through boxed
// trait objects, a plan generates the fused kernel for the universe (hipRTC, cached by source
// text), uploads the time grid, initial row and Sobol tables (in XOR mode with the scramble
// masks folded in) once, and every run is one kernel launch over all paths (plus a one-block
// reduction when terminal moments are asked).
#include <hip/hip_runtime_api.h>
#include <hip/hiprtc.h>

#include <unistd.h>

#include <chrono>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "../../include/sdesim.h"
#include "codegen.hpp"
#include "device/sdesim_args.hpp"
#include "universe.hpp"

// Embedded at build time: the device header text (for hipRTC) and the Joe-Kuo D6 table.
extern "C" const char* sdesim_embedded_device_header;
extern "C" const char* sdesim_embedded_args_header;
extern "C" const int sdesim_jk_dims;
extern "C" const unsigned int sdesim_jk_poly[];
extern "C" const unsigned int sdesim_jk_vinit[];  // [dims][18]

// Ahead-of-time kernels (aot_kernels.hip).
hipError_t sdesim_launch_finalize(const double* partials, int nblocks, int kq, long long S, const double* shifts,
                                  double* out, hipStream_t stream);
hipError_t sdesim_launch_to_long(const void* filtration, int f32, long long S, int rows, double* out,
                                 hipStream_t stream);
hipError_t sdesim_scratch_address(double** ptr, size_t* bytes);

struct sdesim_universe {
  sdesim::Universe u;
};

void sdesim_log(const char* fmt, ...);

namespace {

thread_local std::string g_last_error;

struct Failure {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const std::string& m) { throw Failure{code, m}; }

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(SDESIM_ERUNTIME, std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
int guarded(F&& f) {
  try {
    f();
    return SDESIM_OK;
  } catch (const Failure& e) {
    g_last_error = e.msg;
    return e.code;
  } catch (const sdesim::ParseError& e) {
    g_last_error = e.what();
    return SDESIM_EVALUE;
  } catch (const sdesim::UndefinedName& e) {
    g_last_error = std::string("called `Result::unwrap()` on an `Err` value: ") + e.what();
    return SDESIM_EPANIC;
  } catch (const sdesim::Panic& e) {
    g_last_error = e.what();
    return SDESIM_EPANIC;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return SDESIM_ERUNTIME;
  }
}

// ------------------------------------------------------------------------------------------
// Kernel cache (hipRTC)
// ------------------------------------------------------------------------------------------
struct CompiledKernel {
  hipModule_t module = nullptr;
  hipFunction_t function = nullptr;
};

std::mutex g_cache_mu;
std::map<std::pair<int, std::string>, CompiledKernel> g_kernels;

const char* const kRtcOptions[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=fast",
                                   "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};

// The device header the run-time kernels compile against: the copy embedded at build time.
const std::string& device_header_text() {
  static const std::string text(sdesim_embedded_device_header);
  return text;
}

// Compile generated source to a gfx950 code object (needs no device).
std::vector<char> rtc_compile(const std::string& src) {
  hiprtcProgram prog;
  const char* hdrs[] = {device_header_text().c_str(), sdesim_embedded_args_header};
  const char* names[] = {"sdesim_device.hpp", "sdesim_args.hpp"};
  if (hiprtcCreateProgram(&prog, src.c_str(), "sdesim_kernel.hip", 2, hdrs, names) != HIPRTC_SUCCESS)
    fail(SDESIM_ERUNTIME, "hiprtcCreateProgram failed");
  hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof(kRtcOptions) / sizeof(kRtcOptions[0])), kRtcOptions);
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    fail(SDESIM_ERUNTIME, "kernel compilation failed: " + log);
  }
  size_t code_size = 0;
  hiprtcGetCodeSize(prog, &code_size);
  std::vector<char> code(code_size);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  return code;
}

// Code objects persist across processes in the cache directory (SDESIM_CACHE_DIR, else
// $HOME/.cache/sdesim), one file per kernel named by a 64-bit FNV-1a hash of everything the code
// depends on: the generated source, both device headers, the compile options and the hipRTC
// version. A file holds a 24-byte header (magic, code size, FNV-1a of the code) before the code
// object; one that is missing, short or fails the check is compiled again, before anything is
// handed to the HIP loader. New files are written under a temporary name and renamed into place,
// so concurrent processes never read a partial file. Writing is best effort: a read-only or
// missing home directory only costs the compile.
uint64_t fnv1a(const void* data, size_t n, uint64_t h = 0xcbf29ce484222325ull) {
  const unsigned char* p = (const unsigned char*)data;
  for (size_t i = 0; i < n; ++i) h = (h ^ p[i]) * 0x100000001b3ull;
  return h;
}

constexpr uint64_t kCodeMagic = 0x314b4f43454453ull;  // "SDECOK1"

// The identity of a run-time kernel's code object: FNV-1a over the generated source, both device
// headers, the compile options and the hipRTC version.
uint64_t code_key(const std::string& src) {
  int major = 0, minor = 0;
  hiprtcVersion(&major, &minor);
  uint64_t h = fnv1a(src.data(), src.size());
  const std::string& hdr = device_header_text();
  h = fnv1a(hdr.data(), hdr.size(), h);
  h = fnv1a(sdesim_embedded_args_header, std::strlen(sdesim_embedded_args_header), h);
  for (const char* o : kRtcOptions) h = fnv1a(o, std::strlen(o) + 1, h);
  const int ver[2] = {major, minor};
  return fnv1a(ver, sizeof ver, h);
}

std::string code_cache_file(const std::string& src) {
  const char* dir = std::getenv("SDESIM_CACHE_DIR");
  std::string d;
  if (dir && *dir) d = dir;
  else if (const char* home = std::getenv("HOME")) d = std::string(home) + "/.cache/sdesim";
  else return std::string();
  char name[40];
  std::snprintf(name, sizeof name, "/k_%016llx.co", (unsigned long long)code_key(src));
  return d + name;
}

std::vector<char> read_cached_code(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  uint64_t head[3] = {0, 0, 0};
  if (!in || !in.read((char*)head, sizeof head) || head[0] != kCodeMagic || head[1] == 0 || head[1] > (1ull << 30))
    return {};
  std::vector<char> code(head[1]);
  if (!in.read(code.data(), (std::streamsize)code.size()) || fnv1a(code.data(), code.size()) != head[2]) return {};
  return code;
}

void write_cached_code(const std::string& path, const std::vector<char>& code) {
  std::error_code ec;
  std::filesystem::create_directories(std::filesystem::path(path).parent_path(), ec);
  const std::string tmp = path + ".tmp" + std::to_string((long long)getpid());
  {
    std::ofstream out(tmp, std::ios::binary | std::ios::trunc);
    const uint64_t head[3] = {kCodeMagic, (uint64_t)code.size(), fnv1a(code.data(), code.size())};
    if (!out || !out.write((const char*)head, sizeof head) || !out.write(code.data(), (std::streamsize)code.size()))
      ec = std::make_error_code(std::errc::io_error);
  }
  if (ec || std::rename(tmp.c_str(), path.c_str()) != 0) std::filesystem::remove(tmp, ec);
}

CompiledKernel compile_kernel(const std::string& src, const std::string& name) {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  std::lock_guard<std::mutex> lock(g_cache_mu);
  auto key = std::make_pair(dev, src);
  auto it = g_kernels.find(key);
  if (it != g_kernels.end()) {
    sdesim_log("kernel %s: already loaded on device %d", name.c_str(), dev);
    return it->second;
  }
  const std::string path = code_cache_file(src);
  CompiledKernel k;
  bool loaded = false;
  const auto t0 = std::chrono::steady_clock::now();
  if (!path.empty()) {
    const std::vector<char> code = read_cached_code(path);
    if (!code.empty()) {
      loaded = hipModuleLoadData(&k.module, code.data()) == hipSuccess;
      if (!loaded) (void)hipGetLastError();  // a code object the runtime refuses: compile instead
    }
  }
  if (!loaded) {
    const std::vector<char> code = rtc_compile(src);
    hip_check(hipModuleLoadData(&k.module, code.data()), "hipModuleLoadData");
    if (!path.empty()) write_cached_code(path, code);
  }
  hip_check(hipModuleGetFunction(&k.function, k.module, name.c_str()), "hipModuleGetFunction");
  sdesim_log("kernel %s: %s in %.1f ms on device %d", name.c_str(),
             loaded ? ("loaded from " + path).c_str() : "compiled with hipRTC",
             std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(), dev);
  g_kernels[key] = k;
  return k;
}

// ------------------------------------------------------------------------------------------
// Sobol direction numbers and sliced XOR tables
// ------------------------------------------------------------------------------------------
// Direction numbers V[d][k] (k = 0..31 multiplies bit k of the Gray code) from the Joe-Kuo
// table by the Bratley-Fox recurrence m_k = 2a_1 m_{k-1} ^ ... ^ 2^{s-1}a_{s-1} m_{k-s+1} ^
// 2^s m_{k-s} ^ m_{k-s}; dimension 0 is the van der Corput sequence (all m_k = 1).
void direction_numbers(int d, unsigned int V[32]) {
  unsigned long long m[33];
  if (d == 0) {
    for (int k = 1; k <= 32; ++k) m[k] = 1;
  } else {
    const unsigned int poly = sdesim_jk_poly[d];
    int s = 0;
    while ((poly >> (s + 1)) != 0u) ++s;
    for (int k = 1; k <= s && k <= 32; ++k) m[k] = sdesim_jk_vinit[(size_t)d * 18 + (k - 1)];
    for (int k = s + 1; k <= 32; ++k) {
      unsigned long long v = m[k - s] ^ (m[k - s] << s);
      for (int i = 1; i < s; ++i)
        if ((poly >> (s - i)) & 1u) v ^= m[k - i] << i;
      m[k] = v;
    }
  }
  for (int k = 1; k <= 32; ++k) V[k - 1] = (unsigned int)(m[k] << (32 - k));
}

void build_sobol_tables(int want, int nslices, std::vector<unsigned int>& host);

struct SobolTables {
  int dims = 0, nslices = 0;
  unsigned int* dev = nullptr;
};
std::mutex g_sobol_mu;
std::map<std::pair<int, int>, SobolTables> g_sobol;  // (device, slices) -> largest tables built so far

// Tables covering `dims` dimensions and Gray codes below 2^(6 + 8 (nslices - 1)).
SobolTables sobol_tables(int dims, int nslices) {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  std::lock_guard<std::mutex> lock(g_sobol_mu);
  SobolTables& t = g_sobol[std::make_pair(dev, nslices)];
  if (t.dev && t.dims >= dims) return t;
  int want = std::max(dims, t.dims);
  std::vector<unsigned int> host;
  build_sobol_tables(want, nslices, host);
  // The previous tables may still be read by queued kernels; they are kept until exit.
  unsigned int* d_tab = nullptr;
  hip_check(hipMalloc(&d_tab, host.size() * sizeof(unsigned int)), "hipMalloc(sobol tables)");
  hip_check(hipMemcpy(d_tab, host.data(), host.size() * sizeof(unsigned int), hipMemcpyHostToDevice),
            "hipMemcpy(sobol tables)");
  t.dev = d_tab;
  t.dims = want;
  t.nslices = nslices;
  return t;
}

// Sliced XOR tables [dims][64 + 256 (nslices - 1)] of the direction numbers.
void build_sobol_tables(int want, int nslices, std::vector<unsigned int>& host) {
  if (want > sdesim_jk_dims) throw std::out_of_range("Sobol dimension exceeds the direction-number table");
  const int stride = 64 + 256 * (nslices - 1);
  host.assign((size_t)want * stride, 0u);
  for (int d = 0; d < want; ++d) {
    unsigned int V[32];
    direction_numbers(d, V);
    unsigned int* row = &host[(size_t)d * stride];
    for (int j = 0; j < 64; ++j) {
      unsigned int x = 0;
      for (int b = 0; b < 6; ++b)
        if ((j >> b) & 1) x ^= V[b];
      row[j] = x;
    }
    for (int sl = 1; sl < nslices; ++sl) {
      for (int j = 0; j < 256; ++j) {
        unsigned int x = 0;
        for (int b = 0; b < 8; ++b) {
          int bit = 6 + 8 * (sl - 1) + b;
          if (bit < 32 && ((j >> b) & 1)) x ^= V[bit];
        }
        row[64 + 256 * (sl - 1) + j] = x;
      }
    }
  }
}

// Words of ChaCha8Rng::seed_from_u64(seed): the XOR masks of the digital shift.
void chacha8_words(unsigned long long seed, size_t n, std::vector<unsigned int>& out) {
  unsigned int key[8];
  unsigned long long st = seed;
  for (int i = 0; i < 8; ++i) {
    st = st * 6364136223846793005ULL + 11634580027462260723ULL;
    unsigned int xs = (unsigned int)(((st >> 18) ^ st) >> 27);
    unsigned int rot = (unsigned int)(st >> 59);
    key[i] = (xs >> rot) | (xs << ((32u - rot) & 31u));
  }
  out.resize(n);
  auto rotl = [](unsigned int v, int c) { return (v << c) | (v >> (32 - c)); };
  for (size_t blk = 0; blk * 16 < n; ++blk) {
    unsigned int in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u, key[0], key[1], key[2], key[3],
                           key[4], key[5], key[6], key[7], (unsigned int)blk, (unsigned int)(blk >> 32), 0u, 0u};
    unsigned int x[16];
    std::memcpy(x, in, sizeof x);
    for (int r = 0; r < 8; r += 2) {
      auto qr = [&](int a, int b, int c, int d) {
        x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 16);
        x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 12);
        x[a] += x[b]; x[d] = rotl(x[d] ^ x[a], 8);
        x[c] += x[d]; x[b] = rotl(x[b] ^ x[c], 7);
      };
      qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15);
      qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14);
    }
    for (int w = 0; w < 16 && blk * 16 + w < n; ++w) out[blk * 16 + w] = x[w] + in[w];
  }
}

template <class T>
T* upload(const std::vector<T>& v) {
  T* d = nullptr;
  if (v.empty()) return nullptr;
  hip_check(hipMalloc(&d, v.size() * sizeof(T)), "hipMalloc");
  hip_check(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
  return d;
}

}  // namespace

// Host-side log lines on stderr, one per event (plan creation, kernel compile or cache load, batch
// autotuning, adjoint storage choice), when SDESIM_LOG is set to anything but "" or "0". Logging
// changes no result.
bool sdesim_log_enabled() {
  static const bool on = [] {
    const char* v = std::getenv("SDESIM_LOG");
    return v && *v && std::strcmp(v, "0") != 0;
  }();
  return on;
}

void sdesim_log(const char* fmt, ...) {
  if (!sdesim_log_enabled()) return;
  char line[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(line, sizeof line, fmt, ap);
  va_end(ap);
  std::fprintf(stderr, "[sdesim] %s\n", line);
}

int sdesim_set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

// Tail values of the f32 XOR-mode increment: for qi = 0 .. 1232 (|z| > 5), the A&S inverse
// normal (src/proc/increment.rs:160-179) of q = (qi + 0.5) 2^-32, divided by sqrt(2 ln 2) and
// rounded to f32 (the kernel multiplies the signed sqrt(2 ln 2) sqrt(dt) back in). One copy per
// device, kept until exit.
std::mutex g_tail_mu;
std::map<int, float*> g_tail;

const float* tail_table() {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  std::lock_guard<std::mutex> lock(g_tail_mu);
  float*& d = g_tail[dev];
  if (d) return d;
  std::vector<float> h(1233);
  const double c = std::sqrt(2.0 * std::log(2.0));
  for (int qi = 0; qi < 1233; ++qi) {
    const double q = ((double)qi + 0.5) * 0x1.0p-32;
    const double t = std::sqrt(-2.0 * std::log(q));
    const double x = t - ((0.010328 * t + 0.802853) * t + 2.515517) / (((0.001308 * t + 0.189269) * t + 1.432788) * t + 1.0);
    h[qi] = (float)(x / c);
  }
  hip_check(hipMalloc(&d, h.size() * sizeof(float)), "hipMalloc(tail table)");
  hip_check(hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy(tail table)");
  return d;
}

// Shared with the other runtime files: ChaCha8 words and host Sobol tables.
void sdesim_chacha8_words(unsigned long long seed, size_t n, std::vector<unsigned int>& out) {
  chacha8_words(seed, n, out);
}

// Host copy of the sliced tables [dims][64 + 256 (nslices - 1)], for runtimes that lay them out
// for their own access pattern.
void sdesim_host_sobol_tables(int dims, int nslices, std::vector<unsigned int>& out) {
  build_sobol_tables(dims, nslices, out);
}

struct sdesim_plan {
  CompiledKernel kernel;
  std::string source;
  SimArgs args{};
  sdesim_options opt{};
  int P = 0, M = 0, KQ = 0;
  int block = 256, ppl = 1;
  long long grid = 0;
  std::vector<void*> owned;
  double* partials = nullptr;
  double* moments = nullptr;
  double* shifts = nullptr;
  // Set when a run was issued on a capturing stream: the captured graph holds pointers to the
  // device buffers, and nothing signals the graph's destruction, so they are kept until exit.
  bool captured = false;
  ~sdesim_plan() {
    if (captured) {
      sdesim_log("plan %s freed after a graph capture of its run: its %zu device buffers stay allocated", kname.c_str(),
                 owned.size());
      return;
    }
    for (void* p : owned) (void)hipFree(p);
  }
  std::string kname;
  template <class T>
  T* own(T* p) {
    if (p) owned.push_back((void*)p);
    return p;
  }
};

extern "C" {

void sdesim_default_options(sdesim_options* o) {
  std::memset(o, 0, sizeof *o);
  o->scheme = SDESIM_EULER;
  o->rng = SDESIM_PSEUDO;
  o->scramble = SDESIM_SCRAMBLE_SHIFT;
  o->precision = SDESIM_F64;
  o->compat = SDESIM_COMPAT_REFERENCE;
  o->outputs = SDESIM_OUT_FILTRATION;
  o->sobol_first = -1;
}

int sdesim_parse_equations(const char* const* equations, int32_t n, sdesim_universe** out) {
  return guarded([&] {
    if (!out) fail(SDESIM_EVALUE, "null output handle");
    std::vector<std::string> eqs;
    for (int32_t i = 0; i < n; ++i) eqs.emplace_back(equations[i] ? equations[i] : "");
    std::unique_ptr<sdesim_universe> u(new sdesim_universe());
    u->u = sdesim::parse_equations(eqs);
    *out = u.release();
  });
}

void sdesim_free_universe(sdesim_universe* u) { delete u; }

int32_t sdesim_num_processes(const sdesim_universe* u) { return u ? u->u.P() : 0; }
int32_t sdesim_num_increments(const sdesim_universe* u) { return u ? u->u.M() : 0; }
const char* sdesim_process_name(const sdesim_universe* u, int32_t p) {
  if (!u || p < 0 || p >= u->u.P()) return nullptr;
  return u->u.processes[p].name.c_str();
}
int32_t sdesim_process_is_algebraic(const sdesim_universe* u, int32_t p) {
  if (!u || p < 0 || p >= u->u.P()) return -1;
  return u->u.processes[p].levy ? 0 : 1;
}
int32_t sdesim_process_index(const sdesim_universe* u, const char* name) {
  if (!u || !name) return -1;
  auto it = u->u.registry.find(name);
  return it == u->u.registry.end() ? -1 : it->second;
}

int sdesim_plan_create(const sdesim_universe* uh, const double* times, int32_t T, const double* x0,
                       int64_t S, const sdesim_options* opt_in, const char* payoff, sdesim_plan** out) {
  return guarded([&] {
    if (!uh || !out) fail(SDESIM_EVALUE, "null handle");
    if (S <= 0) fail(SDESIM_EVALUE, "scenarios must be a positive integer");
    if (T <= 0 || !times) fail(SDESIM_EPANIC, "time_steps must not be empty");
    sdesim_options opt;
    if (opt_in) opt = *opt_in; else sdesim_default_options(&opt);
    const sdesim::Universe& U = uh->u;
    {
      std::set<double> seen;
      for (int32_t i = 0; i < T; ++i) {
        if (std::isnan(times[i])) fail(SDESIM_EVALUE, "time_steps must not contain NaN");
        if (!seen.insert(times[i]).second) fail(SDESIM_EVALUE, "time_steps must not contain duplicate values");
      }
    }
    const int N = T - 1, P = U.P(), M = U.M();
    if (opt.scheme != SDESIM_EULER && opt.scheme != SDESIM_RUNGE_KUTTA) fail(SDESIM_EPANIC, "not implemented");
    if (opt.rng == SDESIM_SOBOL && (long long)N * M > sdesim_jk_dims)
      fail(SDESIM_EPANIC, "Sobol dimension (time steps x increments) exceeds " + std::to_string(sdesim_jk_dims));

    std::unique_ptr<sdesim_plan> plan(new sdesim_plan());
    plan->opt = opt;
    plan->P = P;
    plan->M = M;
    plan->args.S = S;
    plan->args.T = T;
    plan->args.N = N;
    if (P == 0) {  // no processes: every scenario has an empty filtration; nothing to launch
      plan->opt.outputs = 0;
      *out = plan.release();
      return;
    }
    const bool has_payoff = payoff && *payoff;
    plan->KQ = P + (has_payoff ? 1 : 0);
    plan->block = opt.block_size > 0 ? opt.block_size : 256;
    if (plan->block % 64 != 0 || plan->block > 1024) fail(SDESIM_EVALUE, "block_size must be a multiple of 64 up to 1024");
    plan->ppl = opt.paths_per_lane > 0 ? opt.paths_per_lane : 1;
    if (plan->ppl > 8) fail(SDESIM_EVALUE, "paths_per_lane must be at most 8");

    int rng_kind;
    if (M == 0) rng_kind = SDE_RNG_NONE;
    else if (opt.rng == SDESIM_PSEUDO) rng_kind = SDE_RNG_PSEUDO;
    else if (opt.rng == SDESIM_SOBOL) rng_kind = opt.scramble == SDESIM_SCRAMBLE_XOR ? SDE_RNG_SOBOL_XOR : SDE_RNG_SOBOL_SHIFT;
    else if (opt.rng == SDESIM_INJECTED) rng_kind = SDE_RNG_INJECTED;
    else rng_kind = SDE_RNG_PSEUDO;  // the reference treats any other string as "pseudo"

    int nslices = 1;
    long long sobol_first = 0;
    if (rng_kind == SDE_RNG_SOBOL_SHIFT || rng_kind == SDE_RNG_SOBOL_XOR) {
      sobol_first = opt.sobol_first >= 0 ? opt.sobol_first : (rng_kind == SDE_RNG_SOBOL_XOR ? 0 : 5);
      const unsigned long long last = (unsigned long long)(sobol_first + S - 1);
      const unsigned long long g = last | (last >> 1);
      int bits = 0;
      while (bits < 64 && (g >> bits) != 0ull) ++bits;
      if (bits > 32) fail(SDESIM_EPANIC, "Sobol sequence exhausted");
      while (6 + 8 * (nslices - 1) < bits) ++nslices;
    }

    sdesim::CodegenOptions co;
    co.scheme = opt.scheme;
    co.compat = opt.compat;
    co.f32 = opt.precision == SDESIM_F32;
    co.rng = rng_kind;
    co.out = opt.outputs & (SDESIM_OUT_FILTRATION | SDESIM_OUT_TERMINAL);
    co.ppl = plan->ppl;
    co.block = plan->block;
    co.nsl = nslices;
    co.wave_aligned = rng_kind == SDE_RNG_SOBOL_XOR && sobol_first % 64 == 0 && plan->block % 64 == 0;
    co.steps = N > 0;
    co.payoff = has_payoff ? payoff : "";

    // Time grid in the reference's arithmetic: dt = t[n+1] - t[n], sqrt(dt) (increment.rs:38-41,75-79).
    std::vector<double> tv(times, times + T), dts(N), sdts(N);
    std::vector<float> tvf(T), dtsf(N), sdtsf(N);
    for (int n = 0; n < N; ++n) { dts[n] = tv[n + 1] - tv[n]; sdts[n] = std::sqrt(dts[n]); }
    for (int i = 0; i < T; ++i) tvf[i] = (float)tv[i];
    for (int n = 0; n < N; ++n) { dtsf[n] = (float)dts[n]; sdtsf[n] = (float)sdts[n]; }
    bool ug = N > 0;
    for (int n = 1; n < N && ug; ++n) {
      if (co.f32) ug = std::memcmp(&dtsf[n], &dtsf[0], 4) == 0 && std::memcmp(&sdtsf[n], &sdtsf[0], 4) == 0;
      else ug = std::memcmp(&dts[n], &dts[0], 8) == 0 && std::memcmp(&sdts[n], &sdts[0], 8) == 0;
    }
    co.uniform_grid = ug;
    co.name = sdesim::kernel_name(co);
    plan->source = sdesim::generate_kernel_source(U, co);
    plan->kname = co.name;
    plan->kernel = compile_kernel(plan->source, co.name);
    std::vector<double> x0v(x0 ? x0 : nullptr, x0 ? x0 + P : nullptr);
    if (!x0) x0v.assign(P, 0.0);
    if (x0v.empty()) x0v.push_back(0.0);

    SimArgs& a = plan->args;
    a.times = plan->own(upload(tv));
    a.dts = plan->own(upload(dts));
    a.sdts = plan->own(upload(sdts));
    a.times_f = plan->own(upload(tvf));
    a.dts_f = plan->own(upload(dtsf));
    a.sdts_f = plan->own(upload(sdtsf));
    a.x0 = plan->own(upload(x0v));
    if (N > 0) {
      a.dt_u = dts[0]; a.sdt_u = sdts[0];
      a.dtf_u = dtsf[0]; a.sdtf_u = sdtsf[0];
    }
    a.S = S;
    a.seed = opt.seed;
    a.N = N;
    a.T = T;
    a.nslices = nslices;
    a.sobol_first = sobol_first;
    if (rng_kind == SDE_RNG_SOBOL_SHIFT) {
      SobolTables st = sobol_tables(std::max(1, N * M), nslices);
      a.sobol_tab = st.dev;
      a.sobol_tab_words = st.dims * (64 + 256 * (nslices - 1));
    } else if (rng_kind == SDE_RNG_SOBOL_XOR) {
      // The plan's own tables with the complement of dimension d's mask XORed into its 64
      // slice-0 entries: every Gray code reads exactly one slice-0 entry, so the slices' XOR is
      // ~(point ^ mask), the form the kernel's uniforms take.
      const int dims = std::max(1, N * M);
      const size_t stride = 64 + 256 * (size_t)(nslices - 1);
      std::vector<unsigned int> tab, masks;
      build_sobol_tables(dims, nslices, tab);
      chacha8_words(opt.scramble_seed, (size_t)dims, masks);
      for (int d = 0; d < dims; ++d)
        for (int j = 0; j < 64; ++j) tab[(size_t)d * stride + j] ^= ~masks[d];
      if (co.wave_aligned && N > 0) {
        // Wave-aligned layout: slice 0 as [N][64][M], slices 1.. as [nslices - 1][256][N * M].
        std::vector<unsigned int> t0((size_t)N * 64 * M), up((size_t)(nslices - 1) * 256 * dims);
        for (int n = 0; n < N; ++n)
          for (int e = 0; e < 64; ++e)
            for (int k = 0; k < M; ++k) t0[((size_t)n * 64 + e) * M + k] = tab[(size_t)(n * M + k) * stride + e];
        for (int sl = 1; sl < nslices; ++sl)
          for (int e = 0; e < 256; ++e)
            for (int d = 0; d < dims; ++d)
              up[((size_t)(sl - 1) * 256 + e) * dims + d] = tab[(size_t)d * stride + 64 + 256 * (sl - 1) + e];
        a.sobol_tab = plan->own(upload(t0));
        a.sobol_tab_words = (int)t0.size();
        a.sobol_up = plan->own(upload(up));
        a.sobol_dims = dims;
      } else {
        a.sobol_tab = plan->own(upload(tab));
        a.sobol_tab_words = (int)(dims * stride);
      }
      a.tail_tab = tail_table();
    }
    if ((opt.outputs & SDESIM_OUT_FILTRATION) &&
        (long long)P * S * (opt.precision == SDESIM_F32 ? 4 : 8) >= (1LL << 31))
      fail(SDESIM_EVALUE, "one filtration row (processes x scenarios) must be below 2 GiB; split the scenarios");
    plan->grid = (S + (long long)plan->block * plan->ppl - 1) / ((long long)plan->block * plan->ppl);
    if (plan->grid > 0x7fffffffLL) fail(SDESIM_EVALUE, "too many scenarios for one launch");
    if (opt.outputs & SDESIM_OUT_TERMINAL) {
      hip_check(hipMalloc(&plan->partials, (size_t)plan->grid * 2 * plan->KQ * sizeof(double)), "hipMalloc(partials)");
      plan->own(plan->partials);
      hip_check(hipMalloc(&plan->moments, (size_t)2 * plan->KQ * sizeof(double)), "hipMalloc(moments)");
      plan->own(plan->moments);
      std::vector<double> shifts(plan->KQ, 0.0);
      for (int p = 0; p < P; ++p) shifts[p] = x0v[p];
      plan->shifts = plan->own(upload(shifts));
      a.partials = plan->partials;
    }
    sdesim_log("plan %s: %lld paths, %d steps, %d processes, %d increments, grid %lld x %d, %d path(s) per lane",
               co.name.c_str(), (long long)S, N, P, M, plan->grid, plan->block, plan->ppl);
    *out = plan.release();
  });
}

void sdesim_plan_free(sdesim_plan* plan) { delete plan; }

int64_t sdesim_plan_filtration_bytes(const sdesim_plan* plan) {
  if (!plan) return 0;
  const long long b = plan->opt.precision == SDESIM_F32 ? 4 : 8;
  return (int64_t)plan->args.T * plan->P * plan->args.S * b;
}

const char* sdesim_plan_source(const sdesim_plan* plan) { return plan ? plan->source.c_str() : nullptr; }

uint64_t sdesim_plan_code_hash(const sdesim_plan* plan) {
  if (!plan || plan->source.empty()) return 0;
  try {
    return code_key(plan->source);
  } catch (...) {
    return 0;
  }
}

// One run of a plan: the fused kernel with its per-block partials in `partials`, then (terminal
// output) the reduction into `moments`. Checks the buffers the plan's outputs need first.
static void launch_plan(sdesim_plan* plan, void* d_filtration, const double* d_injected, double* partials,
                        double* moments, void* stream) {
  if (stream) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing((hipStream_t)stream, &cs) == hipSuccess && cs == hipStreamCaptureStatusActive)
      plan->captured = true;
  }
  SimArgs a = plan->args;
  if ((plan->opt.outputs & SDESIM_OUT_FILTRATION) && !d_filtration)
    fail(SDESIM_EVALUE, "filtration output requested but no buffer given");
  a.filtration = d_filtration;
  if (plan->opt.rng == SDESIM_INJECTED && plan->M > 0) {
    if (!d_injected) fail(SDESIM_EVALUE, "injected uniforms requested but no buffer given");
    a.injected = d_injected;
  }
  a.partials = partials;
  void* params[] = {&a};
  hip_check(hipModuleLaunchKernel(plan->kernel.function, (unsigned)plan->grid, 1, 1, (unsigned)plan->block, 1, 1, 0,
                                  (hipStream_t)stream, params, nullptr),
            "hipModuleLaunchKernel");
  if (plan->opt.outputs & SDESIM_OUT_TERMINAL) {
    hip_check(sdesim_launch_finalize(partials, (int)plan->grid, plan->KQ, a.S, plan->shifts, moments,
                                     (hipStream_t)stream),
              "finalize");
  }
}

int sdesim_plan_run(sdesim_plan* plan, void* d_filtration, const double* d_injected, void* stream) {
  return guarded([&] {
    if (!plan) fail(SDESIM_EVALUE, "null plan");
    if (plan->P == 0) return;
    launch_plan(plan, d_filtration, d_injected, plan->partials, plan->moments, stream);
  });
}

int sdesim_plan_run_scratch(sdesim_plan* plan, void* d_filtration, const double* d_injected, double* d_moments,
                            void* stream) {
  return guarded([&] {
    if (!plan) fail(SDESIM_EVALUE, "null plan");
    if (!(plan->opt.outputs & SDESIM_OUT_TERMINAL)) fail(SDESIM_EVALUE, "plan has no terminal output");
    if (!d_moments) fail(SDESIM_EVALUE, "null moments buffer");
    const unsigned long long need = (unsigned long long)plan->grid * 2ull * (unsigned long long)plan->KQ * 8ull;
    if (need > SDESIM_SCRATCH_BYTES)
      fail(SDESIM_EVALUE, "the plan's terminal partials (" + std::to_string(need) + " bytes) exceed the scratch array (" +
                              std::to_string((unsigned long long)SDESIM_SCRATCH_BYTES) + " bytes); use sdesim_plan_run");
    double* scr = nullptr;
    size_t cap = 0;
    hip_check(sdesim_scratch_address(&scr, &cap), "hipGetSymbolAddress(scratch)");
    if (need > cap) fail(SDESIM_EVALUE, "scratch array smaller than declared");
    launch_plan(plan, d_filtration, d_injected, scr, d_moments, stream);
  });
}

int sdesim_plan_moments(sdesim_plan* plan, double* h_out, void* stream) {
  return guarded([&] {
    if (!plan || !h_out) fail(SDESIM_EVALUE, "null argument");
    if (!(plan->opt.outputs & SDESIM_OUT_TERMINAL)) fail(SDESIM_EVALUE, "plan has no terminal output");
    hip_check(hipMemcpyAsync(h_out, plan->moments, sizeof(double) * 2 * plan->KQ, hipMemcpyDeviceToHost,
                             (hipStream_t)stream),
              "hipMemcpyAsync(moments)");
    hip_check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
  });
}

int sdesim_filtration_to_long(const void* d_filtration, int32_t precision, int64_t S, int32_t T, int32_t P,
                              double* d_long, void* stream) {
  return guarded([&] {
    if (S <= 0 || T <= 0 || P < 0) fail(SDESIM_EVALUE, "bad shape");
    if (precision != SDESIM_F32 && precision != SDESIM_F64) fail(SDESIM_EVALUE, "precision must be f32 or f64");
    if ((int64_t)T * P > INT32_MAX) fail(SDESIM_EVALUE, "n_times * n_processes exceeds 2^31 - 1");
    if (P == 0) return;
    // The transpose runs one block per tile of 256 rows x 32 (f32) or 16 (f64) scenarios.
    const int64_t W = precision == SDESIM_F32 ? 32 : 16;
    if (((S + W - 1) / W) * (((int64_t)T * P + 255) / 256) >= (int64_t(1) << 24))
      fail(SDESIM_EVALUE, "filtration too large for one long-layout transpose (2^24 tiles of 256 rows)");
    if (!d_filtration || !d_long) fail(SDESIM_EVALUE, "null argument");
    hip_check(sdesim_launch_to_long(d_filtration, precision == SDESIM_F32 ? 1 : 0, S, T * P, d_long,
                                    (hipStream_t)stream),
              "to_long");
  });
}

int sdesim_simulate_long(const sdesim_universe* u, const double* times, int32_t T, const char* const* init_names,
                         const double* init_values, int32_t n_init, int64_t S, const sdesim_options* opt_in,
                         double* h_values, void* stream) {
  sdesim_plan* plan = nullptr;
  void* d_filt = nullptr;
  double* d_long = nullptr;
  int rc = guarded([&] {
    if (!u) fail(SDESIM_EVALUE, "null universe");
    if (S <= 0) fail(SDESIM_EVALUE, "scenarios must be a positive integer");
    if (!h_values) fail(SDESIM_EVALUE, "null output buffer");
    if (n_init < 0 || (n_init > 0 && (!init_names || !init_values))) fail(SDESIM_EVALUE, "bad initial values");
    for (int32_t i = 0; i < n_init; ++i)
      if (!init_names[i]) fail(SDESIM_EVALUE, "null initial value name");
    const int P = u->u.P();
    std::vector<double> x0(P, 0.0);
    for (int32_t i = 0; i < n_init; ++i) {
      auto it = u->u.registry.find(init_names[i]);
      if (it != u->u.registry.end()) x0[it->second] = init_values[i];
    }
    sdesim_options opt;
    if (opt_in) opt = *opt_in; else sdesim_default_options(&opt);
    opt.outputs = SDESIM_OUT_FILTRATION;
    int r = sdesim_plan_create(u, times, T, x0.data(), S, &opt, nullptr, &plan);
    if (r != SDESIM_OK) fail(r, g_last_error);
    if (P == 0) return;
    const size_t vals = (size_t)T * P * S;
    size_t free_b = 0, total_b = 0;
    hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    const size_t need = (size_t)sdesim_plan_filtration_bytes(plan) + vals * sizeof(double);
    if (need + (64u << 20) > free_b)
      fail(SDESIM_ERUNTIME, "output of " + std::to_string(need) + " bytes does not fit in free device memory");
    hip_check(hipMalloc(&d_filt, (size_t)sdesim_plan_filtration_bytes(plan)), "hipMalloc(filtration)");
    hip_check(hipMalloc(&d_long, vals * sizeof(double)), "hipMalloc(long)");
    r = sdesim_plan_run(plan, d_filt, nullptr, stream);
    if (r != SDESIM_OK) fail(r, g_last_error);
    r = sdesim_filtration_to_long(d_filt, opt.precision, S, T, P, d_long, stream);
    if (r != SDESIM_OK) fail(r, g_last_error);
    hip_check(hipMemcpyAsync(h_values, d_long, vals * sizeof(double), hipMemcpyDeviceToHost, (hipStream_t)stream),
              "hipMemcpyAsync(values)");
    hip_check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
  });
  if (d_filt) (void)hipFree(d_filt);
  if (d_long) (void)hipFree(d_long);
  sdesim_plan_free(plan);
  return rc;
}

int sdesim_generate_source(const sdesim_universe* u, int32_t n_times, const sdesim_options* opt_in, const char* payoff,
                           char* buf, int64_t buf_len, int64_t* needed) {
  return guarded([&] {
    if (!u) fail(SDESIM_EVALUE, "null universe");
    sdesim_options opt;
    if (opt_in) opt = *opt_in; else sdesim_default_options(&opt);
    sdesim::CodegenOptions co;
    co.scheme = opt.scheme;
    co.compat = opt.compat;
    co.f32 = opt.precision == SDESIM_F32;
    co.rng = u->u.M() == 0 ? SDE_RNG_NONE
             : opt.rng == SDESIM_SOBOL ? (opt.scramble == SDESIM_SCRAMBLE_XOR ? SDE_RNG_SOBOL_XOR : SDE_RNG_SOBOL_SHIFT)
             : opt.rng == SDESIM_INJECTED ? SDE_RNG_INJECTED : SDE_RNG_PSEUDO;
    co.out = opt.outputs & (SDESIM_OUT_FILTRATION | SDESIM_OUT_TERMINAL);
    co.ppl = opt.paths_per_lane > 0 ? opt.paths_per_lane : 1;
    co.block = opt.block_size > 0 ? opt.block_size : 256;
    co.nsl = 3;
    co.uniform_grid = true;
    const long long first = opt.sobol_first >= 0 ? opt.sobol_first : (co.rng == SDE_RNG_SOBOL_XOR ? 0 : 5);
    co.wave_aligned = co.rng == SDE_RNG_SOBOL_XOR && first % 64 == 0 && co.block % 64 == 0;
    co.name = sdesim::kernel_name(co);
    co.steps = n_times > 1;
    co.payoff = payoff ? payoff : "";
    std::string src = sdesim::generate_kernel_source(u->u, co);
    if (needed) *needed = (int64_t)src.size() + 1;
    if (buf && buf_len > 0) std::snprintf(buf, (size_t)buf_len, "%s", src.c_str());
  });
}

void sdesim_kernel_cache_reset(void) {
  std::lock_guard<std::mutex> lock(g_cache_mu);
  g_kernels.clear();
}

int sdesim_compile_source(const char* source, int64_t* code_bytes) {
  return guarded([&] {
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<char> code = rtc_compile(source ? source : "");
    if (code_bytes) *code_bytes = (int64_t)code.size();
    sdesim_log("source compiled with hipRTC: %zu bytes of code in %.1f ms", code.size(),
               std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
  });
}

int sdesim_eval_expression(const char* expr, const char* const* names, const double* values, int32_t n,
                           double* out) {
  return guarded([&] {
    sdesim::ExprPtr e = sdesim::parse_expression(expr ? expr : "");
    std::map<std::string, double> ns;
    for (int32_t i = 0; i < n; ++i) ns[names[i]] = values[i];
    try {
      *out = sdesim::eval_host(*e, ns);
    } catch (const std::runtime_error& err) {
      fail(SDESIM_EPANIC, err.what());
    }
  });
}

int sdesim_sobol_points(int32_t dims, uint64_t first, int64_t n, uint32_t* out) {
  return guarded([&] {
    if (dims <= 0 || dims > sdesim_jk_dims) fail(SDESIM_EVALUE, "dims out of range");
    if (n < 0) fail(SDESIM_EVALUE, "negative count");
    const unsigned long long last = first + (unsigned long long)(n > 0 ? n - 1 : 0);
    const unsigned long long gl = last | (last >> 1);
    int bits = 0;
    while (bits < 64 && (gl >> bits) != 0ull) ++bits;
    if (bits > 32) fail(SDESIM_EPANIC, "Sobol sequence exhausted");
    int nsl = 1;
    while (6 + 8 * (nsl - 1) < bits) ++nsl;
    std::vector<unsigned int> tab;
    build_sobol_tables(dims, nsl, tab);
    const int stride = 64 + 256 * (nsl - 1);
    for (int64_t i = 0; i < n; ++i) {
      const unsigned long long idx = first + (unsigned long long)i;
      const unsigned long long g = idx ^ (idx >> 1);
      for (int d = 0; d < dims; ++d) {
        const unsigned int* row = &tab[(size_t)d * stride];
        unsigned int x = row[g & 63u];
        for (int k = 1; k < nsl; ++k) x ^= row[64 + 256 * (k - 1) + ((g >> (6 + 8 * (k - 1))) & 255u)];
        out[(size_t)i * dims + d] = x;
      }
    }
  });
}

const char* sdesim_last_error(void) { return g_last_error.c_str(); }
int32_t sdesim_version(void) { return 100; }

}  // extern "C"

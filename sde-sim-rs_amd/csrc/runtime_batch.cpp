// runtime_batch.cpp -- host side of batched instances (include/sdesim.h: sdesim_batch_*).
//
// The reference runs one universe per simulate() call (src/sim/mod.rs:20-92); here many
// independent 1-D instances with their own parameters, path counts, step counts, horizons and
// seeds run in one persistent launch (batch_kernels.hip). This file builds the work-item queue
// (longest first), sizes the persistent grid from occupancy, autotunes paths-per-lane and block
// size (timed on the workload itself, remembered per device and workload in memory and in a
// small text cache on disk), and captures reset + kernel + finalize into a hipGraph.
#include <hip/hip_runtime_api.h>
#include <sys/stat.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/sdesim.h"
#include "device/sdesim_args.hpp"

hipError_t sdesim_launch_batch(int f32, int ppl, int grid, int block, const SdeInstance* inst, const SdeItem* items,
                               int n_items, unsigned int* counter, double* partials, void* terminal,
                               const long long* term_offset, hipStream_t stream);
hipError_t sdesim_launch_batch_finalize(const SdeInstance* inst, const int* slot0, int n, const double* partials,
                                        double* out, hipStream_t stream);
hipError_t sdesim_batch_occupancy(int f32, int ppl, int block, int* blocks_per_cu);

// Error plumbing shared with runtime.cpp.
int sdesim_set_error(int code, const std::string& msg);
void sdesim_log(const char* fmt, ...);

namespace {

struct BatchFailure {
  int code;
  std::string msg;
};

void bcheck(hipError_t e, const char* what) {
  if (e != hipSuccess) throw BatchFailure{SDESIM_ERUNTIME, std::string(what) + ": " + hipGetErrorString(e)};
}

template <class F>
int bguard(F&& f) {
  try {
    f();
    return SDESIM_OK;
  } catch (const BatchFailure& e) {
    return sdesim_set_error(e.code, e.msg);
  } catch (const std::exception& e) {
    return sdesim_set_error(SDESIM_ERUNTIME, e.what());
  }
}

std::mutex g_tune_mu;
std::map<std::string, std::pair<int, int>> g_tuned;  // key -> (ppl, block)

std::string tune_cache_path() {
  const char* dir = std::getenv("SDESIM_CACHE_DIR");
  std::string d;
  if (dir && *dir) d = dir;
  else if (const char* home = std::getenv("HOME")) d = std::string(home) + "/.cache/sdesim";
  else return std::string();
  return d + "/autotune.txt";
}

bool tune_lookup(const std::string& key, int& ppl, int& block) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  auto it = g_tuned.find(key);
  if (it != g_tuned.end()) { ppl = it->second.first; block = it->second.second; return true; }
  std::string path = tune_cache_path();
  if (path.empty()) return false;
  std::ifstream in(path);
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    std::string k;
    int p, b;
    if (ls >> k >> p >> b && k == key && (p == 1 || p == 2 || p == 4) && (b == 64 || b == 128 || b == 256)) {
      g_tuned[key] = {p, b};
      ppl = p; block = b;
      return true;
    }
  }
  return false;
}

void tune_store(const std::string& key, int ppl, int block, double ms) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tuned[key] = {ppl, block};
  std::string path = tune_cache_path();
  if (path.empty()) return;
  const std::string dir = path.substr(0, path.rfind('/'));
  for (size_t i = 1; i <= dir.size(); ++i)
    if (i == dir.size() || dir[i] == '/') (void)::mkdir(dir.substr(0, i).c_str(), 0755);
  std::ofstream out(path, std::ios::app);
  if (out) out << key << " " << ppl << " " << block << " " << ms << "\n";
}

}  // namespace

struct sdesim_batch {
  int n = 0, f32 = 0;
  int ppl = 0, block = 0, grid = 0, n_items = 0;
  std::vector<SdeInstance> host;
  std::vector<long long> term_off;
  long long total_paths = 0;
  SdeInstance* d_inst = nullptr;
  SdeItem* d_items = nullptr;
  int* d_slot0 = nullptr;
  unsigned int* d_counter = nullptr;
  double* d_partials = nullptr;
  double* d_moments = nullptr;
  long long* d_term_off = nullptr;
  void* d_terminal = nullptr;
  bool keep_terminal = false;
  hipStream_t capture_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::string key;

  ~sdesim_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
    for (void* p : {(void*)d_inst, (void*)d_items, (void*)d_slot0, (void*)d_counter, (void*)d_partials,
                    (void*)d_moments, (void*)d_term_off, d_terminal})
      if (p) (void)hipFree(p);
  }

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }

  // (Re)build the work queue for a paths-per-lane value and size the persistent grid.
  void configure(int new_ppl, int new_block) {
    const long long per_item = 64LL * new_ppl;
    long long total_chunks = 0;
    for (int i = 0; i < n; ++i) total_chunks += (host[i].paths + per_item - 1) / per_item;
    if (total_chunks > (long long)INT32_MAX)
      throw BatchFailure{SDESIM_EVALUE, "too many work items (paths / (64 * paths_per_lane) summed over the batch)"};
    // Buffers below may be in use by queued launches of the previous configuration.
    if (d_items) bcheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
    drop_graph();
    ppl = new_ppl;
    block = new_block;
    std::vector<int> slot0(n + 1, 0);
    std::vector<SdeItem> items;
    for (int i = 0; i < n; ++i) {
      const long long chunks = (host[i].paths + per_item - 1) / per_item;
      slot0[i + 1] = slot0[i] + (int)chunks;
      for (long long c = 0; c < chunks; ++c) items.push_back(SdeItem{i, (int)c, slot0[i] + (int)c, 0});
    }
    // Longest processing time first: an item costs `steps` wave-steps whatever its live lanes.
    std::stable_sort(items.begin(), items.end(), [&](const SdeItem& a, const SdeItem& b) {
      return host[a.instance].steps > host[b.instance].steps;
    });
    n_items = (int)items.size();
    if (d_items) (void)hipFree(d_items);
    if (d_slot0) (void)hipFree(d_slot0);
    if (d_partials) (void)hipFree(d_partials);
    d_items = nullptr; d_slot0 = nullptr; d_partials = nullptr;
    bcheck(hipMalloc(&d_items, sizeof(SdeItem) * std::max(1, n_items)), "hipMalloc(items)");
    bcheck(hipMemcpy(d_items, items.data(), sizeof(SdeItem) * items.size(), hipMemcpyHostToDevice), "hipMemcpy(items)");
    bcheck(hipMalloc(&d_slot0, sizeof(int) * slot0.size()), "hipMalloc(slots)");
    bcheck(hipMemcpy(d_slot0, slot0.data(), sizeof(int) * slot0.size(), hipMemcpyHostToDevice), "hipMemcpy(slots)");
    bcheck(hipMalloc(&d_partials, sizeof(double) * 2 * std::max(1, n_items)), "hipMalloc(partials)");
    int dev = 0, cus = 0, per_cu = 0;
    bcheck(hipGetDevice(&dev), "hipGetDevice");
    bcheck(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    bcheck(sdesim_batch_occupancy(f32, ppl, block, &per_cu), "occupancy");
    grid = std::max(1, cus * std::max(1, per_cu));
    const long long waves = (long long)grid * (block / 64);
    if (waves > n_items) grid = (int)std::max<long long>(1, (n_items + block / 64 - 1) / (block / 64));
  }

  void launch(hipStream_t s) {
    bcheck(sdesim_launch_batch(f32, ppl, grid, block, d_inst, d_items, n_items, d_counter, d_partials,
                               keep_terminal ? d_terminal : nullptr, d_term_off, s),
           "batch launch");
    bcheck(sdesim_launch_batch_finalize(d_inst, d_slot0, n, d_partials, d_moments, s), "batch finalize");
  }
};

extern "C" {

int sdesim_batch_create(const sdesim_instance* inst, int32_t n, int32_t precision, int32_t keep_terminal,
                        sdesim_batch** out) {
  return bguard([&] {
    if (!out || (!inst && n > 0)) throw BatchFailure{SDESIM_EVALUE, "null argument"};
    if (n <= 0) throw BatchFailure{SDESIM_EVALUE, "a batch needs at least one instance"};
    std::unique_ptr<sdesim_batch> b(new sdesim_batch());
    b->n = n;
    b->f32 = precision == SDESIM_F32 ? 1 : 0;
    b->keep_terminal = keep_terminal != 0;
    b->host.resize(n);
    b->term_off.resize(n + 1, 0);
    unsigned long long sig = 1469598103934665603ULL;
    auto mix = [&](unsigned long long v) { sig = (sig ^ v) * 1099511628211ULL; };
    for (int i = 0; i < n; ++i) {
      const sdesim_instance& a = inst[i];
      if (a.kind < 0 || a.kind > 2) throw BatchFailure{SDESIM_EVALUE, "instance " + std::to_string(i) + ": unknown model"};
      if (a.steps < 1) throw BatchFailure{SDESIM_EVALUE, "instance " + std::to_string(i) + ": steps must be >= 1"};
      if (a.paths < 1) throw BatchFailure{SDESIM_EVALUE, "instance " + std::to_string(i) + ": paths must be >= 1"};
      if (!(a.horizon > 0.0)) throw BatchFailure{SDESIM_EVALUE, "instance " + std::to_string(i) + ": horizon must be > 0"};
      SdeInstance& h = b->host[i];
      h.kind = a.kind;
      h.steps = a.steps;
      h.paths = a.paths;
      h.p0 = a.p0; h.p1 = a.p1; h.p2 = a.p2;
      h.x0 = a.x0;
      h.dt = a.horizon / (double)a.steps;
      h.sdt = std::sqrt(h.dt);
      h.dtf = (float)h.dt;
      h.sdtf = (float)h.sdt;
      h.seed = a.seed;
      b->term_off[i + 1] = b->term_off[i] + a.paths;
      mix((unsigned long long)a.kind); mix((unsigned long long)a.steps); mix((unsigned long long)a.paths);
    }
    b->total_paths = b->term_off[n];
    long long items1 = 0;  // work items at one path per lane (the most)
    for (int i = 0; i < n; ++i) items1 += (inst[i].paths + 63) / 64;
    if (items1 > (long long)INT32_MAX)
      throw BatchFailure{SDESIM_EVALUE, "too many work items (paths / (64 * paths_per_lane) summed over the batch)"};
    char kb[64];
    std::snprintf(kb, sizeof kb, "%s-%d-%016llx", b->f32 ? "f32" : "f64", n, sig);
    b->key = kb;
    bcheck(hipMalloc(&b->d_inst, sizeof(SdeInstance) * n), "hipMalloc(instances)");
    bcheck(hipMemcpy(b->d_inst, b->host.data(), sizeof(SdeInstance) * n, hipMemcpyHostToDevice), "hipMemcpy");
    bcheck(hipMalloc(&b->d_term_off, sizeof(long long) * (n + 1)), "hipMalloc(offsets)");
    bcheck(hipMemcpy(b->d_term_off, b->term_off.data(), sizeof(long long) * (n + 1), hipMemcpyHostToDevice), "hipMemcpy");
    bcheck(hipMalloc(&b->d_counter, sizeof(unsigned int)), "hipMalloc(counter)");
    bcheck(hipMalloc(&b->d_moments, sizeof(double) * 2 * n), "hipMalloc(moments)");
    if (b->keep_terminal) {
      bcheck(hipMalloc(&b->d_terminal, (size_t)b->total_paths * (b->f32 ? 4 : 8)), "hipMalloc(terminal)");
    }
    b->configure(1, 256);
    *out = b.release();
  });
}

void sdesim_batch_free(sdesim_batch* b) { delete b; }

int sdesim_batch_set_config(sdesim_batch* b, int32_t paths_per_lane, int32_t block_size) {
  return bguard([&] {
    if (!b) throw BatchFailure{SDESIM_EVALUE, "null batch"};
    if (paths_per_lane != 1 && paths_per_lane != 2 && paths_per_lane != 4)
      throw BatchFailure{SDESIM_EVALUE, "paths_per_lane must be 1, 2 or 4"};
    if (block_size != 64 && block_size != 128 && block_size != 256)
      throw BatchFailure{SDESIM_EVALUE, "block_size must be 64, 128 or 256"};
    b->configure(paths_per_lane, block_size);
  });
}

int sdesim_batch_autotune(sdesim_batch* b, void* stream, int32_t* ppl_out, int32_t* block_out, double* ms_out) {
  return bguard([&] {
    if (!b) throw BatchFailure{SDESIM_EVALUE, "null batch"};
    hipStream_t s = (hipStream_t)stream;
    int best_p = 1, best_b = 256;
    double best = -1.0;
    int cp, cb;
    if (tune_lookup(b->key, cp, cb)) {
      best_p = cp;
      best_b = cb;
    } else {
      struct Events {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Events() {
          if (e0) (void)hipEventDestroy(e0);
          if (e1) (void)hipEventDestroy(e1);
        }
      } ev;
      bcheck(hipEventCreate(&ev.e0), "hipEventCreate");
      bcheck(hipEventCreate(&ev.e1), "hipEventCreate");
      hipEvent_t e0 = ev.e0, e1 = ev.e1;
      for (int p : {1, 2, 4}) {
        for (int blk : {64, 128, 256}) {
          b->configure(p, blk);
          b->launch(s);  // warm
          float ms_best = 1e30f;
          for (int rep = 0; rep < 2; ++rep) {
            bcheck(hipEventRecord(e0, s), "hipEventRecord");
            b->launch(s);
            bcheck(hipEventRecord(e1, s), "hipEventRecord");
            bcheck(hipEventSynchronize(e1), "hipEventSynchronize");
            float ms = 0.f;
            bcheck(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
            ms_best = std::min(ms_best, ms);
          }
          if (best < 0 || ms_best < best) { best = ms_best; best_p = p; best_b = blk; }
        }
      }
      tune_store(b->key, best_p, best_b, best);
    }
    b->configure(best_p, best_b);
    sdesim_log("batch autotune: %d path(s) per lane, block %d (%s, %.3f ms)", best_p, best_b,
               best < 0 ? "remembered" : "timed", best);
    if (ppl_out) *ppl_out = best_p;
    if (block_out) *block_out = best_b;
    if (ms_out) *ms_out = best;
  });
}

int sdesim_batch_run(sdesim_batch* b, void* stream) {
  return bguard([&] {
    if (!b) throw BatchFailure{SDESIM_EVALUE, "null batch"};
    b->launch((hipStream_t)stream);
  });
}

int sdesim_batch_capture(sdesim_batch* b) {
  return bguard([&] {
    if (!b) throw BatchFailure{SDESIM_EVALUE, "null batch"};
    b->drop_graph();
    if (!b->capture_stream) bcheck(hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking), "hipStreamCreate");
    bcheck(hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
      b->launch(b->capture_stream);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(b->capture_stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    bcheck(hipStreamEndCapture(b->capture_stream, &b->graph), "hipStreamEndCapture");
    bcheck(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  });
}

int sdesim_batch_replay(sdesim_batch* b, void* stream) {
  return bguard([&] {
    if (!b || !b->exec) throw BatchFailure{SDESIM_EVALUE, "batch has no captured graph"};
    bcheck(hipGraphLaunch(b->exec, (hipStream_t)stream), "hipGraphLaunch");
  });
}

int sdesim_batch_moments(sdesim_batch* b, double* h_out, void* stream) {
  return bguard([&] {
    if (!b || !h_out) throw BatchFailure{SDESIM_EVALUE, "null argument"};
    bcheck(hipMemcpyAsync(h_out, b->d_moments, sizeof(double) * 2 * b->n, hipMemcpyDeviceToHost, (hipStream_t)stream),
           "hipMemcpyAsync");
    bcheck(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
  });
}

int sdesim_batch_terminal(sdesim_batch* b, void* h_out, void* stream) {
  return bguard([&] {
    if (!b || !h_out) throw BatchFailure{SDESIM_EVALUE, "null argument"};
    if (!b->keep_terminal) throw BatchFailure{SDESIM_EVALUE, "batch was created without terminal values"};
    bcheck(hipMemcpyAsync(h_out, b->d_terminal, (size_t)b->total_paths * (b->f32 ? 4 : 8), hipMemcpyDeviceToHost,
                          (hipStream_t)stream),
           "hipMemcpyAsync");
    bcheck(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
  });
}

int sdesim_batch_info(const sdesim_batch* b, int32_t* ppl, int32_t* block, int32_t* grid, int32_t* n_items,
                      int64_t* total_paths) {
  if (!b) return SDESIM_EVALUE;
  if (ppl) *ppl = b->ppl;
  if (block) *block = b->block;
  if (grid) *grid = b->grid;
  if (n_items) *n_items = b->n_items;
  if (total_paths) *total_paths = b->total_paths;
  return SDESIM_OK;
}

}  // extern "C"

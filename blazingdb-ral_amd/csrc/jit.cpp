/*
 * jit.cpp — hiprtc compilation and module cache of generated kernels (jit.h).
 *
 * The generated source includes device_common.h and program.h; their text is
 * linked into the library (the build writes it into rtc_headers.cpp) and handed to
 * hiprtc as in-memory headers, so the run-time compiled kernels and the kernels
 * built ahead of time share every device helper.
 */
#include "jit.h"

#include "codegen.h"

#include <hip/hiprtc.h>

#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <unordered_map>
#include <vector>

namespace blz {

extern const char kRtcDeviceCommon[];
extern const char kRtcProgram[];

namespace {

struct Entry {
  std::mutex mu;
  bool built = false;
  std::vector<char> code;                 // code object for gfx950
  std::map<int, hipModule_t> modules;     // per device
  std::map<std::pair<int, std::string>, hipFunction_t> functions;  // per device and name
  // Compiled as part of a combined program: the entry holding its code object, and
  // the suffix of its kernels' names there.
  std::shared_ptr<Entry> shared;
  std::string suffix;
};

std::mutex g_mu;
std::unordered_map<std::string, std::shared_ptr<Entry>> g_cache;
long long g_compiles = 0, g_hits = 0;
double g_compile_ms = 0;

// -ffp-contract=off: each op rounds once, as the ahead-of-time kernels do;
// denormals kept (float32 inputs widen exactly); no packed-FP32 forms.
const char *kOpts[] = {"--offload-arch=gfx950",
                       "-O3",
                       "-std=c++17",
                       "-ffp-contract=off",
                       "-fno-gpu-flush-denormals-to-zero",
                       "-Xclang",
                       "-target-feature",
                       "-Xclang",
                       "-packed-fp32-ops"};

void compile(Entry &e, const std::string &src) {
  const auto t0 = std::chrono::steady_clock::now();
  hiprtcProgram prog;
  const char *headers[] = {kRtcDeviceCommon, kRtcProgram};
  const char *names[] = {"device_common.h", "program.h"};
  if (hiprtcCreateProgram(&prog, src.c_str(), "blz_compiled.hip", 2, headers, names) != HIPRTC_SUCCESS)
    throw std::runtime_error("hiprtcCreateProgram failed");
  const hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof(kOpts) / sizeof(kOpts[0])), kOpts);
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n + 1, '\0');
    hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    throw std::runtime_error(std::string("compiled tier: hiprtc: ") + hiprtcGetErrorString(r) + "\n" +
                             log.substr(0, 4000));
  }
  size_t n = 0;
  hiprtcGetCodeSize(prog, &n);
  e.code.resize(n);
  hiprtcGetCode(prog, e.code.data());
  hiprtcDestroyProgram(&prog);
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  std::lock_guard<std::mutex> g(g_mu);
  g_compiles++;
  g_compile_ms += ms;
}

// The module of `e`'s code object on device `dev`, loaded on first use (e.mu held).
hipModule_t module_on(Entry &e, int dev) {
  auto m = e.modules.find(dev);
  if (m == e.modules.end()) {
    hipModule_t mod = nullptr;
    const hipError_t err = hipModuleLoadData(&mod, e.code.data());
    if (err != hipSuccess)
      throw std::runtime_error(std::string("compiled tier: hipModuleLoadData: ") + hipGetErrorString(err));
    m = e.modules.emplace(dev, mod).first;
  }
  return m->second;
}

}  // namespace

hipFunction_t jit_function(const std::string &src, const std::string &name) {
  std::shared_ptr<Entry> e;
  {
    std::lock_guard<std::mutex> g(g_mu);
    auto &slot = g_cache[src];
    if (!slot) slot = std::make_shared<Entry>();
    else g_hits++;
    e = slot;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) throw std::runtime_error("compiled tier: no current device");
  std::lock_guard<std::mutex> g(e->mu);
  if (!e->built) {
    compile(*e, src);
    e->built = true;
  }
  auto it = e->functions.find({dev, name});
  if (it != e->functions.end()) return it->second;
  const std::string sym = name + e->suffix;
  hipFunction_t f = nullptr;
  hipError_t err = hipSuccess;
  if (e->shared) {
    std::lock_guard<std::mutex> g2(e->shared->mu);
    err = hipModuleGetFunction(&f, module_on(*e->shared, dev), sym.c_str());
  } else {
    err = hipModuleGetFunction(&f, module_on(*e, dev), sym.c_str());
  }
  if (err != hipSuccess)
    throw std::runtime_error(std::string("compiled tier: hipModuleGetFunction(") + sym + "): " + hipGetErrorString(err));
  e->functions[{dev, name}] = f;
  return f;
}

void jit_compile_many(const std::vector<std::string> &srcs) {
  const std::string pre = codegen_prelude();
  std::vector<std::pair<const std::string *, std::shared_ptr<Entry>>> todo;
  std::vector<const std::string *> alone;
  {
    std::lock_guard<std::mutex> g(g_mu);
    for (const std::string &src : srcs) {
      auto &slot = g_cache[src];
      if (!slot) slot = std::make_shared<Entry>();
      if (slot->built) continue;
      if (src.compare(0, pre.size(), pre) == 0) todo.push_back({&src, slot});
      else alone.push_back(&src);
    }
  }
  auto each_alone = [](const std::string &src) {
    try {
      jit_compile_only(src);
    } catch (const std::exception &) {  // reported by jit_function when it is used
    }
  };
  for (const std::string *src : alone) each_alone(*src);
  if (todo.size() < 2) {
    for (auto &t : todo) each_alone(*t.first);
    return;
  }
  std::string all = pre;
  for (size_t k = 0; k < todo.size(); k++) {
    std::string body = todo[k].first->substr(pre.size());
    const std::string suffix = "_m" + std::to_string(k);
    for (const char *kn : {"blz_cg_project", "blz_cg_pred", "blz_cg_gather"}) {
      const std::string from = std::string("__launch_bounds__(256) ") + kn + "(";
      const std::string to = std::string("__launch_bounds__(256) ") + kn + suffix + "(";
      for (size_t at = body.find(from); at != std::string::npos; at = body.find(from, at + to.size()))
        body.replace(at, from.size(), to);
    }
    all += "#undef RS\n#undef GR\n" + body;
  }
  auto shared = std::make_shared<Entry>();
  try {
    compile(*shared, all);
    shared->built = true;
  } catch (const std::exception &) {
    for (auto &t : todo) each_alone(*t.first);
    return;
  }
  for (size_t k = 0; k < todo.size(); k++) {
    Entry &e = *todo[k].second;
    std::lock_guard<std::mutex> g(e.mu);
    if (e.built) continue;
    e.shared = shared;
    e.suffix = "_m" + std::to_string(k);
    e.built = true;
  }
}

void jit_compile_only(const std::string &src) {
  std::shared_ptr<Entry> e;
  {
    std::lock_guard<std::mutex> g(g_mu);
    auto &slot = g_cache[src];
    if (!slot) slot = std::make_shared<Entry>();
    else g_hits++;
    e = slot;
  }
  std::lock_guard<std::mutex> g(e->mu);
  if (!e->built) {
    compile(*e, src);
    e->built = true;
  }
}

int jit_blocks_per_cu(hipFunction_t f) {
  int b = 0;
  if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&b, f, 256, 0) != hipSuccess) return 1;
  return b > 0 ? b : 1;
}

hipError_t launch_jit(hipFunction_t f, int grid, hipStream_t st, const uint8_t *blob, const Workspace &ws) {
  const uint8_t *b = blob;
  Workspace w = ws;
  void *args[] = {(void *)&b, (void *)&w};
  return hipModuleLaunchKernel(f, (unsigned)grid, 1, 1, 256, 1, 1, 0, st, args, nullptr);
}

void jit_counters(long long *compiles, long long *hits, double *compile_ms) {
  std::lock_guard<std::mutex> g(g_mu);
  if (compiles) *compiles = g_compiles;
  if (hits) *hits = g_hits;
  if (compile_ms) *compile_ms = g_compile_ms;
}

}  // namespace blz

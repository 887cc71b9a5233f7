/*
 * capi.cpp — the extern "C" entry points of include/blz/blz_ral.h.
 *
 * Each call: plan (host) -> compile to one device program blob -> upload blob and
 * zero a small workspace -> one kernel launch -> read back counts. Errors never
 * escape as exceptions: they become a status code plus message, the C form of the
 * reference's std::runtime_error convention (Utils.cuh:25-55).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <unordered_map>
#include <thread>
#include <deque>
#include <condition_variable>
#include <chrono>
#include <map>
#include <set>
#include <functional>
#include <mutex>
#include <tuple>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/blz/blz_ral.h"
#include "device_arena.h"
#include "jit.h"
#include "kernels.h"
#include "planner.h"
#include "program_compiler.h"
#include "result_sets.h"

using namespace blz;

namespace {

struct BlzError : std::runtime_error {
  int code;
  BlzError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return;
  if (e == hipErrorOutOfMemory) throw BlzError(BLZ_ERR_OOM, std::string(what) + ": out of device memory");
  throw BlzError(BLZ_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPCHECK(x) hip_check((x), #x)

void invalid(const std::string &m) { throw BlzError(BLZ_ERR_INVALID, m); }

int write_err(char *err, size_t len, const std::string &m) {
  if (err && len > 0) {
    size_t n = std::min(len - 1, m.size());
    std::memcpy(err, m.data(), n);
    err[n] = 0;
  }
  return 0;
}

template <class F>
int guarded(char *err, size_t len, F &&f) {
  // a failure left by an earlier call (hipGetLastError reports it once) must not be
  // taken for a failure of this call's kernel launches
  (void)hipGetLastError();
  try {
    f();
    write_err(err, len, "");
    return BLZ_OK;
  } catch (const BlzError &e) {
    write_err(err, len, e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    write_err(err, len, "host out of memory");
    return BLZ_ERR_OOM;
  } catch (const std::exception &e) {
    write_err(err, len, e.what());
    const std::string m = e.what();
    if (m.find("not supported") != std::string::npos || m.find("unsupported") != std::string::npos)
      return BLZ_ERR_UNSUPPORTED;
    return BLZ_ERR_INVALID;
  }
}

hipStream_t S(blz_stream_t s) { return (hipStream_t)s; }

int device_cus() {
  static int cus = 0;
  static std::once_flag once;
  std::call_once(once, [] {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 256;
  });
  return cus;
}

// Stream-ordered allocation from the current device's default pool (left at its
// default release threshold: keeping freed memory in the pool, 8 GiB of it, made
// a run of 256 one-shot projections 21x slower).
void *pool_alloc(size_t bytes, hipStream_t st) {
  void *p = nullptr;
  HIPCHECK(hipMallocAsync(&p, bytes, st));
  return p;
}

// Launch buffers of the blocking one-shot calls, kept per (device, stream) and grown
// on demand. The stream-ordered pool hands freed memory back at every
// synchronisation, so a 100M-row filter mapped its ~400 MB of kept-row index space
// afresh on every call (0.38 ms of a 1.04 ms call). A lease holds its slot for the
// whole call (calls on one stream from several threads take turns); a buffer larger
// than kLbCacheMax is allocated for the call alone.
constexpr size_t kLbCacheMax = (size_t)4 << 30;
// Pinned host words per slot for the calls' readbacks (kept count, null counts):
// a device-to-host copy into pinned memory needs no staging copy on the host.
constexpr int kLbHostWords = 128;
struct LbSlot {
  uint8_t *p = nullptr;
  size_t bytes = 0;
  unsigned long long *host = nullptr;
  std::mutex busy;
};
std::mutex g_lb_mu;
std::map<std::pair<int, hipStream_t>, std::unique_ptr<LbSlot>> g_lb;

class LbLease {
 public:
  LbLease(size_t bytes, hipStream_t st) : st_(st) {
    if (bytes > kLbCacheMax) {
      p_ = (uint8_t *)pool_alloc(bytes, st);
      return;
    }
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    {
      std::lock_guard<std::mutex> g(g_lb_mu);
      auto &slot = g_lb[{dev, st}];
      if (!slot) slot.reset(new LbSlot());
      slot_ = slot.get();
    }
    lock_ = std::unique_lock<std::mutex>(slot_->busy);
    if (!slot_->host &&
        hipHostMalloc((void **)&slot_->host, sizeof(unsigned long long) * kLbHostWords, hipHostMallocDefault) !=
            hipSuccess)
      slot_->host = nullptr;  // readbacks go through pageable memory
    if (slot_->bytes < bytes) {  // the slot's previous call has completed (calls block)
      if (slot_->p) HIPCHECK(hipFree(slot_->p));
      slot_->p = nullptr;
      slot_->bytes = 0;
      HIPCHECK(hipMalloc((void **)&slot_->p, bytes));
      slot_->bytes = bytes;
    }
    p_ = slot_->p;
  }
  // The call's work on st_ has completed (or the stream was synchronised) by now.
  ~LbLease() {
    if (!slot_ && p_) (void)hipFreeAsync(p_, st_);
  }
  LbLease(const LbLease &) = delete;
  LbLease &operator=(const LbLease &) = delete;
  uint8_t *get() const { return p_; }
  // Pinned readback words for `words` values, or `fallback` (pageable) when the
  // lease has none or too few.
  unsigned long long *host(size_t words, unsigned long long *fallback) const {
    return slot_ && slot_->host && words <= (size_t)kLbHostWords ? slot_->host : fallback;
  }

 private:
  hipStream_t st_;
  LbSlot *slot_ = nullptr;
  uint8_t *p_ = nullptr;
  std::unique_lock<std::mutex> lock_;
};

// A device address for the stand-in output pointers of a sized filter's first
// compilation (the predicate and the slot scan never read or write outputs): 256
// bytes per device, allocated once and kept for the process.
uint8_t *stand_in_address() {
  static std::mutex mu;
  static std::map<int, uint8_t *> addr;
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(mu);
  uint8_t *&p = addr[dev];
  if (!p) HIPCHECK(hipMalloc((void **)&p, 256));
  return p;
}

// Frees the cached launch buffers of `st` (on every device) when the library's
// stream is destroyed.
void lb_cache_drop(hipStream_t st) {
  std::lock_guard<std::mutex> g(g_lb_mu);
  for (auto it = g_lb.begin(); it != g_lb.end();) {
    if (it->first.second != st) { ++it; continue; }
    {
      std::lock_guard<std::mutex> b(it->second->busy);
      if (it->second->p) (void)hipFree(it->second->p);
      if (it->second->host) (void)hipHostFree(it->second->host);
    }
    it = g_lb.erase(it);
  }
}

int64_t rows_per_tile(int r) { return (int64_t)block_threads() * r; }

int64_t tile_rows(const Compiled &c) { return c.T > 0 ? c.T : rows_per_tile(c.R); }

// Workgroups per CU of a tile-VM kernel, cached per (shape, mode, LDS).
int vm_per_cu(const Compiled &c, int mode, int lds) {
  static std::mutex mu;
  static std::map<std::tuple<int, int, int, int, int>, int> cache;
  const auto key = std::make_tuple(c.R, c.S, mode, c.vm_slow, lds);
  std::lock_guard<std::mutex> g(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  const int b = std::max(1, vm_blocks_per_cu(c.R, mode, c.vm_slow, lds));
  cache[key] = b;
  return b;
}

// Persistent grid: every workgroup that fits on the device at once.
int interp_grid(const Compiled &c, bool compact, int64_t n) {
  const int64_t tiles = (n + tile_rows(c) - 1) / tile_rows(c);
  const int per_cu = c.vm ? vm_per_cu(c, compact ? 1 : 0, c.lds_bytes)
                     : interp_blocks_per_cu(c.S, c.R, compact, c.lds_bytes);
  const int64_t resident = (int64_t)device_cus() * per_cu;
  return (int)std::max<int64_t>(1, std::min<int64_t>(tiles, resident));
}

void check_column(const gdf_column *c, const char *what, int64_t n) {
  if (!c) invalid(std::string(what) + " is NULL");
  if (c->size < 0) invalid(std::string(what) + " has a negative size");
  if (n >= 0 && c->size != n) invalid(std::string(what) + " size differs from the input row count");
  const int w = blz_dtype_width(c->dtype);
  if (c->size > 0 && w > 0 && c->data == nullptr) invalid(std::string(what) + " has no data buffer");
  if (w > 0 && ((uintptr_t)c->data % w) != 0) invalid(std::string(what) + " data is not aligned to its width");
  if (c->valid && ((uintptr_t)c->valid % 8) != 0) invalid(std::string(what) + " validity is not 8-byte aligned");
}

// Device scratch of one launch:
//   blob | VmBatch + VmPlanRef | null counts | counters | tile states  (zeroed part
//   starts at the null counts) | keep words | slot counts | kept-row indices
struct LaunchBuffers {
  uint8_t *base = nullptr;
  size_t bytes = 0;
  size_t bd_off = 0;
  size_t zero_off = 0, zero_bytes = 0;
  size_t keep_off = 0, count_off = 0, idx_off = 0, side_off = 0;
  Workspace ws{};
  uint8_t *blob = nullptr;
  VmBatch *bd = nullptr;  // device address of the launch descriptor
  int64_t n_chunks = 0;   // slot-scan chunks
};

size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

void layout(LaunchBuffers &lb, const Compiled &c, int n_out, bool compact, int64_t n) {
  const int64_t n_tiles = compact ? (n + tile_rows(c) - 1) / tile_rows(c) : 0;
  const size_t blob_sz = round_up(c.blob.size(), 256);
  lb.bd_off = blob_sz;
  const size_t bd_sz = round_up(sizeof(VmBatch) + sizeof(VmPlanRef), 256);
  const size_t nc = round_up(sizeof(unsigned long long) * std::max(n_out, 1), 64);
  const size_t cnt = 64;  // tile / chunk counter + n_sel
  const size_t ts = round_up(sizeof(unsigned long long) * std::max<int64_t>(n_tiles, 1), 64);
  lb.zero_off = blob_sz + bd_sz;
  lb.zero_bytes = nc + cnt + ts;
  size_t off = lb.zero_off + lb.zero_bytes;
  const bool vmf = (c.vm || c.jit) && compact;  // predicate, slot scan, gather
  const int64_t slots = vmf ? n_tiles * 4 : 0;
  lb.n_chunks = (slots + scan_chunk_slots(c.T) - 1) / scan_chunk_slots(c.T);
  lb.keep_off = off;
  off += round_up(vmf ? sizeof(uint64_t) * n_tiles * (c.T / 64) : 0, 256);
  lb.count_off = off;
  off += round_up(sizeof(uint32_t) * slots, 256);
  lb.idx_off = off;
  off += round_up(vmf ? sizeof(uint32_t) * n : 0, 256);
  lb.side_off = off;  // compiled filter: staged predicate columns, n_tiles * T rows each
  off += round_up(vmf && c.jit ? (size_t)c.jit_side_bytes * n_tiles * c.T : 0, 256);
  lb.bytes = std::max<size_t>(off, 256);
}
void bind(LaunchBuffers &lb, int n_out) {
  uint8_t *p = lb.base;
  lb.blob = p;
  lb.bd = (VmBatch *)(p + lb.bd_off);
  uint8_t *z = p + lb.zero_off;
  lb.ws.null_counts = (unsigned long long *)z;
  const size_t nc = round_up(sizeof(unsigned long long) * std::max(n_out, 1), 64);
  lb.ws.tile_counter = (unsigned int *)(z + nc);
  lb.ws.n_sel = (unsigned long long *)(z + nc + 8);
  lb.ws.tile_state = (unsigned long long *)(z + nc + 64);
  lb.ws.keep = (uint64_t *)(p + lb.keep_off);
  lb.ws.wave_count = (uint32_t *)(p + lb.count_off);
  lb.ws.idx = (uint32_t *)(p + lb.idx_off);
  lb.ws.side = p + lb.side_off;
}
// Host image of the blob and the single-program launch descriptor, copied to
// lb.blob in one transfer.
std::vector<uint8_t> device_image(const LaunchBuffers &lb, const Compiled &c) {
  std::vector<uint8_t> img(lb.zero_off, 0);
  std::memcpy(img.data(), c.blob.data(), c.blob.size());
  if (c.vm || (c.jit && c.T > 0)) {  // the tile VM's kernels, or a compiled filter's slot scan
    const int64_t tiles = (c.n_rows + c.T - 1) / c.T;
    VmBatch b{};
    b.n_plans = 1;
    b.bufb = c.vm_bufb;
    b.tmp_off = c.vm_tmp_off;
    b.g_buf_off = c.vm_g_buf_off;
    b.g_tmp_off = c.vm_g_tmp_off;
    b.g_cpre_off = c.vm_g_cpre_off;
    b.g_wave_bytes = c.vm_g_wave_bytes;
    b.vr_off = c.vm_vr_off;
    b.g_vr_off = c.vm_g_vr_off;
    b.vr_wave_words = c.n_vslots * c.R;
    b.n_tiles = tiles;
    b.n_chunks = lb.n_chunks;
    b.chunk_counter = lb.ws.tile_counter;
    b.plans = (const VmPlanRef *)(lb.base + lb.bd_off + sizeof(VmBatch));
    VmPlanRef r{};
    r.blob = lb.blob;
    r.ws = lb.ws;
    std::memcpy(img.data() + lb.bd_off, &b, sizeof(b));
    std::memcpy(img.data() + lb.bd_off + sizeof(VmBatch), &r, sizeof(r));
  }
  return img;
}

// Kernel launches of one execution of a program: 3 for a tile-VM or compiled
// filter (predicate, slot scan, gather), 1 otherwise.
int program_stages(const Compiled &c, bool compact) { return (c.vm || c.jit) && compact ? 3 : 1; }

// Enqueues the program's kernels. evs (optional, program_stages + 1 events) are
// recorded before each kernel and after the last.
// sel_hint: rows a compiled filter kept the last time it ran (-1 unknown) — it sizes
// the gather's grid, never its result.
// stages (a tile-VM filter only): bit 0 the predicate, bit 1 the slot scan, bit 2
// the gather, so a caller can read the kept count between the scan and the gather.
hipError_t launch_program(const Compiled &c, bool compact, int grid, hipStream_t st, const uint8_t *blob,
                          const Workspace &ws, const VmBatch *bd, int64_t n_chunks, hipEvent_t *evs = nullptr,
                          int64_t sel_hint = -1, int stages = 7) {
  auto mark = [&](int i) { return evs ? hipEventRecord(evs[i], st) : hipSuccess; };
  hipError_t e = mark(0);
  if (e != hipSuccess) return e;
  if (c.jit) {
    const hipFunction_t f = jit_function(c.jit_src, c.jit_name);
    if (!compact) {
      // Projection: blocks of four 64 R-row units per wave, 8 workgroups per CU at most
      // (the 100M-row wide expression: 1.50 / 1.49 / 1.52 / 2.00 / 3.33 ms at 4 / 8 /
      // 16 / 32 / 64 per CU: a read and write stream runs best with few workgroups
      // per CU walking many blocks).
      const int64_t units = (c.n_rows + 64 * c.R - 1) / (64 * c.R);
      const int64_t waves = (units + 3) / 4;
      const int64_t g = std::min<int64_t>((waves + 3) / 4, (int64_t)device_cus() * 8);
      e = launch_jit(f, (int)std::max<int64_t>(1, g), st, blob, ws);
      return e == hipSuccess ? mark(1) : e;
    }
    // Filter: the predicate kernel over blocks of four slots per wave (32 workgroups
    // per CU at most), the slot scan, then the gather of the kept rows by as many
    // workgroups as stay resident (the kept count is known on the device only).
    const hipFunction_t fg = jit_function(c.jit_src, "blz_cg_gather");
    const int64_t slots = ((c.n_rows + c.T - 1) / c.T) * 4;
    const int64_t g = std::min<int64_t>((slots + 15) / 16, (int64_t)device_cus() * 32);
    e = launch_jit(f, (int)std::max<int64_t>(1, g), st, blob, ws);
    if (e == hipSuccess) e = mark(1);
    if (e == hipSuccess) e = launch_slot_scan(bd, n_chunks, st);
    if (e == hipSuccess) e = mark(2);
    if (e == hipSuccess) {
      // The kept count is known on the device only. Without a hint: as many
      // workgroups as stay resident (each walks chunks of 256 outputs per wave). With
      // the last run's count: one workgroup per 1024 outputs, between the resident
      // count and 256 per CU — many kept rows run faster with more workgroups in
      // flight than stay resident (500M rows at s = 0.5: 8.3 -> 6.7 ms), few rows
      // slower (C1's 2.3%: 0.070 ms at the resident count, 0.129 at 256 per CU).
      // Outputs with validity add one null-count atomic per workgroup each, and
      // atomics on one address serialize: then 8 workgroups per CU at most.
      const int64_t chunks = (c.n_rows + 255) / 256;  // most gather chunks (every row kept)
      const int64_t resident = (int64_t)device_cus() * jit_blocks_per_cu(fg);
      int64_t gg = std::min<int64_t>((chunks + 3) / 4, resident);
      if (sel_hint >= 0)
        gg = std::max<int64_t>(gg, std::min<int64_t>((sel_hint + 1023) / 1024, (int64_t)device_cus() * 256));
      if (c.jit_null_outputs > 0) gg = std::min<int64_t>(gg, (int64_t)device_cus() * 8);
      e = launch_jit(fg, (int)std::max<int64_t>(1, gg), st, blob, ws);
    }
    return e == hipSuccess ? mark(3) : e;
  }
  if (c.vm) {
    const int64_t tiles = (c.n_rows + c.T - 1) / c.T;
    const int64_t chunks = (c.n_rows + 64 * c.R - 1) / (64 * c.R);  // most gather chunks (all rows kept)
    // Predicate tiles: 8x the resident workgroups, so CUs that finish early take
    // queued ones (a general predicate over 100M rows 0.294 -> 0.254 ms). Projection
    // tiles and the gather of kept rows stay at the resident count (8x made a mix of
    // 256 projections 11% slower; 2x and 4x slowed the gather).
    auto grid_for = [&](int mode, int lds, int64_t units) {
      const int64_t resident = (int64_t)device_cus() * vm_per_cu(c, mode, lds) * (mode == 1 ? 8 : 1);
      return (int)std::max<int64_t>(1, std::min<int64_t>(units, resident));
    };
    if (!compact) {
      e = launch_vm(c.R, 0, c.vm_slow, grid_for(0, c.lds_bytes, tiles), c.lds_bytes, st, bd);
      return e == hipSuccess ? mark(1) : e;
    }
    if (!(stages & 1)) {
    } else if (c.vm_direct_pred) {  // 32 workgroups per CU: more than stay resident, so
                                    // CUs that finish early take queued ones (100M rows of
                                    // three columns, non-temporal loads: 0.264 / 0.259 /
                                    // 0.262 / 0.267 / 0.274 ms at 16 / 32 / 64 / 8 / 4 per CU)
      e = launch_pred_direct(bd, c.vm_direct_terms, c.vm_direct_wmask,
                             (int)std::max<int64_t>(1, std::min<int64_t>(tiles, (int64_t)device_cus() * 32)), st);
    } else {
      e = launch_vm(c.R, 1, c.vm_slow, grid_for(1, c.lds_bytes, tiles), c.lds_bytes, st, bd);
    }
    if (e == hipSuccess) e = mark(1);
    if (e == hipSuccess && (stages & 2)) e = launch_slot_scan(bd, n_chunks, st);
    if (e == hipSuccess) e = mark(2);
    if (e == hipSuccess && (stages & 4))
      e = launch_vm(c.R, 2, c.vm_slow, grid_for(2, c.lds_gather, (chunks + 3) / 4), c.lds_gather, st, bd);
    return e == hipSuccess ? mark(3) : e;
  }
  e = launch_interp(c.S, c.R, compact, grid, c.lds_bytes, st, blob, ws);
  return e == hipSuccess ? mark(1) : e;
}

struct RunResult {
  int64_t n_sel = 0;
  std::vector<int32_t> nulls;
};

// One-shot execution of a compiled program (stream-ordered allocations, blocking).
RunResult run_once(const Compiled &c, bool compact, int64_t n, int n_out, hipStream_t st) {
  RunResult r;
  r.nulls.assign(n_out, 0);
  if (n == 0) return r;
  const int64_t tiles = (n + tile_rows(c) - 1) / tile_rows(c);
  LaunchBuffers lb;
  layout(lb, c, n_out, compact, n);
  LbLease lease(lb.bytes, st);
  lb.base = lease.get();
  bind(lb, n_out);
  std::vector<unsigned long long> host_v(n_out + 1, 0);
  unsigned long long *host = lease.host(n_out + 1, host_v.data());
  const std::vector<uint8_t> img = device_image(lb, c);
  hipError_t e = hipMemcpyAsync(lb.blob, img.data(), img.size(), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = launch_fill_zero(lb.base + lb.zero_off, (int64_t)lb.zero_bytes,
                                            (int)std::min<size_t>(1024, lb.zero_bytes / 2048 + 1), st);
  if (e == hipSuccess)
    e = launch_program(c, compact, interp_grid(c, compact, n), st, lb.blob, lb.ws, lb.bd, lb.n_chunks);
  if (e == hipSuccess && n_out > 0)
    e = hipMemcpyAsync(host, lb.ws.null_counts, sizeof(unsigned long long) * n_out, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && compact)
    e = hipMemcpyAsync(host + n_out, lb.ws.n_sel, sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
  hipError_t s = hipStreamSynchronize(st);
  HIPCHECK(e);
  HIPCHECK(s);
  for (int o = 0; o < n_out; o++) r.nulls[o] = (int32_t)host[o];
  r.n_sel = compact ? (int64_t)host[n_out] : n;
  return r;
}

// One-shot tile-VM filter whose outputs are allocated after its slot scan (blocking):
// `c` was compiled from `req` with stand-in output pointers; allocate(rows) fills
// req.outputs with columns of `rows` rows (at least one, so every bitmap exists).
template <typename Alloc>
RunResult run_filter_sized(const Compiled &c, CompileRequest &req, Alloc allocate, int64_t n, int n_out,
                           hipStream_t st) {
  RunResult r;
  r.nulls.assign(n_out, 0);
  LaunchBuffers lb;
  layout(lb, c, n_out, true, n);
  LbLease lease(lb.bytes, st);
  lb.base = lease.get();
  bind(lb, n_out);
  std::vector<unsigned long long> host_v(n_out + 1, 0);
  unsigned long long *host = lease.host(n_out + 1, host_v.data());
  hipError_t e = hipSuccess;
  try {
    const std::vector<uint8_t> img = device_image(lb, c);
    e = hipMemcpyAsync(lb.blob, img.data(), img.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
      e = launch_fill_zero(lb.base + lb.zero_off, (int64_t)lb.zero_bytes,
                           (int)std::min<size_t>(1024, lb.zero_bytes / 2048 + 1), st);
    if (e == hipSuccess)
      e = launch_program(c, true, interp_grid(c, true, n), st, lb.blob, lb.ws, lb.bd, lb.n_chunks, nullptr, -1, 3);
    if (e == hipSuccess)
      e = hipMemcpyAsync(host + n_out, lb.ws.n_sel, sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    HIPCHECK(e);
    r.n_sel = (int64_t)host[n_out];
    allocate(std::max<int64_t>(r.n_sel, 1));
    const Compiled c2 = compile_program(req);
    if (c2.blob.size() != c.blob.size() || !c2.vm || c2.R != c.R || c2.T != c.T)
      throw std::runtime_error("internal: the filter's program changed with its outputs");
    const std::vector<uint8_t> img2 = device_image(lb, c2);
    e = hipMemcpyAsync(lb.blob, img2.data(), img2.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
      e = launch_program(c2, true, interp_grid(c2, true, n), st, lb.blob, lb.ws, lb.bd, lb.n_chunks, nullptr, -1, 4);
    if (e == hipSuccess && n_out > 0)
      e = hipMemcpyAsync(host, lb.ws.null_counts, sizeof(unsigned long long) * n_out,
                         hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    HIPCHECK(e);
  } catch (...) {
    (void)hipStreamSynchronize(st);  // before the lease lets the buffer go
    throw;
  }
  for (int o = 0; o < n_out; o++) r.nulls[o] = (int32_t)host[o];
  return r;
}

// One execution of a projection program (no compaction) whose launch buffers are
// the device's allocation-free scratch: no device allocation, blocking. Calls on
// one device are serialised by the scratch's lock, so two never share it in flight.
RunResult run_in_scratch(const Compiled &c, int64_t n, int n_out, hipStream_t st) {
  RunResult r;
  r.nulls.assign(n_out, 0);
  uint8_t *base = nullptr;
  size_t cap = 0;
  std::mutex *mu = nullptr;
  HIPCHECK(noalloc_scratch(&base, &cap, &mu));
  LaunchBuffers lb;
  layout(lb, c, n_out, false, n);
  if (lb.bytes > cap)
    invalid("the plan's launch buffers need " + std::to_string(lb.bytes) + " bytes; the allocation-free scratch holds " +
            std::to_string(cap));
  if (n == 0) return r;
  std::lock_guard<std::mutex> g(*mu);
  lb.base = base;
  bind(lb, n_out);
  std::vector<unsigned long long> host(n_out + 1, 0);
  const std::vector<uint8_t> img = device_image(lb, c);
  hipError_t e = hipMemcpyAsync(lb.blob, img.data(), img.size(), hipMemcpyHostToDevice, st);
  // the null counts and counters start at zero on every call (the scratch keeps
  // the last call's contents)
  if (e == hipSuccess) e = launch_fill_zero(lb.base + lb.zero_off, (int64_t)lb.zero_bytes, 1, st);
  if (e == hipSuccess) e = launch_program(c, false, interp_grid(c, false, n), st, lb.blob, lb.ws, lb.bd, lb.n_chunks);
  if (e == hipSuccess && n_out > 0)
    e = hipMemcpyAsync(host.data(), lb.ws.null_counts, sizeof(unsigned long long) * n_out, hipMemcpyDeviceToHost, st);
  hipError_t sy = hipStreamSynchronize(st);
  HIPCHECK(e);
  HIPCHECK(sy);
  for (int o = 0; o < n_out; o++) r.nulls[o] = (int32_t)host[o];
  r.n_sel = n;
  return r;
}

std::vector<int> frame_dtypes(gdf_column *const *frame, int n) {
  std::vector<int> d(n);
  for (int i = 0; i < n; i++) d[i] = frame[i]->dtype;
  return d;
}

DevInputInfo dev_info(const gdf_column *c) {
  return DevInputInfo{c->data, c->valid, c->dtype, c->null_count};
}

// Maps frame columns to device inputs (in first-use order).
struct InputMap {
  std::vector<int> frame_to_dev;
  std::vector<DevInputInfo> inputs;
  explicit InputMap(int n_frame) : frame_to_dev(n_frame, -1) {}
  int use(gdf_column *const *frame, int f) {
    if (frame_to_dev[f] < 0) {
      frame_to_dev[f] = (int)inputs.size();
      inputs.push_back(dev_info(frame[f]));
    }
    return frame_to_dev[f];
  }
  void remap(gdf_column *const *frame, PlanSpec &p) {
    for (auto &c : p.in_dev) c = use(frame, c);
  }
};

void check_frame(gdf_column *const *frame, int n_frame) {
  if (n_frame <= 0 || !frame) invalid("empty frame");
  const int64_t n = frame[0] ? frame[0]->size : 0;
  for (int i = 0; i < n_frame; i++) check_column(frame[i], "frame column", n);
}

// Device memory from the serving loop's arena, or from the driver.
void *arena_or_malloc(DeviceArena *arena, size_t bytes) {
  if (bytes == 0) return nullptr;
  if (arena) {
    void *p = arena->alloc(bytes);
    if (!p) throw BlzError(BLZ_ERR_OOM, "serving arena: out of device memory");
    return p;
  }
  void *p = nullptr;
  HIPCHECK(hipMalloc(&p, bytes));
  return p;
}
void arena_or_free(DeviceArena *arena, void *p) {
  if (!p) return;
  if (arena) arena->free(p);
  else (void)hipFree(p);
}

// Device allocation; stream-ordered (no implicit device synchronisation) when
// `async`, for callers that must not wait for other streams' work.
void *dalloc(size_t bytes, hipStream_t st = nullptr, bool async = false) {
  if (bytes == 0) return nullptr;
  void *p = nullptr;
  if (async) p = pool_alloc(bytes, st);
  else HIPCHECK(hipMalloc(&p, bytes));
  return p;
}

// Allocates a result column of capacity n (validity zeroed: compaction ORs words).
gdf_column alloc_column(int dtype, int64_t n, bool with_valid, hipStream_t st, bool async = false) {
  gdf_column c{};
  c.dtype = (gdf_dtype)dtype;
  const int w = blz_dtype_width((gdf_dtype)dtype);
  if (w == 0) throw BlzError(BLZ_ERR_UNSUPPORTED, "dtype without a fixed width is not supported");
  c.data = dalloc(round_up((size_t)n * w, 64), st, async);
  if (with_valid && n > 0) {
    const size_t vb = gdf_valid_allocation_size((gdf_size_type)n);
    c.valid = (gdf_valid_type *)dalloc(vb, st, async);
    HIPCHECK(launch_fill_zero(c.valid, (int64_t)vb, (int)std::min<size_t>(1024, vb / 2048 + 1), st));
  }
  c.size = (gdf_size_type)n;
  return c;
}

std::string condition_of(const char *qp) {
  std::string s = qp ? qp : "";
  if (s.find("condition=[") != std::string::npos) return planner::get_named_expression(s, "condition");
  return s;
}

}  // namespace

extern "C" {

const char *blz_version(void) { return "blz-mi355x 0.1 (gfx950)"; }

int blz_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

}  // extern "C"

namespace {
// perform_operation's plan compiled for the given columns (throws on error).
Compiled compile_register_plan(gdf_column *const *output_columns, int n_outputs, gdf_column *const *input_columns,
                               int n_inputs, const column_index_type *left_inputs,
                               const column_index_type *right_inputs, const column_index_type *outputs,
                               int n_operations, const column_index_type *final_output_positions,
                               const gdf_binary_operator *operators, const gdf_unary_operator *unary_operators,
                               const gdf_scalar *left_scalars, const gdf_scalar *right_scalars, int scalar_mode) {
  if (n_inputs <= 0) invalid("perform_operation needs at least one input column");
  if (n_inputs > kMaxSlots) invalid("more than 64 input columns");
  if (n_outputs < 0 || n_operations < 0) invalid("negative count");
  const int64_t n = input_columns[0]->size;
  for (int i = 0; i < n_inputs; i++) check_column(input_columns[i], "input column", n);
  for (int i = 0; i < n_outputs; i++) check_column(output_columns[i], "output column", n);
  PlanSpec p;
  for (int c = 0; c < n_inputs; c++) p.in_dev.push_back(c);
  p.left.assign(left_inputs, left_inputs + n_operations);
  p.right.assign(right_inputs, right_inputs + n_operations);
  p.outs.assign(outputs, outputs + n_operations);
  p.final_pos.assign(final_output_positions, final_output_positions + n_outputs);
  for (int k = 0; k < n_operations; k++) {
    p.binops.push_back(operators[k]);
    p.unops.push_back(unary_operators[k]);
    p.lsc.push_back(left_scalars[k]);
    p.rsc.push_back(right_scalars[k]);
  }
  CompileRequest req;
  req.n_rows = n;
  for (int c = 0; c < n_inputs; c++) req.inputs.push_back(dev_info(input_columns[c]));
  req.proj = &p;
  req.scalar_mode = scalar_mode;
  for (int o = 0; o < n_outputs; o++)
    req.outputs.push_back({output_columns[o]->data, output_columns[o]->valid, output_columns[o]->dtype, 0, o});
  return compile_program(req);
}
}  // namespace

extern "C" {

int blz_perform_operation_noalloc(gdf_column *const *output_columns, int n_outputs,
                                  gdf_column *const *input_columns, int n_inputs,
                                  const column_index_type *left_inputs, const column_index_type *right_inputs,
                                  const column_index_type *outputs, int n_operations,
                                  const column_index_type *final_output_positions,
                                  const gdf_binary_operator *operators, const gdf_unary_operator *unary_operators,
                                  const gdf_scalar *left_scalars, const gdf_scalar *right_scalars, int scalar_mode,
                                  int32_t *out_null_counts, blz_stream_t stream, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    Compiled c = compile_register_plan(output_columns, n_outputs, input_columns, n_inputs, left_inputs, right_inputs,
                                       outputs, n_operations, final_output_positions, operators, unary_operators,
                                       left_scalars, right_scalars, scalar_mode);
    RunResult r = run_in_scratch(c, input_columns[0]->size, n_outputs, S(stream));
    if (out_null_counts)
      for (int o = 0; o < n_outputs; o++) out_null_counts[o] = output_columns[o]->valid ? r.nulls[o] : 0;
  });
}

int blz_perform_operation(gdf_column *const *output_columns, int n_outputs,
                          gdf_column *const *input_columns, int n_inputs,
                          const column_index_type *left_inputs, const column_index_type *right_inputs,
                          const column_index_type *outputs, int n_operations,
                          const column_index_type *final_output_positions,
                          const gdf_binary_operator *operators, const gdf_unary_operator *unary_operators,
                          const gdf_scalar *left_scalars, const gdf_scalar *right_scalars, int scalar_mode,
                          int32_t *out_null_counts, blz_stream_t stream, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    Compiled c = compile_register_plan(output_columns, n_outputs, input_columns, n_inputs, left_inputs, right_inputs,
                                       outputs, n_operations, final_output_positions, operators, unary_operators,
                                       left_scalars, right_scalars, scalar_mode);
    RunResult r = run_once(c, false, input_columns[0]->size, n_outputs, S(stream));
    if (out_null_counts)
      for (int o = 0; o < n_outputs; o++) out_null_counts[o] = output_columns[o]->valid ? r.nulls[o] : 0;
  });
}

int blz_evaluate_expression(gdf_column *const *frame, int n_frame, const char *expression,
                            gdf_column *output, int scalar_mode, int32_t *out_null_count,
                            blz_stream_t stream, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    check_frame(frame, n_frame);
    if (!expression) invalid("expression is NULL");
    const int64_t n = frame[0]->size;
    check_column(output, "output column", n);
    planner::ExprPlan ep = planner::plan_expression(frame_dtypes(frame, n_frame), expression);
    if (ep.bare) {
      const gdf_column *src = frame[ep.bare_col];
      if (src->dtype != output->dtype) invalid("bare-column expression: output dtype differs from the column");
      const size_t bytes = (size_t)n * blz_dtype_width(src->dtype);
      if (bytes) HIPCHECK(hipMemcpyAsync(output->data, src->data, bytes, hipMemcpyDeviceToDevice, S(stream)));
      if (output->valid && src->valid && n > 0)
        HIPCHECK(hipMemcpyAsync(output->valid, src->valid, ((size_t)n + 7) / 8, hipMemcpyDeviceToDevice, S(stream)));
      else if (output->valid && n > 0)  // a column without a bitmap is all valid
        HIPCHECK(hipMemsetAsync(output->valid, 0xFF, ((size_t)n + 7) / 8, S(stream)));
      HIPCHECK(hipStreamSynchronize(S(stream)));
      if (out_null_count) *out_null_count = src->valid ? src->null_count : 0;
      return;
    }
    InputMap im(n_frame);
    im.remap(frame, ep.spec);
    CompileRequest req;
    req.n_rows = n;
    req.inputs = im.inputs;
    req.proj = &ep.spec;
    req.scalar_mode = scalar_mode;
    req.outputs.push_back({output->data, output->valid, output->dtype, 0, 0});
    Compiled c = compile_program(req);
    RunResult r = run_once(c, false, n, 1, S(stream));
    if (out_null_count) *out_null_count = output->valid ? r.nulls[0] : 0;
  });
}

namespace {

// Fused filter (+ optional projection). `proj_qp` NULL keeps every frame column.
void filter_project_impl(gdf_column *const *frame, int n_frame, const char *filter_qp, const char *proj_qp,
                         gdf_column *out_columns, int max_out, int *n_out, int32_t *is_alias,
                         int64_t *n_kept, int scalar_mode, hipStream_t st) {
  check_frame(frame, n_frame);
  const int64_t n = frame[0]->size;
  const auto dt = frame_dtypes(frame, n_frame);
  InputMap im(n_frame);
  planner::ExprPlan fp;
  const bool has_filter = filter_qp != nullptr;
  if (has_filter) {
    fp = planner::plan_expression(dt, condition_of(filter_qp));
    if (fp.bare) im.use(frame, fp.bare_col);
    else im.remap(frame, fp.spec);
  }
  planner::ProjectPlan pp;
  std::vector<planner::ProjOut> outs;
  if (proj_qp) {
    pp = planner::parse_project(dt, proj_qp);
    if (pp.has_eval) im.remap(frame, pp.spec);
    outs = pp.outputs;
  } else {
    for (int i = 0; i < n_frame; i++) outs.push_back({1, i, dt[i], ""});
  }
  if ((int)outs.size() > max_out) invalid("more outputs than out_columns can hold");
  // Allocate outputs: evaluated outputs always carry validity (gdf_column_cpp
  // allocates it, GDFColumn.cu:302-338); gathered ones iff their input has it.
  std::vector<gdf_column> res(outs.size());
  CompileRequest req;
  req.n_rows = n;
  req.scalar_mode = scalar_mode;
  req.compact = has_filter;
  int64_t kept = n;
  uint8_t *placeholder = nullptr;
  try {
    // Output columns of `rows` rows, their pointers into req.outputs.
    auto allocate_outputs = [&](int64_t rows) {
      req.outputs.clear();
      for (size_t i = 0; i < outs.size(); i++) {
        const auto &o = outs[i];
        const bool with_valid = o.kind == 0 ? true : frame[o.index]->valid != nullptr;
        res[i] = alloc_column(o.dtype, rows, with_valid, st);
        if (o.kind == 0) req.outputs.push_back({res[i].data, res[i].valid, o.dtype, 0, o.index});
        else req.outputs.push_back({res[i].data, res[i].valid, o.dtype, 1, im.use(frame, o.index)});
      }
    };
    // the columns bare outputs copy are inputs too (use() keeps one index per column)
    for (const auto &o : outs)
      if (o.kind != 0) (void)im.use(frame, o.index);
    req.inputs = im.inputs;
    if (has_filter) {
      if (fp.bare) req.bare_stencil_input = im.frame_to_dev[fp.bare_col];
      else req.pred = &fp.spec;
    }
    if (proj_qp && pp.has_eval) req.proj = &pp.spec;
    std::vector<int32_t> nulls(outs.size(), 0);
    if (n > 0 && !req.inputs.empty()) {
      // A filter the tile VM runs in three launches gets outputs of the rows it keeps,
      // as the reference materializes the kept rows (CalciteInterpreter.cpp:543-560):
      // the program is compiled with stand-in output pointers (the predicate and the
      // slot scan never touch the outputs), the kept count is read after the scan,
      // the outputs are allocated to it, and the program compiled again with them
      // (the same plan: the same layout) runs the gather.
      Compiled c;
      bool sized = false;
      if (has_filter) {
        placeholder = stand_in_address();
        for (size_t i = 0; i < outs.size(); i++) {
          const auto &o = outs[i];
          const bool with_valid = o.kind == 0 ? true : frame[o.index]->valid != nullptr;
          uint8_t *v = with_valid ? placeholder : nullptr;
          if (o.kind == 0) req.outputs.push_back({placeholder, v, o.dtype, 0, o.index});
          else req.outputs.push_back({placeholder, v, o.dtype, 1, im.use(frame, o.index)});
        }
        c = compile_program(req);
        sized = c.vm;
      }
      if (sized) {
        RunResult r = run_filter_sized(c, req, allocate_outputs, n, (int)outs.size(), st);
        kept = r.n_sel;
        nulls = r.nulls;
      } else {
        allocate_outputs(n);
        c = compile_program(req);
        RunResult r = run_once(c, has_filter, n, (int)outs.size(), st);
        kept = r.n_sel;
        nulls = r.nulls;
      }
    } else {
      allocate_outputs(n);
      HIPCHECK(hipStreamSynchronize(st));
      if (n > 0 && req.inputs.empty()) invalid("nothing to evaluate");
    }
    for (size_t i = 0; i < outs.size(); i++) {
      res[i].size = (gdf_size_type)kept;
      res[i].null_count = (res[i].valid && kept > 0) ? nulls[i] : 0;
      out_columns[i] = res[i];
      if (is_alias) is_alias[i] = 0;
    }
  } catch (...) {
    (void)hipStreamSynchronize(st);  // the stream may still use what is freed here
    for (auto &c : res) { (void)hipFree(c.data); (void)hipFree(c.valid); }
    throw;
  }
  if (n_out) *n_out = (int)outs.size();
  if (n_kept) *n_kept = kept;
}

}  // namespace

int blz_process_filter(gdf_column *const *frame, int n_frame, const char *query_part,
                       gdf_column *out_columns, int64_t *n_kept, int scalar_mode, blz_stream_t stream,
                       char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!query_part) invalid("query_part is NULL");
    int n_out = 0;
    filter_project_impl(frame, n_frame, query_part, nullptr, out_columns, n_frame, &n_out, nullptr, n_kept,
                        scalar_mode, S(stream));
  });
}

namespace {
// One aggregation of a LogicalAggregate line (the reference's gdf_agg_op, by name).
enum AggOp { AGG_SUM, AGG_MIN, AGG_MAX, AGG_AVG, AGG_COUNT };
struct AggSpec {
  AggOp op;
  std::string input;  // "$k", an expression, or empty
};

std::string trim(const std::string &s) {
  const size_t a = s.find_first_not_of(" \t\n"), b = s.find_last_not_of(" \t\n");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// process_aggregate's parsing (GroupBy.cpp:543-569, get_aggregation_operation
// CalciteExpressionParsing.cpp:483-507, get_string_between_outer_parentheses :756-767).
std::vector<AggSpec> parse_aggregate(const std::string &qp) {
  const size_t a = qp.find('('), b = qp.rfind(')');
  if (a == std::string::npos || b == std::string::npos || b < a) invalid("not a LogicalAggregate line");
  const std::string combined = qp.substr(a + 1, b - a - 1);
  const std::string group = trim(planner::get_named_expression(combined, "group"));
  if (group.size() > 2) throw BlzError(BLZ_ERR_UNSUPPORTED, "aggregation with GROUP BY is not supported");
  std::vector<AggSpec> out;
  for (const std::string &e0 : planner::get_expressions_from_expression_list(combined)) {
    const std::string e = trim(e0);
    if (e.find("group=") != std::string::npos) continue;
    const size_t eq = e.find("=[");
    if (eq == std::string::npos) invalid("aggregation without =[...]: " + e);
    const size_t lp = e.find('(', eq), rp = e.rfind(')');
    if (lp == std::string::npos || rp == std::string::npos || rp < lp) invalid("aggregation without (...): " + e);
    const std::string name = trim(e.substr(eq + 2, lp - eq - 2));
    AggSpec s;
    if (name == "SUM") s.op = AGG_SUM;
    else if (name == "MIN") s.op = AGG_MIN;
    else if (name == "MAX") s.op = AGG_MAX;
    else if (name == "AVG") s.op = AGG_AVG;
    else if (name == "COUNT") s.op = AGG_COUNT;
    else if (name == "COUNT_DISTINCT") throw BlzError(BLZ_ERR_UNSUPPORTED, "COUNT_DISTINCT is not supported");
    else invalid("aggregation type not supported, " + name);
    s.input = trim(e.substr(lp + 1, rp - lp - 1));
    out.push_back(s);
  }
  return out;
}

bool is_float_dtype(int t) { return t == GDF_FLOAT32 || t == GDF_FLOAT64; }
double b2d_host(uint64_t u) {
  double d;
  std::memcpy(&d, &u, 8);
  return d;
}

// get_aggregation_output_type (CalciteExpressionParsing.cpp:92-117), no GROUP BY.
int aggregation_output_type(int input_type, AggOp op) {
  switch (op) {
    case AGG_COUNT: return GDF_INT64;
    case AGG_SUM: return is_float_dtype(input_type) ? GDF_FLOAT64 : GDF_INT64;
    case AGG_AVG: return GDF_FLOAT64;
    default: return input_type;
  }
}

// Reduction of the valid rows of `c` (op 0 SUM, 1 MIN, 2 MAX) on the GPU. A column
// whose null_count is 0 is read as all valid, its bitmap unread (the interpreter's
// rule for its inputs, interpreter_ops.cuh:1312-1327).
RedPart reduce_column(const gdf_column *c, int op, hipStream_t st) {
  const int load = column_load_kind(c->dtype);
  if (load == LD_NONE) throw BlzError(BLZ_ERR_UNSUPPORTED, "aggregation over this dtype is not supported");
  const int64_t n = c->size;
  const int grid = reduce_grid(load, n);
  // the partials live in the stream's launch buffers, the result is read back
  // through its pinned words
  LbLease lease(sizeof(RedPart) * (grid + 1), st);
  RedPart *parts = (RedPart *)lease.get();
  RedPart r_v{};
  RedPart *r = (RedPart *)lease.host(sizeof(RedPart) / sizeof(unsigned long long), (unsigned long long *)&r_v);
  const uint8_t *valid = c->null_count > 0 ? (const uint8_t *)c->valid : nullptr;
  hipError_t e = launch_reduce(op, c->data, load, valid, n, parts, grid, st);
  if (e == hipSuccess) e = hipMemcpyAsync(r, parts + grid, sizeof(RedPart), hipMemcpyDeviceToHost, st);
  hipError_t s = hipStreamSynchronize(st);
  HIPCHECK(e);
  HIPCHECK(s);
  return *r;
}

// A one-row column holding `bits` (the first width bytes) and its validity.
gdf_column scalar_column(int dtype, uint64_t bits, bool valid, hipStream_t st) {
  gdf_column c{};
  c.dtype = (gdf_dtype)dtype;
  c.size = 1;
  const size_t vb = gdf_valid_allocation_size(1);
  c.data = dalloc(64, st);
  c.valid = (gdf_valid_type *)dalloc(vb, st);
  // the whole padded bitmap is written (no separate clearing launch)
  std::vector<uint8_t> vbytes(vb, 0);
  vbytes[0] = valid ? 1 : 0;
  hipError_t e = hipMemcpyAsync(c.data, &bits, blz_dtype_width((gdf_dtype)dtype), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(c.valid, vbytes.data(), vb, hipMemcpyHostToDevice, st);
  hipError_t s = hipStreamSynchronize(st);
  if (e != hipSuccess || s != hipSuccess) {
    (void)hipFree(c.data);
    (void)hipFree(c.valid);
    HIPCHECK(e);
    HIPCHECK(s);
  }
  c.null_count = valid ? 0 : 1;
  return c;
}
}  // namespace

int blz_process_aggregate(gdf_column *const *frame, int n_frame, const char *query_part,
                          gdf_column *out_columns, int max_out, int *n_out, int scalar_mode,
                          blz_stream_t stream, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    check_frame(frame, n_frame);
    if (!query_part || !out_columns || !n_out) invalid("bad arguments");
    const std::vector<AggSpec> aggs = parse_aggregate(query_part);
    if ((int)aggs.size() > max_out) invalid("more aggregations than out_columns can hold");
    hipStream_t st = S(stream);
    const int64_t n = frame[0]->size;
    const auto dt = frame_dtypes(frame, n_frame);
    std::vector<gdf_column> made;
    try {
      for (const AggSpec &a : aggs) {
        // the aggregation's input column (compute_aggregations, GroupBy.cpp:303-315)
        gdf_column temp{};
        const gdf_column *in = nullptr;
        if (planner::contains_evaluation(a.input)) {
          const int t = planner::get_output_type_expression(dt, a.input);
          temp = alloc_column(t, n, true, st);
          int32_t nulls = 0;
          char e2[512] = {0};
          const int rc = blz_evaluate_expression(frame, n_frame, a.input.c_str(), &temp, scalar_mode, &nulls,
                                                 stream, e2, sizeof(e2));
          if (rc != BLZ_OK) {
            blz_free_column(&temp);
            throw BlzError(rc, std::string("aggregation input: ") + e2);
          }
          // evaluate_expression leaves the column's null_count at 0 (the call that
          // would update it is commented out, LogicalFilter.cpp:368), so COUNT and
          // AVG's divisor count every row (GroupBy.cpp:263, :284) and the reduction
          // reads every row, as for any column whose null_count is 0
          (void)nulls;
          temp.null_count = 0;
          in = &temp;
        } else {
          int k = 0;
          if (!a.input.empty()) {
            if (a.input[0] != '$') invalid("aggregation input is neither a column nor an expression: " + a.input);
            k = std::stoi(a.input.substr(1));
          }
          if (k < 0 || k >= n_frame) invalid("aggregation input column out of range: " + a.input);
          in = frame[k];
        }
        const int ot = aggregation_output_type(in->dtype, a.op);
        const int64_t rows = in->size, valid_rows = rows - (in->valid ? in->null_count : 0);
        uint64_t bits = 0;
        bool valid = true;
        try {
          if (a.op == AGG_COUNT) {
            bits = (uint64_t)valid_rows;
          } else if (a.op == AGG_AVG) {
            // the sum over the valid rows in INT64 or FLOAT64, divided in binary64
            // by size - null_count (GroupBy.cpp:251-270)
            valid = rows > 0 && valid_rows > 0;
            if (valid) {
              const RedPart r = reduce_column(in, 0, st);
              const double sum = is_float_dtype(in->dtype) ? b2d_host(r.v) : (double)(int64_t)r.v;
              const double avg = sum / (double)valid_rows;
              std::memcpy(&bits, &avg, 8);
            }
          } else {
            const RedPart r = rows > 0 ? reduce_column(in, a.op == AGG_SUM ? 0 : a.op == AGG_MIN ? 1 : 2, st)
                                       : RedPart{0, 0};
            valid = r.cnt > 0;
            bits = r.v;
            if (valid && a.op != AGG_SUM && ot == GDF_FLOAT32) {  // widened float32 back
              const float f = (float)b2d_host(r.v);
              uint32_t u;
              std::memcpy(&u, &f, 4);
              bits = u;
            }
            if (!valid) bits = 0;
          }
        } catch (...) {
          if (temp.data) blz_free_column(&temp);
          throw;
        }
        if (temp.data) blz_free_column(&temp);
        made.push_back(scalar_column(ot, bits, valid, st));
      }
    } catch (...) {
      for (auto &c : made) blz_free_column(&c);
      throw;
    }
    for (size_t i = 0; i < made.size(); i++) out_columns[i] = made[i];
    *n_out = (int)made.size();
  });
}

int blz_filter_project(gdf_column *const *frame, int n_frame, const char *filter_query_part,
                       const char *project_query_part, gdf_column *out_columns, int max_out, int *n_out,
                       int64_t *n_kept, int scalar_mode, blz_stream_t stream, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!filter_query_part) invalid("filter_query_part is NULL");
    filter_project_impl(frame, n_frame, filter_query_part, project_query_part, out_columns, max_out, n_out,
                        nullptr, n_kept, scalar_mode, S(stream));
  });
}

int blz_execute_project(gdf_column *const *frame, int n_frame, const char *query_part,
                        gdf_column *out_columns, int32_t *is_alias, int max_out, int *n_out,
                        int scalar_mode, blz_stream_t stream, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    check_frame(frame, n_frame);
    if (!query_part) invalid("query_part is NULL");
    const int64_t n = frame[0]->size;
    planner::ProjectPlan pp = planner::parse_project(frame_dtypes(frame, n_frame), query_part);
    if ((int)pp.outputs.size() > max_out) invalid("more outputs than out_columns can hold");
    std::vector<gdf_column> res(pp.outputs.size());
    std::vector<int> eval_pos;
    CompileRequest req;
    req.n_rows = n;
    req.scalar_mode = scalar_mode;
    InputMap im(n_frame);
    try {
      for (size_t i = 0; i < pp.outputs.size(); i++) {
        const auto &o = pp.outputs[i];
        if (o.kind == 0) {
          res[i] = alloc_column(o.dtype, n, true, S(stream));
          req.outputs.push_back({res[i].data, res[i].valid, o.dtype, 0, o.index});
          eval_pos.push_back((int)i);
        } else {
          res[i] = *frame[o.index];
        }
      }
      std::vector<int32_t> nulls(req.outputs.size(), 0);
      if (pp.has_eval && n > 0) {
        im.remap(frame, pp.spec);
        req.inputs = im.inputs;
        req.proj = &pp.spec;
        Compiled c = compile_program(req);
        RunResult r = run_once(c, false, n, (int)req.outputs.size(), S(stream));
        nulls = r.nulls;
      }
      // update_null_count on every output (CalciteInterpreter.cpp:332-335).
      for (size_t k = 0; k < eval_pos.size(); k++) res[eval_pos[k]].null_count = n > 0 ? nulls[k] : 0;
      for (size_t i = 0; i < pp.outputs.size(); i++) {
        if (pp.outputs[i].kind == 1) {
          gdf_column &c = res[i];
          if (c.valid && c.size > 0) {
            unsigned long long *d = nullptr, h = 0;
            d = (unsigned long long *)pool_alloc(8, S(stream));
            HIPCHECK(launch_fill_zero(d, 8, 1, S(stream)));
            HIPCHECK(launch_null_count(c.valid, c.size, d, 256, S(stream)));
            HIPCHECK(hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, S(stream)));
            HIPCHECK(hipFreeAsync(d, S(stream)));
            HIPCHECK(hipStreamSynchronize(S(stream)));
            c.null_count = (gdf_size_type)h;
          } else {
            c.null_count = 0;
          }
        }
        out_columns[i] = res[i];
        if (is_alias) is_alias[i] = pp.outputs[i].kind == 1;
      }
    } catch (...) {
      for (int p : eval_pos) { (void)hipFree(res[p].data); (void)hipFree(res[p].valid); }
      throw;
    }
    if (n_out) *n_out = (int)pp.outputs.size();
  });
}

int blz_materialize_column(const gdf_column *input, gdf_column *output, const gdf_column *row_indices,
                           blz_stream_t stream, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    check_column(input, "input column", -1);
    check_column(row_indices, "row_indices", -1);
    check_column(output, "output column", row_indices->size);
    // the indices are read as a signed integer of their width (ColumnManipulation.cu:120-134)
    const int iw = blz_dtype_width(row_indices->dtype);
    if (iw == 0) throw BlzError(BLZ_ERR_UNSUPPORTED, "row_indices must have a 1, 2, 4 or 8-byte dtype");
    const int w = blz_dtype_width(input->dtype);
    if (w == 0 || blz_dtype_width(output->dtype) != w) invalid("input and output widths differ");
    const int64_t m = row_indices->size;
    unsigned long long *d = nullptr, h = 0;
    d = (unsigned long long *)pool_alloc(8, S(stream));
    hipError_t e = launch_fill_zero(d, 8, 1, S(stream));
    if (e == hipSuccess && m > 0) {
      const int64_t words = (m + 63) / 64;
      const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((words + 15) / 16, (int64_t)device_cus() * 16));
      e = launch_gather(w, input->data, input->valid, input->size, row_indices->data, iw, m, output->data,
                        output->valid, d, grid, S(stream));
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, S(stream));
    (void)hipFreeAsync(d, S(stream));
    hipError_t s = hipStreamSynchronize(S(stream));
    HIPCHECK(e);
    HIPCHECK(s);
    output->null_count = (m == 0 || output->valid == nullptr) ? 0 : (gdf_size_type)h;
  });
}

int blz_materialize_columns(const gdf_column *const *inputs, gdf_column *const *outputs, int n_columns,
                            const gdf_column *row_indices, blz_stream_t stream, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!inputs || !outputs || n_columns < 0) invalid("bad arguments");
    check_column(row_indices, "row_indices", -1);
    const int iw = blz_dtype_width(row_indices->dtype);
    if (iw == 0) throw BlzError(BLZ_ERR_UNSUPPORTED, "row_indices must have a 1, 2, 4 or 8-byte dtype");
    if (n_columns == 0) return;
    const int64_t m = row_indices->size;
    const int64_t n_in = inputs[0] ? inputs[0]->size : 0;
    for (int c = 0; c < n_columns; c++) {
      check_column(inputs[c], "input column", n_in);
      check_column(outputs[c], "output column", m);
      const int w = blz_dtype_width(inputs[c]->dtype);
      if (w == 0 || blz_dtype_width(outputs[c]->dtype) != w) invalid("input and output widths differ");
    }
    // One gather kernel per column, stream-ordered, one synchronisation: a single
    // pass over all columns measured slower on a random permutation (its random
    // accesses spread over every column at once).
    hipStream_t st = S(stream);
    unsigned long long *d = nullptr;
    std::vector<unsigned long long> h(n_columns, 0);
    d = (unsigned long long *)pool_alloc(8 * (size_t)n_columns, st);
    hipError_t e = launch_fill_zero(d, 8 * (int64_t)n_columns, 1, st);
    const int64_t words = (m + 63) / 64;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((words + 15) / 16, (int64_t)device_cus() * 16));
    for (int c = 0; e == hipSuccess && m > 0 && c < n_columns; c++)
      e = launch_gather(blz_dtype_width(inputs[c]->dtype), inputs[c]->data, inputs[c]->valid, n_in,
                        row_indices->data, iw, m, outputs[c]->data, outputs[c]->valid, d + c, grid, st);
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d, 8 * (size_t)n_columns, hipMemcpyDeviceToHost, st);
    (void)hipFreeAsync(d, st);
    hipError_t s = hipStreamSynchronize(st);
    HIPCHECK(e);
    HIPCHECK(s);
    for (int c = 0; c < n_columns; c++)
      outputs[c]->null_count = (m == 0 || outputs[c]->valid == nullptr) ? 0 : (gdf_size_type)h[c];
  });
}

int blz_gdf_sequence(int32_t *data, size_t size, int32_t init_val, int32_t step, blz_stream_t stream,
                     char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (size == 0) return;
    if (!data) invalid("data is NULL");
    const int grid = (int)std::min<size_t>((size + 255) / 256, (size_t)device_cus() * 8);
    HIPCHECK(launch_sequence(data, (int64_t)size, init_val, step, grid, S(stream)));
    HIPCHECK(hipStreamSynchronize(S(stream)));
  });
}

int blz_apply_boolean_mask(const gdf_column *input, const gdf_column *stencil, gdf_column *out,
                           blz_stream_t stream, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    check_column(input, "input column", -1);
    check_column(stencil, "stencil", input->size);
    const int64_t n = input->size;
    CompileRequest req;
    req.n_rows = n;
    req.compact = true;
    req.inputs = {dev_info(input), dev_info(stencil)};
    req.bare_stencil_input = 1;
    gdf_column res = alloc_column(input->dtype, n, input->valid != nullptr, S(stream));
    try {
      req.outputs.push_back({res.data, res.valid, input->dtype, 1, 0});
      int64_t kept = 0;
      int32_t nulls = 0;
      if (n > 0) {
        Compiled c = compile_program(req);
        RunResult r = run_once(c, true, n, 1, S(stream));
        kept = r.n_sel;
        nulls = r.nulls[0];
      }
      res.size = (gdf_size_type)kept;
      res.null_count = (res.valid && kept > 0) ? nulls : 0;
      *out = res;
    } catch (...) {
      (void)hipFree(res.data);
      (void)hipFree(res.valid);
      throw;
    }
  });
}

int blz_update_null_count(gdf_column *column, blz_stream_t stream, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    check_column(column, "column", -1);
    if (column->size == 0 || column->valid == nullptr) { column->null_count = 0; return; }
    unsigned long long *d = nullptr, h = 0;
    d = (unsigned long long *)pool_alloc(8, S(stream));
    hipError_t e = launch_fill_zero(d, 8, 1, S(stream));
    if (e == hipSuccess) e = launch_null_count(column->valid, column->size, d, 256, S(stream));
    if (e == hipSuccess) e = hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, S(stream));
    (void)hipFreeAsync(d, S(stream));
    hipError_t s = hipStreamSynchronize(S(stream));
    HIPCHECK(e);
    HIPCHECK(s);
    column->null_count = (gdf_size_type)h;
  });
}

// ---- prepared programs ----------------------------------------------------------

struct blz_program_s {
  Compiled c;
  bool compact = false;
  int64_t n = 0;
  int n_out = 0;
  int grid = 1;
  LaunchBuffers lb;
  std::vector<gdf_column> outs;
  hipEvent_t evs[4] = {nullptr, nullptr, nullptr, nullptr};  // created by the first timing call
  uint8_t *out_block = nullptr;  // every output's data and bitmap (one allocation)
  std::vector<uint8_t> img;  // host image of the blob + launch descriptor
  int64_t last_sel = -1;     // rows kept by the last launch read back (compiled filter: sizes its gather)
  DeviceArena *arena = nullptr;  // device memory from the serving loop's arena (else the driver)
  hipStream_t mem_st = nullptr;
};

namespace {
// kernel argument of make_program: the interpreters' automatic choice (tile VM, else
// the register-file kernel), never the compiled tier.
constexpr int kInterpAuto = -1;

// Compiles and allocates one program (throws on error); blz_program_create and
// the batch share it.
blz_program make_program(gdf_column *const *frame, int n_frame, const char *filter_qp, const char *proj_qp,
                         int scalar_mode, int kernel, int s, int r, int t, hipStream_t st = nullptr,
                         DeviceArena *arena = nullptr) {
  const bool sync = arena == nullptr;
  check_frame(frame, n_frame);
  const int64_t n = frame[0]->size;
  const auto dt = frame_dtypes(frame, n_frame);
  InputMap im(n_frame);
  planner::ExprPlan fp;
  const bool has_filter = filter_qp != nullptr;
  if (has_filter) {
    fp = planner::plan_expression(dt, condition_of(filter_qp));
    if (fp.bare) im.use(frame, fp.bare_col);
    else im.remap(frame, fp.spec);
  }
  planner::ProjectPlan pp;
  std::vector<planner::ProjOut> outs;
  if (proj_qp) {
    pp = planner::parse_project(dt, proj_qp);
    if (pp.has_eval) im.remap(frame, pp.spec);
    outs = pp.outputs;
  } else {
    for (int i = 0; i < n_frame; i++) outs.push_back({1, i, dt[i], ""});
  }
  auto *p = new blz_program_s();
  try {
    CompileRequest req;
    req.n_rows = n;
    req.scalar_mode = scalar_mode;
    req.compact = has_filter;
    req.force_s = s;
    req.force_r = r;
    req.force_t = t;
    req.kernel = kernel == kInterpAuto ? 0 : kernel;
    // Every output (capacity: every row) in one allocation, freed with the program.
    // No bitmap is cleared here: the kernels write every validity word of their
    // outputs, and a launch of the register-file kernel's compaction (which ORs
    // words at its tile edges) clears its outputs' bitmaps first.
    p->arena = arena;
    p->mem_st = st;
    {
      std::vector<size_t> off_data(outs.size()), off_valid(outs.size(), SIZE_MAX);
      size_t total = 0;
      for (size_t i = 0; i < outs.size(); i++) {
        const auto &o = outs[i];
        const int w = blz_dtype_width((gdf_dtype)o.dtype);
        if (w == 0) throw BlzError(BLZ_ERR_UNSUPPORTED, "dtype without a fixed width is not supported");
        off_data[i] = total;
        total += round_up((size_t)n * w, 256);
        const bool with_valid = o.kind == 0 ? true : frame[o.index]->valid != nullptr;
        if (with_valid && n > 0) {
          off_valid[i] = total;
          total += round_up(gdf_valid_allocation_size((gdf_size_type)n), 256);
        }
      }
      p->out_block = (uint8_t *)arena_or_malloc(arena, std::max<size_t>(total, 256));
      for (size_t i = 0; i < outs.size(); i++) {
        const auto &o = outs[i];
        gdf_column c{};
        c.dtype = (gdf_dtype)o.dtype;
        c.data = p->out_block + off_data[i];
        c.valid = off_valid[i] == SIZE_MAX ? nullptr : (gdf_valid_type *)(p->out_block + off_valid[i]);
        c.size = (gdf_size_type)n;
        p->outs.push_back(c);
        if (o.kind == 0) req.outputs.push_back({c.data, c.valid, o.dtype, 0, o.index});
        else req.outputs.push_back({c.data, c.valid, o.dtype, 1, im.use(frame, o.index)});
      }
    }
    req.inputs = im.inputs;
    if (has_filter) {
      if (fp.bare) req.bare_stencil_input = im.frame_to_dev[fp.bare_col];
      else req.pred = &fp.spec;
    }
    if (proj_qp && pp.has_eval) req.proj = &pp.spec;
    if (req.inputs.empty()) invalid("nothing to evaluate");
    // Automatic (kernel 0, no forced slots or tile rows): a prepared program is
    // launched many times, so it takes the compiled tier; a plan the generator or
    // the run-time compiler does not take falls back to the interpreters.
    if (kernel == 0 && s == 0 && t == 0) {
      req.kernel = 5;
      try {
        p->c = compile_program(req);
        (void)jit_function(p->c.jit_src, p->c.jit_name);
        if (has_filter) (void)jit_function(p->c.jit_src, "blz_cg_gather");
      } catch (const BlzError &) {
        throw;
      } catch (const std::exception &) {
        req.kernel = 0;
        p->c = compile_program(req);
      }
    } else {
      p->c = compile_program(req);
    }
    if (p->c.jit) {  // compiled and loaded now, not at launch
      (void)jit_function(p->c.jit_src, p->c.jit_name);
      if (has_filter) (void)jit_function(p->c.jit_src, "blz_cg_gather");
    }
    p->compact = has_filter;
    p->n = n;
    p->n_out = (int)outs.size();
    p->grid = interp_grid(p->c, has_filter, std::max<int64_t>(n, 1));
    layout(p->lb, p->c, p->n_out, has_filter, n);
    p->lb.base = (uint8_t *)arena_or_malloc(arena, p->lb.bytes);
    bind(p->lb, p->n_out);
    p->img = device_image(p->lb, p->c);
    if (sync) HIPCHECK(hipMemcpy(p->lb.blob, p->img.data(), p->img.size(), hipMemcpyHostToDevice));
    else HIPCHECK(hipMemcpyAsync(p->lb.blob, p->img.data(), p->img.size(), hipMemcpyHostToDevice, st));
    if (sync) HIPCHECK(hipDeviceSynchronize());
  } catch (...) {
    // copies this program enqueued may still target its memory: wait before the
    // arena hands it out again
    if (arena) (void)hipStreamSynchronize(st);
    blz_program_destroy(p);
    throw;
  }
  return p;
}
}  // namespace

int blz_program_create(blz_program *prog, gdf_column *const *frame, int n_frame, const char *filter_qp,
                       const char *proj_qp, gdf_column *out_columns, int n_out_cap, int *n_out,
                       int scalar_mode, int kernel, int s, int r, int t, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!prog) invalid("prog is NULL");
    *prog = nullptr;
    if (kernel < 0 || kernel > 5) invalid("kernel must be 0, 3, 4 or 5");
    blz_program p = make_program(frame, n_frame, filter_qp, proj_qp, scalar_mode, kernel, s, r, t);
    if ((int)p->outs.size() > n_out_cap) {
      blz_program_destroy(p);
      invalid("more outputs than out_columns can hold");
    }
    for (size_t i = 0; i < p->outs.size(); i++) out_columns[i] = p->outs[i];
    if (n_out) *n_out = (int)p->outs.size();
    *prog = p;
  });
}

namespace {
hipError_t program_enqueue(blz_program p, hipStream_t st) {
  if (p->n == 0) return hipSuccess;
  hipError_t e = launch_fill_zero(p->lb.base + p->lb.zero_off, (int64_t)p->lb.zero_bytes,
                                  (int)std::min<size_t>(1024, p->lb.zero_bytes / 2048 + 1), st);
  // the direct kernel's compaction ORs boundary validity words: clear the outputs'
  // validity first (the tile VM writes every word of its outputs whole; the
  // compiled filter clears each bit range it owns before setting it)
  for (size_t i = 0; e == hipSuccess && p->compact && !p->c.vm && !p->c.jit && i < p->outs.size(); i++)
    if (p->outs[i].valid) {
      const size_t vb = gdf_valid_allocation_size((gdf_size_type)p->n);
      e = launch_fill_zero(p->outs[i].valid, (int64_t)vb, (int)std::min<size_t>(1024, vb / 2048 + 1), st);
    }
  if (e == hipSuccess)
    e = launch_program(p->c, p->compact, p->grid, st, p->lb.blob, p->lb.ws, p->lb.bd, p->lb.n_chunks, nullptr,
                       p->last_sel);
  return e;
}
}  // namespace

int blz_program_launch(blz_program prog, blz_stream_t stream, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!prog) invalid("prog is NULL");
    HIPCHECK(program_enqueue(prog, S(stream)));
  });
}

int blz_program_result(blz_program prog, blz_stream_t stream, int64_t *n_kept, int32_t *null_counts,
                       char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!prog) invalid("prog is NULL");
    std::vector<unsigned long long> h(prog->n_out + 1, 0);
    if (prog->n > 0) {
      if (prog->n_out > 0)
        HIPCHECK(hipMemcpyAsync(h.data(), prog->lb.ws.null_counts, 8 * prog->n_out, hipMemcpyDeviceToHost,
                                S(stream)));
      if (prog->compact)
        HIPCHECK(hipMemcpyAsync(h.data() + prog->n_out, prog->lb.ws.n_sel, 8, hipMemcpyDeviceToHost, S(stream)));
    }
    HIPCHECK(hipStreamSynchronize(S(stream)));
    const int64_t kept = prog->compact ? (int64_t)h[prog->n_out] : prog->n;
    prog->last_sel = kept;
    if (n_kept) *n_kept = kept;
    if (null_counts)
      for (int o = 0; o < prog->n_out; o++)
        null_counts[o] = (prog->outs[o].valid && kept > 0) ? (int32_t)h[o] : 0;
  });
}

int blz_program_time_stages(blz_program prog, blz_stream_t stream, int iters, double *stage_ms, int *n_stages,
                            char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!prog || iters <= 0 || !stage_ms) invalid("bad arguments");
    for (auto &e : prog->evs)
      if (!e) HIPCHECK(hipEventCreate(&e));
    const int ns = program_stages(prog->c, prog->compact);
    std::vector<double> total(ns, 0.0);
    for (int i = 0; i < iters; i++) {
      // the zero-fill of counters is part of a launch but not of the kernel time
      HIPCHECK(launch_fill_zero(prog->lb.base + prog->lb.zero_off, (int64_t)prog->lb.zero_bytes,
                                (int)std::min<size_t>(1024, prog->lb.zero_bytes / 2048 + 1), S(stream)));
      for (size_t o = 0; prog->compact && !prog->c.vm && !prog->c.jit && o < prog->outs.size(); o++)
        if (prog->outs[o].valid) {
          const size_t vb = gdf_valid_allocation_size((gdf_size_type)prog->n);
          HIPCHECK(launch_fill_zero(prog->outs[o].valid, (int64_t)vb, (int)std::min<size_t>(1024, vb / 2048 + 1),
                                    S(stream)));
        }
      HIPCHECK(launch_program(prog->c, prog->compact, prog->grid, S(stream), prog->lb.blob, prog->lb.ws,
                              prog->lb.bd, prog->lb.n_chunks, prog->evs, prog->last_sel));
      HIPCHECK(hipEventSynchronize(prog->evs[ns]));
      for (int k = 0; k < ns; k++) {
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, prog->evs[k], prog->evs[k + 1]));
        total[k] += ms;
      }
    }
    for (int k = 0; k < ns; k++) stage_ms[k] = total[k] / iters;
    if (n_stages) *n_stages = ns;
  });
}

int blz_program_time(blz_program prog, blz_stream_t stream, int iters, double *mean_ms, char *err,
                     size_t err_len) {
  double st[3] = {0, 0, 0};
  int ns = 0;
  const int rc = blz_program_time_stages(prog, stream, iters, st, &ns, err, err_len);
  if (rc == BLZ_OK && mean_ms) *mean_ms = st[0] + st[1] + st[2];
  return rc;
}

int blz_program_info(blz_program prog, int *s, int *r, int *n_ops, int *n_vops, int *n_slots, int *grid) {
  if (!prog) return BLZ_ERR_INVALID;
  if (s) *s = prog->c.S;
  if (r) *r = prog->c.R;
  if (n_ops) *n_ops = prog->c.n_ops;
  if (n_vops) *n_vops = prog->c.n_vops;
  if (n_slots) *n_slots = prog->c.n_slots;
  if (grid) *grid = prog->grid;
  return BLZ_OK;
}

int blz_program_tile(blz_program prog, int *tile_rows, int *lds_bytes, int *kernel) {
  if (!prog) return BLZ_ERR_INVALID;
  if (tile_rows) *tile_rows = prog->c.T;
  if (lds_bytes) *lds_bytes = prog->c.lds_bytes;
  if (kernel) *kernel = prog->c.jit ? 5 : prog->c.vm ? 4 : 3;
  return BLZ_OK;
}

int blz_program_predicate(blz_program prog, int *direct, int *terms, int *wmask) {
  if (!prog) return BLZ_ERR_INVALID;
  const bool d = prog->compact && prog->c.vm && prog->c.vm_direct_pred;
  if (direct) *direct = d ? 1 : 0;
  if (terms) *terms = d ? prog->c.vm_direct_terms : 0;
  if (wmask) *wmask = d ? prog->c.vm_direct_wmask : 0;
  return BLZ_OK;
}

// The kernel source the compiled tier writes for a plan over a frame of this shape
// (dtypes, bitmaps present, null counts, rows); nothing is allocated or launched,
// column data is not read.
namespace {
// The compiled tier's source of a filter / projection over `frame` (the text
// make_program compiles for kernel 5). Throws as the planner and generator do.
std::string plan_jit_source(gdf_column *const *frame, int n_frame, const char *filter_qp, const char *proj_qp,
                            int scalar_mode) {
  check_frame(frame, n_frame);
  const int64_t n = frame[0]->size;
  const auto dt = frame_dtypes(frame, n_frame);
  InputMap im(n_frame);
  planner::ExprPlan fp;
  const bool has_filter = filter_qp != nullptr;
  if (has_filter) {
    fp = planner::plan_expression(dt, condition_of(filter_qp));
    if (fp.bare) im.use(frame, fp.bare_col);
    else im.remap(frame, fp.spec);
  }
  planner::ProjectPlan pp;
  std::vector<planner::ProjOut> outs;
  if (proj_qp) {
    pp = planner::parse_project(dt, proj_qp);
    if (pp.has_eval) im.remap(frame, pp.spec);
    outs = pp.outputs;
  } else {
    for (int i = 0; i < n_frame; i++) outs.push_back({1, i, dt[i], ""});
  }
  CompileRequest req;
  req.n_rows = n;
  req.scalar_mode = scalar_mode;
  req.compact = has_filter;
  req.kernel = 5;
  // stand-in output buffers: the source depends only on which are present
  static uint8_t stand_in[64];
  for (const auto &o : outs) {
    const bool with_valid = o.kind == 0 ? true : frame[o.index]->valid != nullptr;
    if (o.kind == 0) req.outputs.push_back({stand_in, with_valid ? stand_in : nullptr, o.dtype, 0, o.index});
    else req.outputs.push_back({stand_in, with_valid ? stand_in : nullptr, o.dtype, 1, im.use(frame, o.index)});
  }
  req.inputs = im.inputs;
  if (has_filter) {
    if (fp.bare) req.bare_stencil_input = im.frame_to_dev[fp.bare_col];
    else req.pred = &fp.spec;
  }
  if (proj_qp && pp.has_eval) req.proj = &pp.spec;
  if (req.inputs.empty()) invalid("nothing to evaluate");
  return compile_program(req).jit_src;
}
}  // namespace

int blz_program_source(gdf_column *const *frame, int n_frame, const char *filter_qp, const char *proj_qp,
                       int scalar_mode, char *src, size_t src_cap, size_t *src_len, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    const std::string js = plan_jit_source(frame, n_frame, filter_qp, proj_qp, scalar_mode);
    if (src_len) *src_len = js.size();
    if (src && src_cap > 0) {
      const size_t k = std::min(src_cap - 1, js.size());
      std::memcpy(src, js.data(), k);
      src[k] = 0;
    }
  });
}

// Compiles a kernel source for the device (no device needed): 0 when it compiles,
// the compiler's log in err otherwise.
int blz_jit_compile(const char *src, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!src) invalid("src is NULL");
    jit_compile_only(src);
  });
}

int blz_jit_compile_many(const char *const *srcs, int n, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!srcs || n < 0) invalid("bad arguments");
    std::vector<std::string> v;
    for (int i = 0; i < n; i++) {
      if (!srcs[i]) invalid("a source is NULL");
      v.emplace_back(srcs[i]);
    }
    jit_compile_many(v);
  });
}

int blz_jit_counters(long long *compiles, long long *hits, double *compile_ms) {
  jit_counters(compiles, hits, compile_ms);
  return BLZ_OK;
}

int blz_program_destroy(blz_program prog) {
  if (!prog) return BLZ_OK;
  // the serving loop destroys a program only after its batch completed; elsewhere
  // the device finishes its work first
  if (!prog->arena) (void)hipDeviceSynchronize();
  arena_or_free(prog->arena, prog->out_block);
  arena_or_free(prog->arena, prog->lb.base);
  for (auto &e : prog->evs)
    if (e) (void)hipEventDestroy(e);
  delete prog;
  return BLZ_OK;
}

// ---- memory / streams --------------------------------------------------------------

int blz_device_malloc(void **ptr, size_t bytes) {
  return hipMalloc(ptr, bytes) == hipSuccess ? BLZ_OK : BLZ_ERR_OOM;
}
int blz_device_free(void *ptr) { return hipFree(ptr) == hipSuccess ? BLZ_OK : BLZ_ERR_HIP; }
int blz_memcpy_htod(void *dst, const void *src, size_t bytes, blz_stream_t stream) {
  if (bytes == 0) return BLZ_OK;
  if (hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(stream)) != hipSuccess) return BLZ_ERR_HIP;
  return hipStreamSynchronize(S(stream)) == hipSuccess ? BLZ_OK : BLZ_ERR_HIP;
}
int blz_memcpy_dtoh(void *dst, const void *src, size_t bytes, blz_stream_t stream) {
  if (bytes == 0) return BLZ_OK;
  if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(stream)) != hipSuccess) return BLZ_ERR_HIP;
  return hipStreamSynchronize(S(stream)) == hipSuccess ? BLZ_OK : BLZ_ERR_HIP;
}
int blz_memset_device(void *dst, int value, size_t bytes, blz_stream_t stream) {
  if (bytes == 0) return BLZ_OK;
  if (hipMemsetAsync(dst, value, bytes, S(stream)) != hipSuccess) return BLZ_ERR_HIP;
  return hipStreamSynchronize(S(stream)) == hipSuccess ? BLZ_OK : BLZ_ERR_HIP;
}
int blz_device_copy(void *dst, const void *src, size_t bytes, blz_stream_t stream) {
  if (bytes == 0) return BLZ_OK;
  if (bytes % 16 || (uintptr_t)dst % 16 || (uintptr_t)src % 16) return BLZ_ERR_INVALID;
  const int grid = (int)std::min<size_t>(bytes / 16 / 256 + 1, (size_t)device_cus() * 16);
  if (launch_copy(dst, src, (int64_t)bytes, grid, S(stream)) != hipSuccess) return BLZ_ERR_HIP;
  return hipStreamSynchronize(S(stream)) == hipSuccess ? BLZ_OK : BLZ_ERR_HIP;
}

int blz_stream_create(blz_stream_t *stream) {
  hipStream_t s;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return BLZ_ERR_HIP;
  *stream = (blz_stream_t)s;
  return BLZ_OK;
}
int blz_stream_synchronize(blz_stream_t stream) {
  return hipStreamSynchronize(S(stream)) == hipSuccess ? BLZ_OK : BLZ_ERR_HIP;
}
int blz_release_launch_buffers(void) {
  std::lock_guard<std::mutex> g(g_lb_mu);
  for (auto &e : g_lb) {
    std::lock_guard<std::mutex> b(e.second->busy);  // after any call in flight
    if (e.second->p) (void)hipFree(e.second->p);
    if (e.second->host) (void)hipHostFree(e.second->host);
    e.second->p = nullptr;
    e.second->bytes = 0;
    e.second->host = nullptr;
  }
  return BLZ_OK;
}

int blz_stream_destroy(blz_stream_t stream) {
  lb_cache_drop(S(stream));
  return hipStreamDestroy(S(stream)) == hipSuccess ? BLZ_OK : BLZ_ERR_HIP;
}
int blz_free_column(gdf_column *column) {
  if (!column) return BLZ_OK;
  hipError_t a = column->data ? hipFree(column->data) : hipSuccess;
  hipError_t b = column->valid ? hipFree(column->valid) : hipSuccess;
  std::memset(column, 0, sizeof(*column));
  return (a == hipSuccess && b == hipSuccess) ? BLZ_OK : BLZ_ERR_HIP;
}

int blz_plan_expression(const int32_t *frame_dtypes_in, int n_frame, const char *expression, int cap,
                        column_index_type *left, column_index_type *right, column_index_type *outputs,
                        int32_t *binops, int32_t *unops, gdf_scalar *lscalars, gdf_scalar *rscalars,
                        int *n_ops, int32_t *in_cols, int *n_in, column_index_type *final_pos, char *err,
                        size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!expression) invalid("expression is NULL");
    std::vector<int> dt(frame_dtypes_in, frame_dtypes_in + n_frame);
    planner::ExprPlan ep = planner::plan_expression(dt, expression);
    if (ep.bare) {
      *n_ops = 0;
      *n_in = 1;
      in_cols[0] = ep.bare_col;
      *final_pos = -1;
      return;
    }
    const int k = (int)ep.spec.left.size();
    if (k > cap || (int)ep.spec.in_dev.size() > cap) invalid("plan larger than cap");
    for (int i = 0; i < k; i++) {
      left[i] = ep.spec.left[i];
      right[i] = ep.spec.right[i];
      outputs[i] = ep.spec.outs[i];
      binops[i] = ep.spec.binops[i];
      unops[i] = ep.spec.unops[i];
      lscalars[i] = ep.spec.lsc[i];
      rscalars[i] = ep.spec.rsc[i];
    }
    *n_ops = k;
    for (size_t i = 0; i < ep.spec.in_dev.size(); i++) in_cols[i] = ep.spec.in_dev[i];
    *n_in = (int)ep.spec.in_dev.size();
    *final_pos = ep.spec.final_pos[0];
  });
}

namespace {
// Copies the ops of p from op `first` on into the caller's arrays (capacity cap).
void export_ops(const PlanSpec &p, size_t first, int cap, column_index_type *left, column_index_type *right,
                column_index_type *outputs, int32_t *binops, int32_t *unops, gdf_scalar *lscalars,
                gdf_scalar *rscalars, int *n_ops) {
  const int k = (int)(p.left.size() - first);
  if (k > cap) invalid("plan larger than cap (" + std::to_string(k) + " ops)");
  for (int i = 0; i < k; i++) {
    left[i] = p.left[first + i];
    right[i] = p.right[first + i];
    outputs[i] = p.outs[first + i];
    binops[i] = p.binops[first + i];
    unops[i] = p.unops[first + i];
    lscalars[i] = p.lsc[first + i];
    rscalars[i] = p.rsc[first + i];
  }
  *n_ops = k;
}
}  // namespace

int blz_add_expression_to_plan(const char *expression, column_index_type expression_position,
                               column_index_type num_outputs, column_index_type num_inputs,
                               const column_index_type *new_input_indices, int n_frame, int cap,
                               column_index_type *left, column_index_type *right, column_index_type *outputs,
                               int32_t *binops, int32_t *unops, gdf_scalar *lscalars, gdf_scalar *rscalars,
                               int *n_ops, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!expression || !n_ops || n_frame < 0 || (n_frame > 0 && !new_input_indices)) invalid("bad arguments");
    std::vector<int> idx(new_input_indices, new_input_indices + n_frame);
    PlanSpec p;
    planner::add_expression_to_plan(p, expression, expression_position, num_outputs, num_inputs, idx);
    export_ops(p, 0, cap, left, right, outputs, binops, unops, lscalars, rscalars, n_ops);
  });
}

int blz_parse_project_plan(const int32_t *frame_dtypes_in, int n_frame, const char *query_part, int cap,
                           column_index_type *left, column_index_type *right, column_index_type *outputs,
                           int32_t *binops, int32_t *unops, gdf_scalar *lscalars, gdf_scalar *rscalars, int *n_ops,
                           int32_t *in_cols, int *n_in, column_index_type *final_pos, int *n_eval,
                           int32_t *out_kind, int32_t *out_index, int32_t *out_dtype, int max_out, int *n_out,
                           char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!query_part || !n_ops || !n_in || !n_eval || !n_out) invalid("bad arguments");
    std::vector<int> dt(frame_dtypes_in, frame_dtypes_in + n_frame);
    const planner::ProjectPlan pp = planner::parse_project(dt, query_part);
    if ((int)pp.outputs.size() > max_out) invalid("more outputs than max_out");
    if ((int)pp.spec.in_dev.size() > cap) invalid("more input columns than cap");
    export_ops(pp.spec, 0, cap, left, right, outputs, binops, unops, lscalars, rscalars, n_ops);
    for (size_t i = 0; i < pp.spec.in_dev.size(); i++) in_cols[i] = pp.spec.in_dev[i];
    *n_in = (int)pp.spec.in_dev.size();
    for (size_t i = 0; i < pp.spec.final_pos.size(); i++) final_pos[i] = pp.spec.final_pos[i];
    *n_eval = (int)pp.spec.final_pos.size();
    for (size_t i = 0; i < pp.outputs.size(); i++) {
      out_kind[i] = pp.outputs[i].kind;
      out_index[i] = pp.outputs[i].index;
      out_dtype[i] = pp.outputs[i].dtype;
    }
    *n_out = (int)pp.outputs.size();
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches: many independent queries, one persistent launch per kernel kind.

namespace {

struct SubBatch {
  std::vector<int> members;   // program indices
  VmBatch *d_bd = nullptr;    // device descriptor (VmBatch + VmPlanRef[])
  int lds = 0, lds_gather = 0;
  int slow = 0;  // op set level (max over the programs)
  int64_t n_tiles = 0, n_chunks = 0;
  int grid = 1, grid_gather = 1;
};

}  // namespace

struct blz_batch_s {
  std::vector<blz_program> progs;
  int r = 0, t = 0;
  SubBatch proj, filt;
  std::vector<int> others;           // queries the tile VM cannot run: launched one by one
  uint8_t *d_mem = nullptr;          // descriptors, zero-range table, scan counter
  void **d_range_ptrs = nullptr;
  int64_t *d_range_bytes = nullptr;
  int n_ranges = 0;
  unsigned int *d_counter = nullptr;
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cap = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int tuned = 0;
  int n_compiled = 0;        // queries on the compiled tier (their own launches)
  std::vector<uint8_t> img;  // host image of the descriptors
  DeviceArena *arena = nullptr;  // device memory from the serving loop's arena (else the driver)
  hipStream_t mem_st = nullptr;
};

namespace {

hipError_t batch_enqueue_direct(blz_batch b, hipStream_t st) {
  hipError_t e = launch_fill_ranges(b->d_range_ptrs, b->d_range_bytes, b->n_ranges, 512, st);
  for (size_t i = 0; e == hipSuccess && i < b->others.size(); i++) e = program_enqueue(b->progs[b->others[i]], st);
  if (e == hipSuccess && !b->proj.members.empty())
    e = launch_vm(b->r, 0, b->proj.slow, b->proj.grid, b->proj.lds, st, b->proj.d_bd);
  if (e == hipSuccess && !b->filt.members.empty()) {
    e = launch_vm(b->r, 1, b->filt.slow, b->filt.grid, b->filt.lds, st, b->filt.d_bd);
    if (e == hipSuccess) e = launch_slot_scan(b->filt.d_bd, b->filt.n_chunks, st);
    if (e == hipSuccess)
      e = launch_vm(b->r, 2, b->filt.slow, b->filt.grid_gather, b->filt.lds_gather, st, b->filt.d_bd);
  }
  return e;
}

hipError_t batch_enqueue(blz_batch b, hipStream_t st) {
  if (b->exec) return hipGraphLaunch(b->exec, st);
  return batch_enqueue_direct(b, st);
}

void batch_free(blz_batch b) {
  if (!b) return;
  if (!b->arena) (void)hipDeviceSynchronize();
  arena_or_free(b->arena, b->d_mem);
  b->d_mem = nullptr;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->d_mem) (void)hipFree(b->d_mem);
  for (auto p : b->progs) blz_program_destroy(p);
  delete b;
}

// Builds one batch of fixed shape (r, t). Queries flagged in `compiled` run on the
// compiled tier, each as its own launch (its source compiled beforehand, or now).
// Throws on error.
blz_batch batch_build(const blz_query *qs, int n, int r, int t, bool use_graph, hipStream_t st = nullptr,
                      DeviceArena *arena = nullptr, const std::vector<char> *compiled = nullptr) {
  const bool sync = arena == nullptr;
  auto *b = new blz_batch_s();
  try {
    b->r = r;
    b->t = t;
    for (int i = 0; i < n; i++) {
      const blz_query &q = qs[i];
      if (compiled && (*compiled)[i]) {  // kernel 0: the compiled tier, else the interpreters
        b->progs.push_back(make_program(q.frame, q.n_frame, q.filter_query_part, q.project_query_part,
                                        q.scalar_mode, 0, 0, 0, 0, st, arena));
        b->others.push_back(i);
        b->n_compiled += b->progs.back()->c.jit ? 1 : 0;
        continue;
      }
      blz_program p = nullptr;
      try {
        p = make_program(q.frame, q.n_frame, q.filter_query_part, q.project_query_part, q.scalar_mode, 4, 0, r, t,
                         st, arena);
      } catch (const BlzError &e) {
        throw;  // device or argument errors
      } catch (const std::exception &e) {
        if (std::string(e.what()).find("fit") == std::string::npos) throw;  // not a shape limit
      }
      if (p) {
        b->progs.push_back(p);
        (p->compact ? b->filt : b->proj).members.push_back(i);
      } else {  // a plan outside the shared shape: its own launch in the batch (the tile VM with
                // a shape of its own, else the register-file kernel; not the compiled tier, whose
                // compilation would stall the batch's construction)
        b->progs.push_back(make_program(q.frame, q.n_frame, q.filter_query_part, q.project_query_part,
                                        q.scalar_mode, kInterpAuto, 0, 0, 0, st, arena));
        b->others.push_back(i);
      }
    }
    // One persistent launch sizes its LDS by the maxima over its programs (tile
    // buffers, validity registers, temporaries): while that exceeds 64 KiB, the
    // program with the largest own LDS leaves the shared launch for one of its own.
    auto shared_lds = [&](const std::vector<int> &m) {
      int maxbuf = 0, maxtmp = 0, maxgw = 0, maxgtmp = 0, maxvs = 0;
      for (int i : m) {
        const Compiled &c = b->progs[i]->c;
        maxbuf = std::max(maxbuf, c.vm_bufb);
        maxtmp = std::max(maxtmp, c.lds_bytes - c.vm_tmp_off);
        maxvs = std::max(maxvs, c.n_vslots);
        maxgw = std::max(maxgw, c.vm_g_wave_bytes);
        maxgtmp = std::max(maxgtmp, c.lds_gather - c.vm_g_tmp_off);
      }
      const int vr = 4 * 8 * maxvs * r;
      const int staged = kTileBuffers * maxbuf + vr + maxtmp;
      const int gather = (int)round_up((size_t)(4 * maxgw + vr + maxgtmp), 8) + 8 * ((int)m.size() + 1);
      return std::max(staged, gather);
    };
    for (SubBatch *sb : {&b->proj, &b->filt}) {
      while (sb->members.size() > 1 && shared_lds(sb->members) > 64 * 1024) {
        auto big = std::max_element(sb->members.begin(), sb->members.end(), [&](int x, int y) {
          const Compiled &cx = b->progs[x]->c, &cy = b->progs[y]->c;
          return std::max(cx.lds_bytes, cx.lds_gather) < std::max(cy.lds_bytes, cy.lds_gather);
        });
        b->others.push_back(*big);
        sb->members.erase(big);
      }
    }
    std::sort(b->others.begin(), b->others.end());
    // Host images of the two descriptors and the zero-range table.
    auto build_sub = [&](SubBatch &sb, std::vector<uint8_t> &img, size_t off) {
      VmBatch bd{};
      std::vector<VmPlanRef> refs;
      int maxbuf = 0, maxtmp = 0, maxgw = 0, maxgtmp = 0, maxvs = 0;
      int64_t tile0 = 0, chunk0 = 0, gchunks = 0;
      for (int i : sb.members) {
        blz_program p = b->progs[i];
        const Compiled &c = p->c;
        VmPlanRef ref{};
        ref.blob = p->lb.blob;
        ref.ws = p->lb.ws;
        ref.tile0 = tile0;
        ref.chunk0 = chunk0;
        refs.push_back(ref);
        const int64_t tiles = (c.n_rows + c.T - 1) / c.T;
        tile0 += tiles;
        gchunks += (c.n_rows + 64 * b->r - 1) / (64 * b->r);
        chunk0 += (tiles * 4 + scan_chunk_slots(c.T) - 1) / scan_chunk_slots(c.T);
        maxbuf = std::max(maxbuf, c.vm_bufb);
        maxtmp = std::max(maxtmp, c.lds_bytes - c.vm_tmp_off);
        maxvs = std::max(maxvs, c.n_vslots);
        maxgw = std::max(maxgw, c.vm_g_wave_bytes);
        maxgtmp = std::max(maxgtmp, c.lds_gather - c.vm_g_tmp_off);
        sb.slow = std::max(sb.slow, c.vm_slow);
      }
      bd.n_plans = (int)refs.size();
      bd.bufb = maxbuf;
      bd.vr_wave_words = maxvs * b->r;
      bd.vr_off = kTileBuffers * maxbuf;
      bd.tmp_off = bd.vr_off + 4 * 8 * bd.vr_wave_words;
      bd.g_buf_off = 0;
      bd.g_wave_bytes = maxgw;
      bd.g_vr_off = 4 * maxgw;
      bd.g_tmp_off = bd.g_vr_off + 4 * 8 * bd.vr_wave_words;
      bd.g_cpre_off = (int32_t)round_up((size_t)(bd.g_tmp_off + maxgtmp), 8);
      bd.n_tiles = tile0;
      bd.n_chunks = chunk0;
      bd.chunk_counter = b->d_counter;
      bd.plans = (const VmPlanRef *)(b->d_mem + off + sizeof(VmBatch));
      sb.lds = bd.tmp_off + maxtmp;
      sb.lds_gather = bd.g_cpre_off + 8 * (bd.n_plans + 1);
      if (sb.lds > 64 * 1024 || sb.lds_gather > 64 * 1024) invalid("batch needs more than 64 KiB of LDS");
      sb.n_tiles = tile0;
      sb.n_chunks = chunk0;
      sb.d_bd = (VmBatch *)(b->d_mem + off);
      std::memcpy(img.data() + off, &bd, sizeof(bd));
      std::memcpy(img.data() + off + sizeof(VmBatch), refs.data(), refs.size() * sizeof(VmPlanRef));
      const int cus = device_cus();
      if (!sb.members.empty()) {
        const Compiled &c0 = b->progs[sb.members[0]]->c;
        Compiled probe = c0;  // occupancy depends on (r, mode, slow, lds) only
        probe.vm_slow = sb.slow;
        const int mode = sb.members.empty() ? 0 : (b->progs[sb.members[0]]->compact ? 1 : 0);
        sb.grid = (int)std::max<int64_t>(
            1, std::min<int64_t>(tile0, (int64_t)cus * vm_per_cu(probe, mode, sb.lds) * (mode == 1 ? 8 : 1)));
        sb.grid_gather = (int)std::max<int64_t>(
            1, std::min<int64_t>((gchunks + 3) / 4, (int64_t)cus * vm_per_cu(probe, 2, std::max(sb.lds_gather, 1))));
      }
    };
    const size_t proj_sz = round_up(sizeof(VmBatch) + sizeof(VmPlanRef) * b->proj.members.size(), 256);
    const size_t filt_sz = round_up(sizeof(VmBatch) + sizeof(VmPlanRef) * b->filt.members.size(), 256);
    // zero ranges: each program's counters, the scan counter
    std::vector<void *> rp;
    std::vector<int64_t> rb;
    for (size_t i = 0; i < b->progs.size(); i++) {
      blz_program p = b->progs[i];
      if (std::find(b->others.begin(), b->others.end(), (int)i) != b->others.end()) continue;
      rp.push_back(p->lb.base + p->lb.zero_off);
      rb.push_back((int64_t)p->lb.zero_bytes);
    }
    const size_t ctr_off = proj_sz + filt_sz;
    const size_t ptr_off = ctr_off + 256;
    const size_t bytes_off = round_up(ptr_off + sizeof(void *) * (rp.size() + 1), 256);
    const size_t total = round_up(bytes_off + sizeof(int64_t) * (rb.size() + 1), 256);
    b->d_mem = (uint8_t *)arena_or_malloc(arena, total);
    b->arena = arena;
    b->mem_st = st;
    b->d_counter = (unsigned int *)(b->d_mem + ctr_off);
    rp.push_back(b->d_counter);
    rb.push_back(8);
    b->n_ranges = (int)rp.size();
    b->d_range_ptrs = (void **)(b->d_mem + ptr_off);
    b->d_range_bytes = (int64_t *)(b->d_mem + bytes_off);
    std::vector<uint8_t> &img = b->img;
    img.assign(total, 0);
    build_sub(b->proj, img, 0);
    build_sub(b->filt, img, proj_sz);
    std::memcpy(img.data() + ptr_off, rp.data(), sizeof(void *) * rp.size());
    std::memcpy(img.data() + bytes_off, rb.data(), sizeof(int64_t) * rb.size());
    if (sync) HIPCHECK(hipMemcpy(b->d_mem, img.data(), total, hipMemcpyHostToDevice));
    else HIPCHECK(hipMemcpyAsync(b->d_mem, img.data(), total, hipMemcpyHostToDevice, st));
    HIPCHECK(hipEventCreate(&b->ev0));
    HIPCHECK(hipEventCreate(&b->ev1));
    if (use_graph) {
      // Capture the launch sequence once (kernels only: no memset nodes).
      HIPCHECK(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
      HIPCHECK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
      const hipError_t e = batch_enqueue_direct(b, b->cap);
      hipGraph_t g = nullptr;
      const hipError_t e2 = hipStreamEndCapture(b->cap, &g);
      HIPCHECK(e);
      HIPCHECK(e2);
      b->graph = g;
      HIPCHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->use_graph = true;
    }
    if (sync) HIPCHECK(hipDeviceSynchronize());
  } catch (...) {
    if (arena) (void)hipStreamSynchronize(st);  // as in make_program
    batch_free(b);
    throw;
  }
  return b;
}

double batch_time_ms(blz_batch b, hipStream_t st, int iters) {
  double total = 0;
  for (int i = 0; i < iters; i++) {
    HIPCHECK(hipEventRecord(b->ev0, st));
    HIPCHECK(batch_enqueue(b, st));
    HIPCHECK(hipEventRecord(b->ev1, st));
    HIPCHECK(hipEventSynchronize(b->ev1));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    total += ms;
  }
  return total / iters;
}

// The autotuner's per-query tier. The largest queries (rows x bytes per row of
// their frame), each of at least kJitMinRows rows, at most kJitMaxPlans of them and
// until they hold 90 % of the batch's bytes, are compiled for the compiled tier in
// one program; each is then timed alone on the compiled tier and on the tile VM at
// the batch's shape (r, t), and flagged in `compiled` when the compiled kernel is
// faster (every one that compiles when `always`). Returns the number flagged.
constexpr int kJitMaxPlans = 40;
constexpr int64_t kJitMinRows = 1 << 20;

int choose_compiled(const blz_query *qs, int n, int r, int t, bool always, std::vector<char> &compiled) {
  compiled.assign(n, 0);
  std::vector<std::pair<double, int>> by;
  double total = 0;
  for (int i = 0; i < n; i++) {
    double bpr = 0;
    for (int c = 0; c < qs[i].n_frame; c++)
      bpr += blz_dtype_width(qs[i].frame[c]->dtype) + (qs[i].frame[c]->valid ? 0.125 : 0.0);
    const double bytes = (double)qs[i].frame[0]->size * bpr;
    total += bytes;
    by.push_back({bytes, i});
  }
  std::sort(by.begin(), by.end(), [](const auto &x, const auto &y) { return x.first > y.first; });
  std::vector<int> cand;
  std::vector<std::string> srcs;
  double covered = 0;
  for (const auto &bi : by) {
    const blz_query &q = qs[bi.second];
    if ((int)cand.size() >= kJitMaxPlans || q.frame[0]->size < kJitMinRows || covered >= 0.9 * total) break;
    covered += bi.first;
    try {
      srcs.push_back(plan_jit_source(q.frame, q.n_frame, q.filter_query_part, q.project_query_part, q.scalar_mode));
      cand.push_back(bi.second);
    } catch (const BlzError &) {
      throw;
    } catch (const std::exception &) {  // a plan the generator does not take stays on the tile VM
    }
  }
  if (cand.empty()) return 0;
  jit_compile_many(srcs);
  auto timed = [](blz_program p) {
    double ms = 0;
    char e2[256];
    int rc = blz_program_time(p, nullptr, 1, &ms, e2, sizeof(e2));  // warm-up
    if (rc == BLZ_OK) rc = blz_program_time(p, nullptr, 3, &ms, e2, sizeof(e2));
    if (rc != BLZ_OK) throw BlzError(rc, e2);
    return ms;
  };
  int k = 0;
  for (int i : cand) {
    const blz_query &q = qs[i];
    blz_program pc = nullptr, pv = nullptr;
    try {
      pc = make_program(q.frame, q.n_frame, q.filter_query_part, q.project_query_part, q.scalar_mode, 5, 0, 0, 0);
      if (always) {
        compiled[i] = 1;
        k++;
        blz_program_destroy(pc);
        continue;
      }
      try {
        pv = make_program(q.frame, q.n_frame, q.filter_query_part, q.project_query_part, q.scalar_mode, 4, 0, r, t);
      } catch (const BlzError &) {
        throw;
      } catch (const std::exception &) {  // outside the shared shape: its own interpreter launch
        pv = make_program(q.frame, q.n_frame, q.filter_query_part, q.project_query_part, q.scalar_mode,
                          kInterpAuto, 0, 0, 0);
      }
      if (timed(pc) < timed(pv)) {
        compiled[i] = 1;
        k++;
      }
    } catch (const BlzError &) {
      blz_program_destroy(pc);
      blz_program_destroy(pv);
      throw;
    } catch (const std::exception &) {  // not compiled after all: stays on the tile VM
    }
    blz_program_destroy(pc);
    blz_program_destroy(pv);
  }
  return k;
}

}  // namespace

int blz_batch_create(blz_batch *batch, const blz_query *queries, int n_queries, int r, int t, int use_graph,
                     char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!batch || !queries || n_queries <= 0) invalid("bad arguments");
    *batch = nullptr;
    if (r != 0) {
      *batch = batch_build(queries, n_queries, r, t, use_graph != 0);
      return;
    }
    // Autotuner: every (rows per lane, tile rows) candidate on this batch; the
    // fastest of three timed executions after one warm-up wins.
    static const int cand[][2] = {{4, 1024}, {4, 2048}, {8, 2048}, {2, 512}, {2, 1024}};
    blz_batch best = nullptr;
    double best_ms = 0;
    int tried = 0;
    for (auto &c : cand) {
      blz_batch b = nullptr;
      try {
        b = batch_build(queries, n_queries, c[0], c[1], false);
      } catch (const std::exception &) {
        continue;  // a shape that does not fit some query
      }
      tried++;
      HIPCHECK(batch_enqueue(b, nullptr));
      const double ms = batch_time_ms(b, nullptr, 3);
      if (!best || ms < best_ms) {
        if (best) batch_free(best);
        best = b;
        best_ms = ms;
      } else {
        batch_free(b);
      }
    }
    if (!best) invalid("no tile VM shape fits every query of the batch");
    const int br = best->r, bt = best->t;
    // Then the tier of each large query; the batch with the chosen ones compiled
    // replaces the tile-VM batch only if it runs faster.
    std::vector<char> compiled;
    const bool always = t < 0;
    if (choose_compiled(queries, n_queries, br, bt, always, compiled) > 0) {
      blz_batch b = batch_build(queries, n_queries, br, bt, false, nullptr, nullptr, &compiled);
      HIPCHECK(batch_enqueue(b, nullptr));
      const double ms = batch_time_ms(b, nullptr, 3);
      if (always || ms < best_ms) {
        batch_free(best);
        best = b;
        best_ms = ms;
      } else {
        batch_free(b);
        compiled.assign(n_queries, 0);
      }
    }
    if (use_graph) {
      batch_free(best);
      best = batch_build(queries, n_queries, br, bt, true, nullptr, nullptr, compiled.empty() ? nullptr : &compiled);
    }
    best->tuned = tried;
    *batch = best;
  });
}

int blz_batch_launch(blz_batch batch, blz_stream_t stream, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!batch) invalid("batch is NULL");
    HIPCHECK(batch_enqueue(batch, S(stream)));
  });
}

int blz_batch_result(blz_batch batch, blz_stream_t stream, int q, int64_t *n_kept, int32_t *null_counts,
                     gdf_column *out_columns, int max_out, int *n_out, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!batch || q < 0 || q >= (int)batch->progs.size()) invalid("bad query index");
    blz_program p = batch->progs[q];
    if ((int)p->outs.size() > max_out) invalid("more outputs than out_columns can hold");
    int64_t kept = 0;
    std::vector<int32_t> nulls(std::max(p->n_out, 1), 0);
    char e2[256];
    const int rc = blz_program_result(p, stream, &kept, nulls.data(), e2, sizeof(e2));
    if (rc != BLZ_OK) throw BlzError(rc, e2);
    if (n_kept) *n_kept = kept;
    for (int o = 0; o < p->n_out; o++) {
      if (null_counts) null_counts[o] = nulls[o];
      if (out_columns) {
        out_columns[o] = p->outs[o];
        out_columns[o].size = (gdf_size_type)kept;
        out_columns[o].null_count = nulls[o];
      }
    }
    if (n_out) *n_out = p->n_out;
  });
}

int blz_batch_time(blz_batch batch, blz_stream_t stream, int iters, double *mean_ms, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!batch || iters <= 0 || !mean_ms) invalid("bad arguments");
    *mean_ms = batch_time_ms(batch, S(stream), iters);
  });
}

int blz_batch_info(blz_batch batch, int *r, int *t, int *n_queries, int *graph, int *tuned_candidates) {
  if (!batch) return BLZ_ERR_INVALID;
  if (r) *r = batch->r;
  if (t) *t = batch->t;
  if (n_queries) *n_queries = (int)batch->progs.size();
  if (graph) *graph = batch->use_graph ? 1 : 0;
  if (tuned_candidates) *tuned_candidates = batch->tuned;
  return BLZ_OK;
}

int blz_batch_query_kernel(blz_batch batch, int q, int *kernel, int *shared) {
  if (!batch || q < 0 || q >= (int)batch->progs.size()) return BLZ_ERR_INVALID;
  const bool own = std::find(batch->others.begin(), batch->others.end(), q) != batch->others.end();
  if (shared) *shared = own ? 0 : 1;
  if (kernel) return blz_program_tile(batch->progs[q], nullptr, nullptr, kernel);
  return BLZ_OK;
}

int blz_batch_destroy(blz_batch batch) {
  batch_free(batch);
  return BLZ_OK;
}

int blz_synth_column(gdf_column *column, uint64_t seed, int dist, double a, double b, double null_rate,
                     blz_stream_t stream, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!column || column->size < 0) invalid("bad column");
    if (column->size == 0) return;
    if (blz_dtype_width(column->dtype) == 0 || column->dtype == GDF_BOOL8 || column->dtype > GDF_FLOAT64)
      invalid("unsupported dtype");
    if (column->valid == nullptr && null_rate > 0) invalid("null_rate > 0 needs a validity buffer");
    HIPCHECK(launch_synth(column->data, null_rate > 0 ? column->valid : nullptr, column->size, column->dtype, seed,
                          dist, a, b, null_rate, S(stream)));
  });
}

// ---------------------------------------------------------------------------
// Serving: a request loop over device-resident tables. Queries are admitted
// against an HBM budget (their outputs and workspace), batched as they arrive and
// launched on one of several streams, so batches admitted while others run
// overlap on the device (continuous batching). A query's outputs stay resident
// until it is released.

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct SrvBatch {
  blz_batch b = nullptr;
  hipEvent_t done = nullptr;
  hipStream_t st = nullptr;
  std::vector<int64_t> tickets;
  int live = 0;          // queries not yet released
  size_t ws_bytes = 0;   // budget held for workspace until the batch is freed
};

struct SrvQuery {
  std::vector<gdf_column *> frame;
  std::string filt, proj;
  bool has_filt = false, has_proj = false;
  int scalar_mode = 0;
  double t_submit = 0, t_done = 0;
  int state = 0;  // 0 pending, 1 running, 2 done, 3 failed
  size_t out_bytes = 0, ws_bytes = 0;
  std::string err;
  SrvBatch *batch = nullptr;
  int index = -1;
};

}  // namespace

// A served query whose result goes to the result repository under `token`.
struct SrvHandoff {
  int64_t ticket = 0;
  uint64_t token = 0;
  std::vector<std::string> names;
};

struct blz_server_s {
  std::mutex mu;
  std::condition_variable cv_work, cv_done, cv_handoff;
  std::deque<SrvHandoff> handoff;  // tickets to park in the result repository, in submission order
  std::thread finisher;            // started by the first blz_server_submit_token
  std::deque<int64_t> pending;
  std::unordered_map<int64_t, SrvQuery> qs;
  std::vector<SrvBatch *> inflight, held;
  size_t budget = 0, used = 0;
  DeviceArena *arena = nullptr;  // outputs, workspaces and descriptors of admitted queries
  int max_batch = 64, n_streams = 4, r = 4, next_stream = 0, device = 0;
  std::vector<hipStream_t> streams;
  std::thread worker;
  bool stop = false;
  int64_t next_ticket = 1;
  int64_t n_batches = 0, peak_inflight = 0;
  std::string fatal;
};

namespace {

// Servers alive: a parked result releases its ticket only while its server exists.
std::mutex g_live_mu;
std::set<blz_server> g_live;

void srv_free_batch(blz_server s, SrvBatch *sb) {
  batch_free(sb->b);
  if (sb->done) (void)hipEventDestroy(sb->done);
  s->used -= std::min(s->used, sb->ws_bytes);
  s->held.erase(std::remove(s->held.begin(), s->held.end(), sb), s->held.end());
  delete sb;
}

void srv_loop(blz_server s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  while (!s->stop) {
    // completions
    bool any_done = false;
    for (size_t i = 0; i < s->inflight.size();) {
      SrvBatch *sb = s->inflight[i];
      const hipError_t q = hipEventQuery(sb->done);
      if (q == hipErrorNotReady) { i++; continue; }
      const double t = now_s();
      for (int64_t tk : sb->tickets) {
        SrvQuery &sq = s->qs[tk];
        sq.t_done = t;
        sq.state = q == hipSuccess ? 2 : 3;
        if (q != hipSuccess) sq.err = hipGetErrorString(q);
      }
      s->held.push_back(sb);
      s->inflight.erase(s->inflight.begin() + i);
      any_done = true;
    }
    if (any_done) s->cv_done.notify_all();
    // admission: FIFO while the budget and a stream slot allow
    if (!s->pending.empty() && (int)s->inflight.size() < s->n_streams) {
      std::vector<int64_t> take;
      size_t add = 0;
      while (!s->pending.empty() && (int)take.size() < s->max_batch) {
        const SrvQuery &sq = s->qs[s->pending.front()];
        const size_t need = sq.out_bytes + sq.ws_bytes;
        const bool alone = s->inflight.empty() && s->held.empty() && take.empty();
        if (s->used + add + need > s->budget && !alone) break;
        add += need;
        take.push_back(s->pending.front());
        s->pending.pop_front();
      }
      if (!take.empty()) {
        s->used += add;
        std::vector<blz_query> bq;
        std::vector<std::string *> keep;
        for (int64_t tk : take) {
          SrvQuery &sq = s->qs[tk];
          sq.state = 1;
          blz_query q{};
          q.frame = sq.frame.data();
          q.n_frame = (int)sq.frame.size();
          q.filter_query_part = sq.has_filt ? sq.filt.c_str() : nullptr;
          q.project_query_part = sq.has_proj ? sq.proj.c_str() : nullptr;
          q.scalar_mode = sq.scalar_mode;
          bq.push_back(q);
        }
        hipStream_t st = s->streams[s->next_stream++ % s->streams.size()];
        lk.unlock();
        auto *sb = new SrvBatch();
        sb->st = st;
        sb->tickets = take;
        sb->live = (int)take.size();
        std::string err;
        try {
          sb->b = batch_build(bq.data(), (int)bq.size(), s->r, 256 * s->r, false, st, s->arena);
          HIPCHECK(batch_enqueue(sb->b, st));
          HIPCHECK(hipEventCreateWithFlags(&sb->done, hipEventDisableTiming));
          HIPCHECK(hipEventRecord(sb->done, st));
        } catch (const std::exception &e) {
          err = e.what();
        }
        lk.lock();
        if (!err.empty()) {
          for (int64_t tk : take) {
            s->qs[tk].state = 3;
            s->qs[tk].err = err;
            s->qs[tk].t_done = now_s();
          }
          s->used -= std::min(s->used, add);
          if (sb->b) {
            (void)hipStreamSynchronize(st);  // its enqueued work may use the memory batch_free returns
            batch_free(sb->b);
          }
          if (sb->done) (void)hipEventDestroy(sb->done);
          delete sb;
          s->cv_done.notify_all();
        } else {
          for (size_t i = 0; i < take.size(); i++) {
            SrvQuery &sq = s->qs[take[i]];
            sq.batch = sb;
            sq.index = (int)i;
            sb->ws_bytes += sq.ws_bytes;
          }
          s->inflight.push_back(sb);
          s->n_batches++;
          s->peak_inflight = std::max<int64_t>(s->peak_inflight, (int64_t)s->inflight.size());
        }
        continue;
      }
    }
    // nothing admitted: sleep until a submission or a release (budget freed), or
    // poll the running batches
    if (s->inflight.empty()) s->cv_work.wait_for(lk, std::chrono::milliseconds(2));
    else s->cv_work.wait_for(lk, std::chrono::microseconds(20));
  }
}

size_t srv_estimate(gdf_column *const *frame, int n_frame, const char *filt, const char *proj, size_t *ws) {
  const int64_t n = frame[0]->size;
  const auto dt = frame_dtypes(frame, n_frame);
  size_t out = 0;
  if (proj) {
    const auto pp = planner::parse_project(dt, proj);
    for (auto &o : pp.outputs) out += (size_t)n * std::max(1, blz_dtype_width((gdf_dtype)o.dtype)) + n / 8 + 64;
  } else {
    for (int i = 0; i < n_frame; i++) out += (size_t)n * std::max(1, blz_dtype_width((gdf_dtype)dt[i])) + n / 8 + 64;
  }
  if (filt) (void)planner::plan_expression(dt, condition_of(filt));
  // keep words, slot counts and (a filter) the kept-row indices, plus descriptors
  *ws = (size_t)n / 8 + (size_t)n / 64 + (filt ? (size_t)n * 4 : 0) + 64 * 1024;
  return out;
}

}  // namespace

int blz_server_create(blz_server *server, size_t hbm_budget_bytes, int max_batch, int n_streams, int r, char *err,
                      size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!server || max_batch <= 0 || n_streams <= 0) invalid("bad arguments");
    if (r != 2 && r != 4 && r != 8) invalid("r must be 2, 4 or 8");
    auto *s = new blz_server_s();
    s->budget = hbm_budget_bytes;
    s->max_batch = max_batch;
    s->n_streams = n_streams;
    s->r = r;
    HIPCHECK(hipGetDevice(&s->device));
    for (int i = 0; i < n_streams; i++) {
      hipStream_t st;
      HIPCHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      s->streams.push_back(st);
    }
    s->arena = new DeviceArena((size_t)4 << 30);
    s->worker = std::thread(srv_loop, s);
    {
      std::lock_guard<std::mutex> g(g_live_mu);
      g_live.insert(s);
    }
    *server = s;
  });
}

int blz_server_submit(blz_server s, const blz_query *q, int64_t *ticket, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!s || !q || !ticket) invalid("bad arguments");
    check_frame(q->frame, q->n_frame);
    SrvQuery sq;
    sq.frame.assign(q->frame, q->frame + q->n_frame);
    sq.has_filt = q->filter_query_part != nullptr;
    sq.has_proj = q->project_query_part != nullptr;
    if (sq.has_filt) sq.filt = q->filter_query_part;
    if (sq.has_proj) sq.proj = q->project_query_part;
    sq.scalar_mode = q->scalar_mode;
    sq.out_bytes = srv_estimate(q->frame, q->n_frame, q->filter_query_part, q->project_query_part, &sq.ws_bytes);
    std::lock_guard<std::mutex> g(s->mu);
    sq.t_submit = now_s();
    const int64_t tk = s->next_ticket++;
    s->qs.emplace(tk, std::move(sq));
    s->pending.push_back(tk);
    *ticket = tk;
    s->cv_work.notify_one();
  });
}

int blz_server_wait(blz_server s, int64_t ticket, int64_t *n_kept, int32_t *null_counts, gdf_column *out_columns,
                    int max_out, int *n_out, double *latency_ms, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!s) invalid("server is NULL");
    SrvBatch *sb = nullptr;
    int idx = -1;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      auto it = s->qs.find(ticket);
      if (it == s->qs.end()) invalid("unknown ticket");
      // a reference stays valid when a concurrent submission rehashes the map (an
      // iterator does not)
      SrvQuery &sq = it->second;
      s->cv_done.wait(lk, [&] { return sq.state >= 2; });
      if (sq.state == 3) throw BlzError(BLZ_ERR_HIP, "query failed: " + sq.err);
      if (latency_ms) *latency_ms = (sq.t_done - sq.t_submit) * 1e3;
      sb = sq.batch;
      idx = sq.index;
    }
    char e2[256];
    const int rc = blz_batch_result(sb->b, sb->st, idx, n_kept, null_counts, out_columns, max_out, n_out, e2,
                                    sizeof(e2));
    if (rc != BLZ_OK) throw BlzError(rc, e2);
  });
}

int blz_server_release(blz_server s, int64_t ticket) {
  if (!s) return BLZ_ERR_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->qs.find(ticket);
  if (it == s->qs.end() || it->second.state < 2) return BLZ_ERR_INVALID;
  SrvBatch *sb = it->second.batch;
  // a query whose batch never started returned its budget when the build failed
  if (sb) s->used -= std::min(s->used, it->second.out_bytes);
  if (sb && it->second.index >= 0) {
    blz_program p = sb->b->progs[it->second.index];
    arena_or_free(p->arena, p->out_block);  // the query's outputs (its batch has completed)
    p->out_block = nullptr;
    for (auto &c : p->outs) {
      c.data = nullptr;
      c.valid = nullptr;
    }
    if (--sb->live == 0) srv_free_batch(s, sb);
  }
  s->qs.erase(it);
  s->cv_work.notify_one();
  return BLZ_OK;
}

int blz_server_replay(blz_server s, const blz_query *queries, const double *arrival_s, int n, double *latency_ms,
                      int64_t *n_kept, double *wall_s, int64_t *tickets_out, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!s || !queries || !arrival_s || n <= 0) invalid("bad arguments");
    std::vector<int64_t> tickets(n, 0);
    if (tickets_out) std::fill(tickets_out, tickets_out + n, (int64_t)0);
    std::thread waiter;
    std::string werr;
    std::mutex tm;
    std::condition_variable tcv;
    int submitted = 0;
    const double t0 = now_s();
    waiter = std::thread([&] {
      for (int i = 0; i < n; i++) {
        int64_t tk;
        {
          std::unique_lock<std::mutex> lk(tm);
          tcv.wait(lk, [&] { return submitted > i; });
          tk = tickets[i];
        }
        if (tk == 0) continue;
        char e2[256];
        int64_t kept = 0;
        double lat = 0;
        if (blz_server_wait(s, tk, &kept, nullptr, nullptr, 64, nullptr, &lat, e2, sizeof(e2)) != BLZ_OK) {
          if (werr.empty()) werr = e2;
        }
        if (latency_ms) latency_ms[i] = lat;
        if (n_kept) n_kept[i] = kept;
        if (tickets_out) tickets_out[i] = tk;  // the caller reads the results and releases
        else blz_server_release(s, tk);
      }
    });
    for (int i = 0; i < n; i++) {
      const double due = t0 + arrival_s[i];
      while (now_s() < due) {
        const double left = due - now_s();
        if (left > 200e-6) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)((left - 100e-6) * 1e6)));
      }
      int64_t tk = 0;
      char e2[256];
      const int rc = blz_server_submit(s, &queries[i], &tk, e2, sizeof(e2));
      {
        std::lock_guard<std::mutex> lk(tm);
        tickets[i] = rc == BLZ_OK ? tk : 0;
        if (rc != BLZ_OK && werr.empty()) werr = e2;
        submitted = i + 1;
      }
      tcv.notify_all();
    }
    waiter.join();
    if (wall_s) *wall_s = now_s() - t0;
    if (!werr.empty()) invalid("replay: " + werr);
  });
}

namespace {

// The result repository's producer (CalciteInterpreter.cpp:920-960: the query runs
// on its own thread, then update_token parks its frame or its error). Hand-offs are
// taken in submission order; each waits for its query, describes its outputs and
// parks them. The parked frame holds the ticket: freeing the result releases it,
// which frees the outputs and returns their budget.
void srv_finisher(blz_server s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  while (true) {
    s->cv_handoff.wait(lk, [&] { return s->stop || !s->handoff.empty(); });
    if (s->handoff.empty()) break;  // stopping, nothing left
    SrvHandoff h = std::move(s->handoff.front());
    s->handoff.pop_front();
    auto it = s->qs.find(h.ticket);
    bool ran = false;
    if (it != s->qs.end()) {
      SrvQuery &sq = it->second;
      s->cv_done.wait(lk, [&] { return sq.state >= 2 || s->stop; });
      ran = sq.state >= 2;
    }
    lk.unlock();
    std::vector<gdf_column> cols(64);
    int n_out = 0;
    int64_t kept = 0;
    double lat = 0;
    char e2[512] = {0};
    std::string error;
    if (!ran) error = "the server was destroyed before the query ran";
    else if (blz_server_wait(s, h.ticket, &kept, nullptr, cols.data(), (int)cols.size(), &n_out, &lat, e2,
                             sizeof(e2)) != BLZ_OK)
      error = e2;
    cols.resize(error.empty() ? n_out : 0);
    h.names.resize(cols.size());
    const int64_t tk = h.ticket;
    std::function<void()> release;
    if (ran)
      release = [s, tk] {
        std::lock_guard<std::mutex> g(g_live_mu);
        if (g_live.count(s)) blz_server_release(s, tk);
      };
    try {
      rs_park(h.token, std::move(cols), std::move(h.names), error.empty() ? lat : 0.0, error, std::move(release));
    } catch (const std::exception &) {
      // the token was freed before its query finished: the release above ran
    }
    lk.lock();
  }
}

}  // namespace

int blz_server_submit_token(blz_server s, const blz_query *q, uint64_t token, char *err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!s || !q) invalid("bad arguments");
    rs_require_token(token);
    // as in the reference, a query that cannot run parks its error under the token
    // (the call itself succeeds)
    std::vector<std::string> names;
    int64_t tk = 0;
    try {
      if (q->project_query_part) {
        const auto pp = planner::parse_project(frame_dtypes(q->frame, q->n_frame), q->project_query_part);
        for (auto &o : pp.outputs) names.push_back(o.name);
      } else {
        for (int i = 0; i < q->n_frame; i++) names.push_back("$" + std::to_string(i));
      }
      char e2[512];
      const int rc = blz_server_submit(s, q, &tk, e2, sizeof(e2));
      if (rc != BLZ_OK) throw std::runtime_error(e2);
    } catch (const std::exception &e) {
      rs_park(token, {}, {}, 0.0, e.what(), nullptr);
      return;
    }
    std::lock_guard<std::mutex> g(s->mu);
    if (!s->finisher.joinable()) s->finisher = std::thread(srv_finisher, s);
    s->handoff.push_back(SrvHandoff{tk, token, std::move(names)});
    s->cv_handoff.notify_one();
  });
}

int blz_server_stats(blz_server s, int64_t *n_batches, int64_t *peak_inflight, size_t *used_bytes) {
  if (!s) return BLZ_ERR_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  if (n_batches) *n_batches = s->n_batches;
  if (peak_inflight) *peak_inflight = s->peak_inflight;
  if (used_bytes) *used_bytes = s->used;
  return BLZ_OK;
}

int blz_server_destroy(blz_server s) {
  if (!s) return BLZ_OK;
  {
    std::lock_guard<std::mutex> g(g_live_mu);
    g_live.erase(s);
  }
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    s->cv_handoff.notify_all();
  }
  if (s->worker.joinable()) s->worker.join();
  if (s->finisher.joinable()) s->finisher.join();
  (void)hipDeviceSynchronize();
  for (auto sb : s->inflight) {
    batch_free(sb->b);
    if (sb->done) (void)hipEventDestroy(sb->done);
    delete sb;
  }
  for (auto sb : s->held) {
    batch_free(sb->b);
    if (sb->done) (void)hipEventDestroy(sb->done);
    delete sb;
  }
  for (auto st : s->streams) (void)hipStreamDestroy(st);
  delete s->arena;  // every batch has returned its memory
  delete s;
  return BLZ_OK;
}

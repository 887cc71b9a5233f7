/*
 * vm_kernel.hip — the tile VM's slot scan (the single-pass exclusive prefix of the
 * kept-row counts between the predicate and gather kernels) and the launch
 * dispatch over rows per lane. The VM kernels themselves are in vm_device.h,
 * instantiated per rows-per-lane value by vm_r2.hip, vm_r4.hip and vm_r8.hip.
 */
#include "vm_device.h"

namespace blz {

#define BLZ_VM_SHAPES(X) X(8) X(4) X(2)
#define BLZ_VM_DECL(RR)                                                                     \
  hipError_t launch_vm_r##RR(int mode, int slow, int grid, int lds_bytes, hipStream_t st,   \
                             const VmBatch *bd);                                            \
  int vm_occupancy_r##RR(int mode, int slow, int lds_bytes);
BLZ_VM_SHAPES(BLZ_VM_DECL)
#undef BLZ_VM_DECL

// Exclusive prefix of every program's per-wave-slot kept counts (VM_PRED) in one
// pass, expanded into the row index of every kept row in output order: each
// workgroup scans a chunk of 512 counts of one program and combines its total
// with the chunks before it by a single-pass decoupled look-back over that
// program's state words. Chunks are taken in the order workgroups start, so
// every chunk a look-back waits for belongs to a workgroup that is running.
constexpr int kScanItems = 2;
static_assert(kScanItems * kBlock == kScanChunk, "scan chunk");
__global__ void __launch_bounds__(kBlock)
slot_scan_kernel(const VmBatch *__restrict__ bd_arg) {
  __shared__ uint64_t s_wave[kWaves];
  __shared__ uint64_t s_prefix;
  __shared__ int64_t s_chunk;
  __shared__ uint32_t s_pre[kScanChunk];  // exclusive prefix of each slot of the chunk
  const CONSTANT VmBatch &bd = *(const CONSTANT VmBatch *)bd_arg;
  const CONSTANT VmPlanRef *plans = (const CONSTANT VmPlanRef *)bd.plans;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (threadIdx.x == 0) s_chunk = (int64_t)atomicAdd(bd.chunk_counter, 1u);
  __syncthreads();
  const int64_t gch = s_chunk;
  if (gch >= bd.n_chunks) return;
  const int p = find_plan(bd.n_plans, gch, [&](int i) { return plans[i].chunk0; });
  const Workspace ws = load_ws(plans[p].ws);
  const CONSTANT ProgHeader &h = *(const CONSTANT ProgHeader *)plans[p].blob;
  const int64_t n_slots = ((h.n_rows + h.tile_rows - 1) / h.tile_rows) * kWaves;
  const int64_t b = gch - plans[p].chunk0;                 // chunk within the program
  const int64_t n_blocks = (n_slots + kScanChunk - 1) / kScanChunk;
  const int64_t base = b * kScanChunk + (int64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint64_t sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    v[i] = base + i < n_slots ? ws.wave_count[base + i] : 0u;
    sum += v[i];
  }
  uint64_t inc = sum;
  for (int off = 1; off < 64; off <<= 1) {
    const uint64_t y = __shfl_up(inc, off);
    if (lane >= off) inc += y;
  }
  if (lane == 63) s_wave[wave] = inc;
  __syncthreads();
  uint64_t wbase = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kWaves; w++) {
    const uint64_t c = s_wave[w];
    if (w < wave) wbase += c;
    total += c;
  }
  if (wave == 0) {
    const uint64_t pre = lookback_prefix(ws.tile_state, b, total, lane);
    if (lane == 0) {
      s_prefix = pre;
      if (b == n_blocks - 1) *ws.n_sel = pre + total;
    }
  }
  __syncthreads();
  // Slot prefixes of the chunk into LDS, then the row of every kept row of the
  // chunk's keep words, written in output order. The words go 64 at a time to a
  // wave (one per lane). Sparse words (at most 8 kept rows each): every lane
  // writes its word's rows; denser: for each non-zero word in turn, the lanes of
  // its set bits write their rows to consecutive outputs (coalesced).
  uint64_t run = s_prefix + wbase + inc - sum;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    s_pre[threadIdx.x * kScanItems + i] = (uint32_t)run;
    run += v[i];
  }
  __syncthreads();
  const int wps = h.tile_rows / 256;  // keep words per slot (a power of two <= 64)
  const int64_t ns = n_slots - b * kScanChunk < kScanChunk ? n_slots - b * kScanChunk : kScanChunk;
  const int64_t nw = ns * wps;              // keep words of the chunk
  const int64_t w0 = b * kScanChunk * wps;  // the chunk's first keep word in the program
  for (int64_t g0 = (int64_t)wave * 64; g0 < nw; g0 += kBlock) {
    const int64_t k = g0 + lane;
    const uint64_t word = k < nw ? ws.keep[w0 + k] : 0ull;
    const uint32_t c = (uint32_t)__builtin_popcountll(word);
    uint32_t incw = c;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(incw, off);
      if (lane >= off) incw += y;
    }
    const uint32_t excl = incw - c;
    const uint32_t gbase = s_pre[g0 / wps];  // 64 words start at a slot boundary
    uint32_t maxc = c;
    for (int off = 32; off > 0; off >>= 1) maxc = max(maxc, (uint32_t)__shfl_xor(maxc, off));
    if (maxc <= 8) {
      uint64_t wv = word;
      uint32_t o = gbase + excl;
      const uint32_t r0 = (uint32_t)((w0 + k) * 64);
      for (uint32_t it = 0; it < maxc; it++) {
        if (wv) {
          ws.idx[o++] = r0 + (uint32_t)__builtin_ctzll(wv);
          wv &= wv - 1;
        }
      }
      continue;
    }
    uint64_t nz = __builtin_amdgcn_ballot_w64(word != 0);
    while (nz) {
      const int j = __builtin_ctzll(nz);
      nz &= nz - 1;
      const uint64_t wj = readlane64(word, j);
      const uint32_t oj = gbase + (uint32_t)__builtin_amdgcn_readlane((int)excl, j);
      if ((wj >> lane) & 1)
        ws.idx[oj + __builtin_popcountll(wj & ((1ull << lane) - 1))] = (uint32_t)((w0 + g0 + j) * 64 + lane);
    }
  }
}

hipError_t launch_vm(int r, int mode, int slow, int grid, int lds_bytes, hipStream_t st, const VmBatch *bd) {
#define BLZ_VM_LAUNCH(RR) \
  if (r == RR) return launch_vm_r##RR(mode, slow, grid, lds_bytes, st, bd);
  BLZ_VM_SHAPES(BLZ_VM_LAUNCH)
#undef BLZ_VM_LAUNCH
  return hipErrorInvalidValue;
}

int vm_blocks_per_cu(int r, int mode, int slow, int lds_bytes) {
#define BLZ_VM_OCC(RR) \
  if (r == RR) return vm_occupancy_r##RR(mode, slow, lds_bytes);
  BLZ_VM_SHAPES(BLZ_VM_OCC)
#undef BLZ_VM_OCC
  return 0;
}

hipError_t launch_slot_scan(const VmBatch *bd, int64_t n_chunks, hipStream_t st) {
  hipLaunchKernelGGL(slot_scan_kernel, dim3((unsigned)(n_chunks > 0 ? n_chunks : 1)), dim3(kBlock), 0, st, bd);
  return hipGetLastError();
}

}  // namespace blz

/*
 * vm_kernel.hip — the tile VM's slot scan (the single-pass exclusive prefix of the
 * kept-row counts between the predicate and gather kernels) and the launch
 * dispatch over rows per lane. The VM kernels themselves are in vm_device.h,
 * instantiated per rows-per-lane value by vm_r2.hip, vm_r4.hip and vm_r8.hip.
 */
#include <type_traits>

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
// pass (written over the counts), expanded into the row index of every kept row in
// output order: each
// workgroup scans a chunk of counts of one program (scan_chunk_slots) and combines its total
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
  const int64_t cs = scan_chunk_slots(h.tile_rows);        // slots per chunk
  const int64_t n_blocks = (n_slots + cs - 1) / cs;
  const int64_t base = b * cs + (int64_t)threadIdx.x * kScanItems;
  const int64_t lim = n_slots < (b + 1) * cs ? n_slots : (b + 1) * cs;  // the chunk's end
  uint32_t v[kScanItems];
  uint64_t sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    v[i] = base + i < lim ? ws.wave_count[base + i] : 0u;
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
    // each slot's count is replaced by its exclusive prefix (the compiled gather
    // finds a kept row's place in its slot's staged values by it)
    if (base + i < lim) ws.wave_count[base + i] = (uint32_t)run;
    run += v[i];
  }
  __syncthreads();
  const int wps = h.tile_rows / 256;  // keep words per slot (a power of two <= 64)
  const int64_t ns = lim - b * cs;          // slots of the chunk
  const int64_t nw = ns * wps;              // keep words of the chunk
  const int64_t w0 = b * cs * wps;          // the chunk's first keep word in the program
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

// Predicate of one program whose ops are all compare-and / range leaves over
// columns (a conjunction of column-versus-constant comparisons, at most
// kDirectTerms ops; its validity, an AND of column bitmaps, is applied by the
// terms whose vcol names a column, so a row is kept only where each such column
// is valid): the operands are loaded from the columns straight into
// registers instead of being staged through LDS, every op's loads of a step in
// flight together and widened only after the last is issued; one instantiation
// per number of terms and pattern of 8-byte / 4-byte columns (WMASK bit k: term k
// reads 8 bytes), so no load sits behind a branch. The keep words and per-slot
// kept counts are the predicate kernel's (slot s = rows s T/4 .. (s + 1) T/4, keep
// word s GW + g), so the slot scan and gather follow unchanged. The comparisons
// are the tile VM's leaves.
constexpr int kDirectTerms = 3;
// Units per wave block: a wave takes this many consecutive units (100M rows of
// three columns: 0.259 / 0.255 / 0.251 / 0.262 ms at 1 / 2 / 4 / 8).
constexpr int kDirectBlock = 4;

// 64-row groups per wave step: 16 when the terms read exactly 8 bytes per row,
// else 8 (measured on 500M rows: one 4-byte column 0.39 ms at 8 against 0.42 at
// 16 and 0.69 at 32; one 8-byte column 0.78 ms at 16 against 0.79 at 8).
template <int NT, int WMASK>
constexpr int direct_step() {
  int bytes = 0;
  for (int k = 0; k < NT; k++) bytes += ((WMASK >> k) & 1) ? 8 : 4;
  return bytes == 8 ? 16 : 8;
}

// One term over the step's rows: the operands widened as the VM's operand class
// widens them, then the comparison ANDed into keep (one uniform branch per term,
// straight-line code per row).
template <int RS>
DEV void direct_term(int ho, int cls, const uint64_t (&raw)[RS], uint64_t im, uint64_t im2, int rf,
                     bool (&keep)[RS]) {
  uint64_t b[RS];
#pragma unroll
  for (int u = 0; u < RS; u++) {
    switch (cls) {
      case HC_L64: b[u] = raw[u]; break;
      case HC_L32I: b[u] = (uint64_t)(int64_t)(int32_t)(uint32_t)raw[u]; break;
      case HC_L32F: b[u] = d2b((double)__builtin_bit_cast(float, (uint32_t)raw[u])); break;
      case HC_L32I_F: b[u] = d2b((double)(int32_t)(uint32_t)raw[u]); break;
      default: b[u] = d2b((double)(int64_t)raw[u]); break;  // HC_L64_F
    }
  }
  const int64_t ii = (int64_t)im2;
  const double fi = b2d(im2), hf = b2d(im);
#define BLZ_TERM(CASE, TEST)                                                    \
  case CASE:                                                                    \
    _Pragma("unroll") for (int u = 0; u < RS; u++) keep[u] = keep[u] & (TEST); \
    break;
  switch (ho) {
    BLZ_TERM(HO_CMPAND + 0, (int64_t)b[u] < ii)
    BLZ_TERM(HO_CMPAND + 1, (int64_t)b[u] <= ii)
    BLZ_TERM(HO_CMPAND + 2, (int64_t)b[u] > ii)
    BLZ_TERM(HO_CMPAND + 3, (int64_t)b[u] >= ii)
    BLZ_TERM(HO_CMPAND + 4, b[u] == im2)
    BLZ_TERM(HO_CMPAND + 5, b[u] != im2)
    BLZ_TERM(HO_CMPAND + 6, b2d(b[u]) < fi)
    BLZ_TERM(HO_CMPAND + 7, b2d(b[u]) <= fi)
    BLZ_TERM(HO_CMPAND + 8, b2d(b[u]) > fi)
    BLZ_TERM(HO_CMPAND + 9, b2d(b[u]) >= fi)
    BLZ_TERM(HO_CMPAND + 10, b2d(b[u]) == fi)
    BLZ_TERM(HO_CMPAND + 11, b2d(b[u]) != fi)
    BLZ_TERM(HO_RANGE_I, ((int64_t)b[u] > ii) & ((int64_t)b[u] < (int64_t)im))
    BLZ_TERM(HO_RANGE_F, ((b2d(b[u]) > fi) | ((rf & 1) != 0 && b2d(b[u]) == fi)) &
                             ((b2d(b[u]) < hf) | ((rf & 2) != 0 && b2d(b[u]) == hf)))
    default: break;
  }
#undef BLZ_TERM
}

template <int NT, int WMASK>
__global__ void __launch_bounds__(kBlock) pred_direct_kernel(const VmBatch *__restrict__ bd_arg) {
  typedef uint32_t u32x8 __attribute__((ext_vector_type(8)));
  const CONSTANT VmBatch &bd = *(const CONSTANT VmBatch *)bd_arg;
  const CONSTANT VmPlanRef *plans = (const CONSTANT VmPlanRef *)bd.plans;
  const VmCtx cx = load_ctx(plans[0].blob);
  const Workspace ws = load_ws(plans[0].ws);
  constexpr int RS = direct_step<NT, WMASK>();
  const int lane = threadIdx.x & 63;
  const int64_t n = cx.n;
  const int GW = cx.T / 256;  // 64-row groups per slot
  const int64_t n_slots = ((n + cx.T - 1) / cx.T) * kWaves;
  // the terms, decoded once (scalar registers): column, operand class, leaf, bounds
  const CONSTANT u32x8 *opw = (const CONSTANT u32x8 *)cx.ops;
  const void *col[NT];
  const uint8_t *valid[NT];
  int cls[NT], ho[NT], rf[NT];
  uint64_t im[NT], im2[NT];
#pragma unroll
  for (int k = 0; k < NT; k++) {
    const u32x8 w = opw[k];
    cls[k] = (int)w.s0 % HC_COUNT_;
    ho[k] = (int)w.s0 / HC_COUNT_;
    rf[k] = (int)(w.s3 & 0xFF);
    col[k] = cx.in[(w.s3 >> 16) & 0xFF].data;
    const int vc = (int)((w.s3 >> 8) & 0xFF);
    valid[k] = vc != 0xFF ? (const uint8_t *)cx.in[vc].valid : nullptr;
    im[k] = (uint64_t)w.s4 | ((uint64_t)w.s5 << 32);
    im2[k] = (uint64_t)w.s6 | ((uint64_t)w.s7 << 32);
  }
  // Work unit: whole slots and whole steps (max(GW, RS) groups); a step takes RS
  // consecutive groups, which may span several slots (GW < 8) or be part of one.
  const int U = GW > RS ? GW : RS;
  const int64_t n_groups = n_slots * GW;
  const int64_t n_units = (n_groups + U - 1) / U;
  const int64_t n_waves = ((int64_t)gridDim.x * kBlock) >> 6;
  // the wave's index made visibly uniform, so step bases live in scalar registers
  const int64_t wave = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * kBlock + threadIdx.x) >> 6));
  // One unit's steps. CLAMP: the unit reaches past the column's last row (row
  // indices are clamped; those rows are not kept); every other unit loads from a
  // scalar base plus the lane offset and a constant per group. The two are separate
  // loops, so the fast one keeps its immediate offsets. Every load is non-temporal:
  // each row is read once (measured: 0.277 -> 0.259 ms on 100M rows of three
  // columns).
  auto run_unit = [&](auto clamp_tag, int64_t unit) {
    constexpr bool CLAMP = decltype(clamp_tag)::value;
    uint64_t slot_cnt = 0;  // GW >= RS: the unit's one slot
    for (int64_t g0 = unit * U; g0 < (unit + 1) * U && g0 < n_groups; g0 += RS) {
      uint64_t raw[NT][RS];
      if (!CLAMP) {
#pragma unroll
        for (int k = 0; k < NT; k++) {
          if ((WMASK >> k) & 1) {
            const GLOBAL uint64_t *p = (const GLOBAL uint64_t *)col[k] + g0 * 64;
#pragma unroll
            for (int u = 0; u < RS; u++) raw[k][u] = __builtin_nontemporal_load(p + u * 64 + lane);
          } else {
            const GLOBAL uint32_t *p = (const GLOBAL uint32_t *)col[k] + g0 * 64;
#pragma unroll
            for (int u = 0; u < RS; u++) raw[k][u] = __builtin_nontemporal_load(p + u * 64 + lane);
          }
        }
      } else {
#pragma unroll
        for (int k = 0; k < NT; k++) {
#pragma unroll
          for (int u = 0; u < RS; u++) {
            const int64_t r = (g0 + u) * 64 + lane;
            const int64_t rc = r < n ? r : n - 1;  // loads stay in bounds; the row is not kept
            raw[k][u] = ((WMASK >> k) & 1)
                            ? __builtin_nontemporal_load((const GLOBAL uint64_t *)col[k] + rc)
                            : (uint64_t)__builtin_nontemporal_load((const GLOBAL uint32_t *)col[k] + rc);
          }
        }
      }
      bool keep[RS];
#pragma unroll
      for (int u = 0; u < RS; u++) keep[u] = CLAMP ? (g0 + u) * 64 + lane < n : true;
#pragma unroll
      for (int k = 0; k < NT; k++) {
        if (valid[k] == nullptr) continue;  // uniform
#pragma unroll
        for (int u = 0; u < RS; u++) {
          const int64_t r = (g0 + u) * 64 + lane;
          keep[u] = keep[u] & valid_bit(valid[k], CLAMP && r >= n ? n - 1 : r);
        }
      }
#pragma unroll
      for (int k = 0; k < NT; k++) direct_term(ho[k], cls[k], raw[k], im[k], im2[k], rf[k], keep);
      uint64_t myword = 0;
      uint32_t c[RS];
#pragma unroll
      for (int u = 0; u < RS; u++) {
        const uint64_t km = __builtin_amdgcn_ballot_w64(keep[u]);
        if (lane == u) myword = km;
        c[u] = (uint32_t)__builtin_popcountll(km);
      }
      if (lane < RS && g0 + lane < n_groups) gstore<uint64_t>(ws.keep, g0 + lane, myword);
      if (GW >= RS) {
#pragma unroll
        for (int u = 0; u < RS; u++) slot_cnt += c[u];
      } else if (lane == 0) {  // RS / GW slots in this step
        for (int j = 0; j * GW < RS && g0 + j * GW < n_groups; j++) {
          uint32_t t = 0;
#pragma unroll
          for (int u = 0; u < RS; u++) t += (u / GW == j) ? c[u] : 0u;
          gstore<uint32_t>(ws.wave_count, g0 / GW + j, t);
        }
      }
    }
    if (GW >= RS && lane == 0) gstore<uint32_t>(ws.wave_count, unit * U / GW, (uint32_t)slot_cnt);
  };
  // Units in blocks of kDirectBlock consecutive units per wave (blocks strided over
  // the waves): a wave's loads and stores stay in consecutive lines.
  const int64_t n_blocks = (n_units + kDirectBlock - 1) / kDirectBlock;
  for (int64_t blk = wave; blk < n_blocks; blk += n_waves) {
    for (int64_t unit = blk * kDirectBlock; unit < (blk + 1) * kDirectBlock && unit < n_units; unit++) {
      if ((unit + 1) * U * 64 <= n) run_unit(std::false_type{}, unit);
      else run_unit(std::true_type{}, unit);
    }
  }
}

template <int NT, int WMASK>
static void launch_pred_direct_t(const VmBatch *bd, int grid, hipStream_t st) {
  hipLaunchKernelGGL((pred_direct_kernel<NT, WMASK>), dim3(grid), dim3(kBlock), 0, st, bd);
}
hipError_t launch_pred_direct(const VmBatch *bd, int n_terms, int wmask, int grid, hipStream_t st) {
  typedef void (*Fn)(const VmBatch *, int, hipStream_t);
  static const Fn t1[2] = {launch_pred_direct_t<1, 0>, launch_pred_direct_t<1, 1>};
  static const Fn t2[4] = {launch_pred_direct_t<2, 0>, launch_pred_direct_t<2, 1>, launch_pred_direct_t<2, 2>,
                           launch_pred_direct_t<2, 3>};
  static const Fn t3[8] = {launch_pred_direct_t<3, 0>, launch_pred_direct_t<3, 1>, launch_pred_direct_t<3, 2>,
                           launch_pred_direct_t<3, 3>, launch_pred_direct_t<3, 4>, launch_pred_direct_t<3, 5>,
                           launch_pred_direct_t<3, 6>, launch_pred_direct_t<3, 7>};
  if (n_terms < 1 || n_terms > kDirectTerms || wmask < 0 || wmask >= (1 << n_terms)) return hipErrorInvalidValue;
  (n_terms == 1 ? t1 : n_terms == 2 ? t2 : t3)[wmask](bd, grid, st);
  return hipGetLastError();
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

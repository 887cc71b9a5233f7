/*
 * kernels.hip — gfx950 kernels of the expression / filter / gather path.
 *
 * interp_kernel<S, R, COMPACT>
 *   The register-file interpreter of src/Interpreter (transformKernel +
 *   InterpreterFunctor, interpreter_ops.cuh:1272-1458), redesigned for CDNA4:
 *   - 256-thread workgroups = 4 wave64s; a wave owns R groups of 64 consecutive
 *     rows, so each group's validity is exactly one 64-bit word and group u of
 *     lane l is row  tile*256R + wave*64R + u*64 + l;
 *   - the register file lives in VGPRs (S slots x R rows x 8 bytes), addressed
 *     with a wave-uniform slot index (GPR-indexing mode), not in LDS as the
 *     reference's `smem[slot*blockDim + tid]` (interpreter_ops.cuh:266);
 *   - the program (ops, immediates, column descriptors) is staged in LDS once per
 *     workgroup and read with uniform addresses into SGPRs; ops are fully typed on
 *     the host, so dispatch is one uniform switch per op, amortised over 64R rows;
 *   - every input column of a tile is loaded (raw, coalesced) before any is
 *     widened, so all loads of a tile are in flight together;
 *   - validity is computed in the same pass on 64-row words, one word per lane;
 *   - COMPACT=true fuses the filter: predicate -> keep mask (ballot) -> workgroup
 *     scan -> single-pass decoupled look-back across tiles -> in-order scatter of
 *     the projected / gathered outputs and their validity bits. This replaces the
 *     reference's stencil + gdf_sequence + apply_boolean_mask + per-column
 *     materialize_column (CalciteInterpreter.cpp:509-573).
 *
 * gather_kernel
 *   materialize_column (ColumnManipulation.cu:83-150): out[i] = in[idx[i]] and the
 *   validity-bit gather (gather_bits, :26-68) in one pass, 64 indices per wave so
 *   each output validity word is one ballot.
 *
 * All floating-point code is compiled with -ffp-contract=off.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <utility>

#include "program.h"
#include "kernels.h"
#include "device_common.h"

namespace blz {

// Runs `nops` device ops starting at `ops` over the register file.
template <int S, int R>
DEV void run_ops(const CONSTANT DevOp *ops, int nops, uint64_t (&rf)[S * R], const int64_t (&row)[R],
                 const CONSTANT InCol *in) {
  for (int k = 0; k < nops; k++) {
    // code | dst | a | b | flags | conv | vcol | pad: one scalar load
    const uint64_t ctl = *(const CONSTANT uint64_t *)(ops + k);
    const int code = (int)(ctl & 0xFF), dst = (int)((ctl >> 8) & 0xFF);
    const int fa = (int)((ctl >> 32) & 0xFF), cv = (int)((ctl >> 40) & 0xFF);
    const int vcol = (int)((ctl >> 48) & 0xFF);
    uint64_t A[R], B[R], Z[R];
    bool vb[R];
    if (fa & F_A_IMM) {
      const uint64_t ia = ops[k].imm_a;
#pragma unroll
      for (int u = 0; u < R; u++) A[u] = ia;
    } else {
      const int a = (int)((ctl >> 16) & 0xFF);
#pragma unroll
      for (int u = 0; u < R; u++) A[u] = rf[a * R + u];
    }
    if (fa & F_B_IMM) {
      const uint64_t ib = ops[k].imm_b;
#pragma unroll
      for (int u = 0; u < R; u++) B[u] = ib;
    } else {
      const int b = (int)((ctl >> 24) & 0xFF);
#pragma unroll
      for (int u = 0; u < R; u++) B[u] = rf[b * R + u];
    }
    const int ca = cv & 3, cb = (cv >> 2) & 3, co = (cv >> 4) & 3;
    if (ca) {
#pragma nounroll
      for (int u = 0; u < R; u++) A[u] = conv_slow(A[u], ca);
    }
    if (cb) {
#pragma nounroll
      for (int u = 0; u < R; u++) B[u] = conv_slow(B[u], cb);
    }
    if (vcol != 0xFF) {
      const void *vp = in[vcol].valid;
#pragma unroll
      for (int u = 0; u < R; u++) vb[u] = valid_bit(vp, row[u]);
    } else {
#pragma unroll
      for (int u = 0; u < R; u++) vb[u] = true;
    }
    exec_op<R>(code, A, B, vb, Z);
    if (co) {
#pragma nounroll
      for (int u = 0; u < R; u++) Z[u] = conv_slow(Z[u], co);
    }
#pragma unroll
    for (int u = 0; u < R; u++) rf[dst * R + u] = Z[u];
  }
}

template <int S, int R, bool COMPACT>
__global__ void __launch_bounds__(kBlock)
interp_kernel(const uint8_t *__restrict__ blob_arg, Workspace ws) {
  __shared__ int64_t s_tile;
  __shared__ uint64_t s_wave_cnt[kWaves];
  __shared__ uint64_t s_prefix;
  __shared__ uint64_t s_vbits[kWaves * R + 2];

  const CONSTANT uint8_t *blob = (const CONSTANT uint8_t *)blob_arg;
  const CONSTANT ProgHeader &h = *(const CONSTANT ProgHeader *)blob;
  const CONSTANT InCol *in = (const CONSTANT InCol *)(blob + h.off_in);
  const CONSTANT OutCol *out = (const CONSTANT OutCol *)(blob + h.off_out);
  const CONSTANT DevOp *ops = (const CONSTANT DevOp *)(blob + h.off_ops);
  const CONSTANT VOp *vops = (const CONSTANT VOp *)(blob + h.off_vops);
  const int64_t n = h.n_rows;
  const int n_in = h.n_in, n_out = h.n_out;
  const int n_ops_p = h.n_ops_p, n_ops_q = h.n_ops_q;
  const int n_vops_p = h.n_vops_p, n_vops_q = h.n_vops_q;
  const int any_valid_out = h.any_valid_out;

  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t rows_per_tile = (int64_t)kBlock * R;
  const int64_t n_tiles = (n + rows_per_tile - 1) / rows_per_tile;
  constexpr int MI = S < kMaxUnrolledInputs ? S : kMaxUnrolledInputs;

  uint64_t rf[S * R];
  uint64_t vrf[S];
#pragma unroll
  for (int i = 0; i < S * R; i++) rf[i] = 0;
#pragma unroll
  for (int i = 0; i < S; i++) vrf[i] = 0;
  // lane o: nulls of output o (< 64) counted by this wave over its tiles, added once
  // at the end (atomics on one address serialize: one per tile would cost more than
  // the tile)
  unsigned long long nacc = 0;

  int64_t tile = blockIdx.x;
  if (COMPACT) {
    if (threadIdx.x == 0) s_tile = (int64_t)atomicAdd(ws.tile_counter, 1u);
    __syncthreads();
    tile = s_tile;
  }
  while (tile < n_tiles) {
    const int64_t wave_base = tile * rows_per_tile + (int64_t)wave * 64 * R;
    int64_t row[R], lrow[R];
#pragma unroll
    for (int u = 0; u < R; u++) {
      row[u] = wave_base + u * 64 + lane;
      lrow[u] = row[u] < n ? row[u] : n - 1;  // clamped: loads stay in bounds
    }

    // 1. Every input column's raw loads are issued before any is widened: wide
    //    columns land in w64, narrow ones in w32 (distinct registers, so no move
    //    and no wait sits between the loads).
    {
      uint64_t w64[MI][R];
      uint32_t w32[MI][R];
#pragma unroll
      for (int c = 0; c < MI; c++) {
        if (c < n_in) {
          const int kind = in[c].load;
          const void *base = in[c].data;
          if (is_wide(kind)) {
#pragma unroll
            for (int u = 0; u < R; u++) w64[c][u] = gload<uint64_t>(base, lrow[u]);
          } else {
#pragma unroll
            for (int u = 0; u < R; u++) w32[c][u] = load_narrow(kind, base, lrow[u]);
          }
        }
      }
#pragma unroll
      for (int c = 0; c < MI; c++) {
        if (c < n_in) {
          const int kind = in[c].load;
#pragma unroll
          for (int u = 0; u < R; u++) rf[c * R + u] = is_wide(kind) ? w64[c][u] : widen32(kind, w32[c][u]);
        }
      }
    }
    for (int c = kMaxUnrolledInputs; c < n_in; c++) {  // plans with more than 16 inputs
      const int kind = in[c].load;
      const void *base = in[c].data;
#pragma unroll
      for (int u = 0; u < R; u++)
        rf[c * R + u] = is_wide(kind) ? gload<uint64_t>(base, lrow[u]) : widen32(kind, load_narrow(kind, base, lrow[u]));
    }
    // Validity words of the wave's R groups: lane u < R holds the word of group u.
    const int64_t wave_word = wave_base >> 6;
    if (any_valid_out || COMPACT) {
      const int64_t w = wave_word + (lane < R ? lane : 0);
      for (int c = 0; c < n_in; c++) {
        const int vs = in[c].vslot;
        if (vs >= 0) {
          const void *vp = in[c].valid;
          vrf[vs] = (lane < R && w * 64 < n) ? load_vword(vp, w, n) : 0ull;
        }
      }
    }

    if (!COMPACT) {
      // 2. Projection program, then narrowing stores of every output.
      run_ops<S, R>(ops, n_ops_q, rf, lrow, in);
      if (any_valid_out) run_vops<S>(vops, n_vops_q, vrf);
      for (int o = 0; o < n_out; o++) {
        const int kind = out[o].store;
        const int slot = out[o].slot;
        void *base = out[o].data;
        if (out[o].src_kind == 1) {  // a bare column (its slot is the input's)
#pragma unroll
          for (int u = 0; u < R; u++)
            if (row[u] < n) copy_alias(kind, base, row[u], in[slot].data, row[u]);
        } else {
#pragma unroll
          for (int u = 0; u < R; u++)
            if (row[u] < n) store_narrow(kind, base, row[u], rf[slot * R + u]);
        }
        void *vp = out[o].valid;
        if (vp != nullptr) {
          const uint64_t word = vsrc_word<S>(out[o].vsrc, vrf);
          const int64_t w = wave_word + lane;
          uint64_t nulls = 0;
          if (lane < R && w * 64 < n) {
            store_vword(vp, w, n, word);
            nulls = __builtin_popcountll(~word & low_mask(n - w * 64));
          }
          for (int off = 32; off > 0; off >>= 1) nulls += __shfl_xor(nulls, off);
          if (o < 64) nacc += lane == o ? nulls : 0;
          else if (lane == 0 && nulls) atomicAdd(ws.null_counts + o, (unsigned long long)nulls);
        }
      }
      tile += gridDim.x;
      continue;
    }

    // ---- COMPACT: filter + project/gather in one pass ----
    run_ops<S, R>(ops, n_ops_p, rf, lrow, in);
    run_vops<S>(vops, n_vops_p, vrf);
    const int st_slot = h.stencil_slot, st_width = h.stencil_width;
    const int st_vsrc = h.stencil_vsrc;
    const uint64_t st_mask = low_mask(8 * st_width);
    bool keep[R];
    uint64_t kmask[R];
    uint64_t wave_cnt = 0;
    {
      // Predicate validity: a validity slot or a constant.
      const uint64_t vstencil = vsrc_word<S>(st_vsrc, vrf);
#pragma unroll
      for (int u = 0; u < R; u++) {
        const bool vbit = (readlane64(vstencil, u) >> lane) & 1;
        keep[u] = row[u] < n && ((rf[st_slot * R + u] & st_mask) != 0) && vbit;
        kmask[u] = __builtin_amdgcn_ballot_w64(keep[u]);
        wave_cnt += __builtin_popcountll(kmask[u]);
      }
    }
    if (lane == 0) s_wave_cnt[wave] = wave_cnt;
    __syncthreads();
    uint64_t tile_cnt = 0, wave_prefix = 0;
#pragma unroll
    for (int w = 0; w < kWaves; w++) {
      const uint64_t c = s_wave_cnt[w];
      if (w < wave) wave_prefix += c;
      tile_cnt += c;
    }
    // Single-pass decoupled look-back (wave 0). State word: bits 62-63 flag
    // (0 = not ready, 1 = aggregate, 2 = inclusive prefix), bits 0-61 value.
    if (wave == 0) {
      constexpr uint64_t FA = 1ull << 62, FP = 2ull << 62, VM = (1ull << 62) - 1;
      uint64_t prefix = 0;
      if (tile == 0) {
        if (lane == 0)
          __hip_atomic_store(ws.tile_state, FP | tile_cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        if (lane == 0)
          __hip_atomic_store(ws.tile_state + tile, FA | tile_cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int64_t look = tile - 1;
        while (true) {
          const int64_t idx = look - lane;
          uint64_t s = FP;  // virtual inclusive prefix 0 before tile 0
          if (idx >= 0)
            s = __hip_atomic_load(ws.tile_state + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          const uint64_t flag = s >> 62;
          const uint64_t is_x = __builtin_amdgcn_ballot_w64(flag == 0);
          const uint64_t is_p = __builtin_amdgcn_ballot_w64(flag == 2);
          const uint64_t upto = is_p ? (((is_p & (~is_p + 1)) << 1) - 1) : ~0ull;
          if (is_x & upto) {
            __builtin_amdgcn_s_sleep(1);
            continue;
          }
          uint64_t v = ((upto >> lane) & 1) ? (s & VM) : 0;
          for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
          prefix += v;
          if (is_p) break;
          look -= 64;
        }
        if (lane == 0)
          __hip_atomic_store(ws.tile_state + tile, FP | (prefix + tile_cnt), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
      }
      if (lane == 0) {
        s_prefix = prefix;
        if (tile == n_tiles - 1) *ws.n_sel = prefix + tile_cnt;
      }
    }
    __syncthreads();
    const uint64_t tile_prefix = s_prefix;

    if (tile_cnt > 0) {
      // Output positions of the kept rows, in row order.
      uint64_t pos[R];
      uint64_t run = tile_prefix + wave_prefix;
#pragma unroll
      for (int u = 0; u < R; u++) {
        const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(kmask[u] >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)kmask[u], 0u));
        pos[u] = run + below;
        run += __builtin_popcountll(kmask[u]);
      }
      // Projection program only if some row of this wave survives.
      if (wave_cnt > 0) run_ops<S, R>(ops + n_ops_p, n_ops_q, rf, lrow, in);
      if (any_valid_out) run_vops<S>(vops + n_vops_p, n_vops_q, vrf);
      for (int o = 0; o < n_out; o++) {
        const int kind = out[o].store;
        const int slot = out[o].slot;
        void *base = out[o].data;
        if (wave_cnt > 0 && out[o].src_kind == 1) {  // a bare column (its slot is the input's)
#pragma unroll
          for (int u = 0; u < R; u++)
            if (keep[u]) copy_alias(kind, base, (int64_t)pos[u], in[slot].data, row[u]);
        } else if (wave_cnt > 0) {
#pragma unroll
          for (int u = 0; u < R; u++)
            if (keep[u]) store_narrow(kind, base, (int64_t)pos[u], rf[slot * R + u]);
        }
        void *vp = out[o].valid;
        if (vp != nullptr) {
          // Compacted validity bits: tile-local bit buffer in LDS, then words.
          const uint64_t vword = vsrc_word<S>(out[o].vsrc, vrf);
          const int nw = (int)(((tile_prefix & 63) + tile_cnt + 63) >> 6);
          for (int i = threadIdx.x; i < nw; i += kBlock) s_vbits[i] = 0;
          __syncthreads();
          uint64_t nulls = 0;
#pragma unroll
          for (int u = 0; u < R; u++) {
            const bool vbit = (readlane64(vword, u) >> lane) & 1;
            if (keep[u]) {
              if (vbit) {
                const uint64_t local = (tile_prefix & 63) + (pos[u] - tile_prefix);
                atomicOr((unsigned long long *)&s_vbits[local >> 6], 1ull << (local & 63));
              } else {
                nulls++;
              }
            }
          }
          __syncthreads();
          const int64_t w0 = (int64_t)(tile_prefix >> 6);
          const uint64_t end = tile_prefix + tile_cnt;  // exclusive bit end
          for (int i = threadIdx.x; i < nw; i += kBlock) {
            const int64_t w = w0 + i;
            const bool first_shared = (i == 0) && (tile_prefix & 63);
            const bool last_shared = (i == nw - 1) && (end & 63);
            unsigned long long *gw = (unsigned long long *)((uint8_t *)vp + 8 * w);
            if (first_shared || last_shared) atomicOr(gw, (unsigned long long)s_vbits[i]);
            else gstore<uint64_t>(gw, 0, s_vbits[i]);
          }
          for (int off = 32; off > 0; off >>= 1) nulls += __shfl_xor(nulls, off);
          if (o < 64) nacc += lane == o ? nulls : 0;
          else if (lane == 0 && nulls) atomicAdd(ws.null_counts + o, (unsigned long long)nulls);
          __syncthreads();
        }
      }
    }
    // Next tile.
    __syncthreads();
    if (threadIdx.x == 0) s_tile = (int64_t)atomicAdd(ws.tile_counter, 1u);
    __syncthreads();
    tile = s_tile;
  }
  if (nacc) atomicAdd(ws.null_counts + lane, nacc);
}

// Adds each wave's `v` (lane 0's) to *dst, summed over the workgroup first: one
// atomic per workgroup (atomics on one address serialize). Every thread of the
// workgroup calls it.
DEV void block_add(unsigned long long *dst, unsigned long long v) {
  __shared__ unsigned long long s_part[kWaves];
  if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kWaves; w++) t += s_part[w];
    if (t) atomicAdd(dst, t);
  }
}

// ---------------------------------------------------------------------------
// gather: out[i] = in[idx[i]] (idx >= 0), validity bits gathered, null count. A
// wave takes four 64-index words at a time and issues every index load, then every
// value load, before its stores: random accesses need many requests in flight.
constexpr int kGatherWords = 4;
template <typename T, typename I>
__global__ void __launch_bounds__(kBlock)
gather_kernel(const T *__restrict__ in, const uint8_t *__restrict__ in_valid, int64_t n_in,
              const I *__restrict__ idx, int64_t n_idx, T *__restrict__ out,
              uint8_t *__restrict__ out_valid, unsigned long long *null_count) {
  const int lane = threadIdx.x & 63;
  const int64_t n_words = (n_idx + 63) >> 6;
  const int64_t wave_id = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t n_waves = ((int64_t)gridDim.x * kBlock) >> 6;
  uint64_t nulls = 0;
  for (int64_t w0 = wave_id * kGatherWords; w0 < n_words; w0 += n_waves * kGatherWords) {
    int64_t j[kGatherWords];
    bool ok[kGatherWords];
    T v[kGatherWords];
#pragma unroll
    for (int u = 0; u < kGatherWords; u++) {
      const int64_t i = (w0 + u) * 64 + lane;
      j[u] = i < n_idx ? (int64_t)idx[i] : -1;
    }
#pragma unroll
    for (int u = 0; u < kGatherWords; u++) {
      ok[u] = j[u] >= 0 && j[u] < n_in;  // an index past the input is treated like a negative one
      v[u] = ok[u] ? in[j[u]] : T(0);
    }
#pragma unroll
    for (int u = 0; u < kGatherWords; u++) {
      const int64_t w = w0 + u;
      if (w >= n_words) break;  // wave-uniform
      const int64_t i = w * 64 + lane;
      if (ok[u]) __builtin_nontemporal_store(v[u], out + i);  // written once
      if (out_valid != nullptr) {
        const uint64_t word = __builtin_amdgcn_ballot_w64(ok[u] && valid_bit(in_valid, j[u]));
        if (lane == 0) {
          store_vword(out_valid, w, n_idx, word);
          nulls += (uint64_t)__builtin_popcountll(~word & low_mask(n_idx - w * 64));
        }
      }
    }
  }
  if (out_valid != nullptr) block_add(null_count, lane == 0 ? nulls : 0);  // out_valid: uniform
}

// idx[i] = init + i * step (gdf_sequence, Utils.cu:4-12)
__global__ void __launch_bounds__(kBlock)
sequence_kernel(int32_t *out, int64_t n, int32_t init, int32_t step) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    out[i] = init + (int32_t)i * step;
}

// Zero-fill (graph-capturable replacement for hipMemsetAsync).
__global__ void __launch_bounds__(kBlock)
fill_zero_kernel(uint64_t *p, int64_t n_words) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n_words; i += (int64_t)gridDim.x * kBlock)
    p[i] = 0;
}

// Seeded synthetic column (benchmark inputs generated in place, not copied from
// the host): row i's value and validity come from splitmix64 of (seed, i).
// dist: 0 integer uniform [a, b]; 1 real uniform [a, b); 2 normal(a, b);
// 3 lognormal(a, b); 4 k / 100 for k integer uniform [a, b]; 5 a + i.
DEV uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
DEV double unit(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

__global__ void __launch_bounds__(kBlock)
synth_kernel(void *data, uint8_t *valid, int64_t n, int dtype, uint64_t seed, int dist, double a, double b,
             double null_rate) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const uint64_t h1 = mix64(seed * 0x100000001B3ull + (uint64_t)i);
    const uint64_t h2 = mix64(h1 ^ 0xD6E8FEB86659FD93ull);
    double v;
    int64_t iv;
    bool is_int = false;
    switch (dist) {
      case 0: iv = (int64_t)a + (int64_t)(h1 % (uint64_t)((int64_t)b - (int64_t)a + 1)); is_int = true; v = 0; break;
      case 1: v = a + (b - a) * unit(h1); break;
      case 2: v = a + b * sqrt(-2.0 * log(1.0 - unit(h1))) * cos(6.283185307179586 * unit(h2)); break;
      case 3: v = exp(a + b * sqrt(-2.0 * log(1.0 - unit(h1))) * cos(6.283185307179586 * unit(h2))); break;
      case 4: v = (double)((int64_t)a + (int64_t)(h1 % (uint64_t)((int64_t)b - (int64_t)a + 1))) / 100.0; break;
      default: iv = (int64_t)a + i; is_int = true; v = 0; break;
    }
    switch (dtype) {
      case 1: ((int8_t *)data)[i] = (int8_t)(is_int ? iv : (int64_t)v); break;
      case 2: ((int16_t *)data)[i] = (int16_t)(is_int ? iv : (int64_t)v); break;
      case 3: ((int32_t *)data)[i] = (int32_t)(is_int ? iv : (int64_t)v); break;
      case 4: ((int64_t *)data)[i] = is_int ? iv : (int64_t)v; break;
      case 5: ((float *)data)[i] = (float)(is_int ? (double)iv : v); break;
      case 6: ((double *)data)[i] = is_int ? (double)iv : v; break;
      default: break;
    }
  }
  // validity words: one lane per 64 rows
  if (valid) {
    const int64_t nw = (n + 63) / 64;
    for (int64_t w = (int64_t)blockIdx.x * kBlock + threadIdx.x; w < nw; w += (int64_t)gridDim.x * kBlock) {
      uint64_t word = 0;
      for (int k = 0; k < 64; k++) {
        const int64_t i = w * 64 + k;
        if (i >= n) break;
        const uint64_t h3 = mix64((seed ^ 0xA5A5A5A55A5A5A5Aull) * 0x100000001B3ull + (uint64_t)i);
        if (unit(h3) >= null_rate) word |= 1ull << k;
      }
      store_vword(valid, w, n, word);
    }
  }
}

// Zeroes several byte ranges (8-byte multiples) in one launch.
__global__ void __launch_bounds__(kBlock)
fill_ranges_kernel(void *const *dst, const int64_t *bytes, int n) {
  for (int r = 0; r < n; r++) {
    uint64_t *p = (uint64_t *)dst[r];
    const int64_t w = bytes[r] / 8;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < w; i += (int64_t)gridDim.x * kBlock) p[i] = 0;
  }
}

// Plain 16-byte-wide copy (also the reference point for HBM byte counters).
__global__ void __launch_bounds__(kBlock)
copy_kernel(uint4 *__restrict__ dst, const uint4 *__restrict__ src, int64_t n16) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n16; i += (int64_t)gridDim.x * kBlock)
    dst[i] = src[i];
}

// popcount of validity bits < n -> null count (gdf_count_nonzero_mask's role).
__global__ void __launch_bounds__(kBlock)
null_count_kernel(const uint8_t *valid, int64_t n, unsigned long long *nulls_out) {
  const int64_t n_words = (n + 63) >> 6;
  uint64_t nulls = 0;
  for (int64_t w = (int64_t)blockIdx.x * kBlock + threadIdx.x; w < n_words; w += (int64_t)gridDim.x * kBlock) {
    const uint64_t word = load_vword(valid, w, n);
    nulls += __builtin_popcountll(~word & low_mask(n - w * 64));
  }
  for (int off = 32; off > 0; off >>= 1) nulls += __shfl_xor(nulls, off);
  block_add(nulls_out, nulls);
}

// ---------------------------------------------------------------------------
// Reduction of one column over its valid rows, cudf::reduce's role in an
// aggregation without GROUP BY (GroupBy.cpp:230-290): op 0 SUM, 1 MIN, 2 MAX.
// Values are widened as the interpreter widens them (integers to int64, floats to
// binary64); SUM accumulates in int64 (wrapping) or binary64, MIN / MAX compare in
// the widened type (a NaN never replaces the running value).
//
// A step is 256 lanes x kRedLoads loads of 16 bytes (E = 16 / width rows per
// load): in load u lane t holds rows step*S + u*256*E + t*E + [0, E), S = the
// step's rows. Workgroup b takes steps b, b + grid, ... in that order and combines
// its rows in the order above; the rows past the last whole step go to workgroup
// (whole steps) % grid, lane t taking rows t, t + 256, ... of them; then a fixed
// tree over the lanes and waves, and the final pass combines the partials in
// workgroup order. A float sum is therefore the same on every run with the same
// row count (the grid is a function of the row count and the dtype).

template <int OP, bool FLT>
DEV uint64_t red_identity() {
  if (FLT) return d2b(OP == 0 ? 0.0 : OP == 1 ? __builtin_inf() : -__builtin_inf());
  return OP == 0 ? 0ull : OP == 1 ? (uint64_t)INT64_MAX : (uint64_t)INT64_MIN;
}
template <int OP, bool FLT>
DEV uint64_t red_combine(uint64_t a, uint64_t b) {
  if (FLT) {
    const double x = b2d(a), y = b2d(b);
    return d2b(OP == 0 ? x + y : OP == 1 ? (y < x ? y : x) : (y > x ? y : x));
  }
  const int64_t x = (int64_t)a, y = (int64_t)b;
  return OP == 0 ? a + b : (uint64_t)(OP == 1 ? (y < x ? y : x) : (y > x ? y : x));
}
// value and valid-row count of this workgroup, combined over its waves in order
template <int OP, bool FLT>
DEV void red_block(uint64_t acc, unsigned long long cnt, RedPart *dst) {
  __shared__ uint64_t s_v[kWaves];
  __shared__ unsigned long long s_c[kWaves];
  for (int off = 32; off > 0; off >>= 1) {
    acc = red_combine<OP, FLT>(acc, __shfl_xor(acc, off));
    cnt += __shfl_xor(cnt, off);
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_v[wave] = acc;
    s_c[wave] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t v = s_v[0];
    unsigned long long c = s_c[0];
    for (int w = 1; w < kWaves; w++) {
      v = red_combine<OP, FLT>(v, s_v[w]);
      c += s_c[w];
    }
    dst->v = v;
    dst->cnt = c;
  }
}

template <int LD>
constexpr int red_width() {
  return LD == LD_I64 || LD == LD_F64 ? 8 : LD == LD_I32 || LD == LD_F32 ? 4 : LD == LD_I16 ? 2 : 1;
}
// Element e (of E) of a 16-byte load, widened.
template <int LD>
DEV uint64_t red_elem(const u32x4 &v, int e) {
  constexpr int W = red_width<LD>();
  if constexpr (W == 8) return (uint64_t)v[2 * e] | ((uint64_t)v[2 * e + 1] << 32);
  else if constexpr (W == 4) return widen32(LD, v[e]);
  else if constexpr (W == 2) return widen32(LD, (v[e >> 1] >> (16 * (e & 1))) & 0xffffu);
  else return widen32(LD, (v[e >> 2] >> (8 * (e & 3))) & 0xffu);
}
// The 16 bytes from row `row` (a multiple of E): one load when the column is
// 16-byte aligned, else element loads packed the same way.
template <int LD, bool VEC>
DEV u32x4 red_load(const void *data, int64_t row) {
  constexpr int W = red_width<LD>();
  const GLOBAL uint8_t *p = (const GLOBAL uint8_t *)data + row * W;
  if constexpr (VEC) {
    return __builtin_nontemporal_load((const GLOBAL u32x4 *)p);
  } else {
    u32x4 v = {0u, 0u, 0u, 0u};
    if constexpr (W == 8) {
      for (int e = 0; e < 2; e++) {
        const uint64_t x = ((const GLOBAL uint64_t *)p)[e];
        v[2 * e] = (uint32_t)x;
        v[2 * e + 1] = (uint32_t)(x >> 32);
      }
    } else if constexpr (W == 4) {
      for (int e = 0; e < 4; e++) v[e] = ((const GLOBAL uint32_t *)p)[e];
    } else if constexpr (W == 2) {
      for (int e = 0; e < 8; e++) v[e >> 1] |= (uint32_t)((const GLOBAL uint16_t *)p)[e] << (16 * (e & 1));
    } else {
      for (int e = 0; e < 16; e++) v[e >> 2] |= (uint32_t)p[e] << (8 * (e & 3));
    }
    return v;
  }
}
// Validity bits of rows row .. row + E - 1 (row a multiple of E).
template <int E>
DEV uint32_t red_valid(const uint8_t *valid, int64_t row) {
  const GLOBAL uint8_t *v = (const GLOBAL uint8_t *)valid;
  if constexpr (E == 16) return (uint32_t)v[row >> 3] | ((uint32_t)v[(row >> 3) + 1] << 8);
  else if constexpr (E == 8) return v[row >> 3];
  else return (v[row >> 3] >> (row & 7)) & ((1u << E) - 1);
}

template <int OP, bool FLT, int LD, bool VEC>
__global__ void __launch_bounds__(kBlock)
reduce_kernel(const void *data, const uint8_t *valid, int64_t n, RedPart *parts) {
  constexpr int W = red_width<LD>(), E = 16 / W;
  constexpr int64_t S = (int64_t)kBlock * E * kRedLoads;
  const int t = threadIdx.x;
  const int64_t whole = n / S;
  uint64_t acc = red_identity<OP, FLT>();
  unsigned long long cnt = 0;
  for (int64_t s = blockIdx.x; s < whole; s += gridDim.x) {
    const int64_t r0 = s * S + (int64_t)t * E;
    u32x4 v[kRedLoads];
    uint32_t vb[kRedLoads];
#pragma unroll
    for (int u = 0; u < kRedLoads; u++) v[u] = red_load<LD, VEC>(data, r0 + (int64_t)u * kBlock * E);
#pragma unroll
    for (int u = 0; u < kRedLoads; u++)
      vb[u] = valid ? red_valid<E>(valid, r0 + (int64_t)u * kBlock * E) : (1u << E) - 1;
#pragma unroll
    for (int u = 0; u < kRedLoads; u++) {
      cnt += __builtin_popcount(vb[u]);
#pragma unroll
      for (int e = 0; e < E; e++) {
        const uint64_t x = red_combine<OP, FLT>(acc, red_elem<LD>(v[u], e));
        acc = (vb[u] >> e) & 1 ? x : acc;
      }
    }
  }
  if (blockIdx.x == whole % gridDim.x) {
    for (int64_t i = whole * S + t; i < n; i += kBlock) {
      if (!valid_bit(valid, i)) continue;
      const uint64_t x = W == 8 ? gload<uint64_t>(data, i) : widen32(LD, load_narrow(LD, data, i));
      acc = red_combine<OP, FLT>(acc, x);
      cnt++;
    }
  }
  red_block<OP, FLT>(acc, cnt, parts + blockIdx.x);
}
template <int OP, bool FLT>
__global__ void __launch_bounds__(kBlock) reduce_final_kernel(const RedPart *parts, int n_parts, RedPart *out) {
  uint64_t acc = red_identity<OP, FLT>();
  unsigned long long cnt = 0;
  for (int i = threadIdx.x; i < n_parts; i += kBlock) {
    acc = red_combine<OP, FLT>(acc, parts[i].v);
    cnt += parts[i].cnt;
  }
  red_block<OP, FLT>(acc, cnt, out);
}

// ---------------------------------------------------------------------------
// Host-side launchers.

template <int S, int R>
static hipError_t launch_sr(bool compact, int grid, hipStream_t st, const uint8_t *blob, Workspace ws) {
  if (compact)
    hipLaunchKernelGGL((interp_kernel<S, R, true>), dim3(grid), dim3(kBlock), 0, st, blob, ws);
  else
    hipLaunchKernelGGL((interp_kernel<S, R, false>), dim3(grid), dim3(kBlock), 0, st, blob, ws);
  return hipGetLastError();
}

// Register-file shapes: S slots x R rows per lane. S*R <= 16 keeps the file in
// VGPRs; S = 32 / 64 (R = 1) serve large plans with the file in scratch.
hipError_t launch_interp(int s, int r, bool compact, int grid, int lds_bytes, hipStream_t st,
                         const uint8_t *blob, Workspace ws) {
  if (lds_bytes > 0) return hipErrorInvalidValue;
  if (s == 8 && r == 1) return launch_sr<8, 1>(compact, grid, st, blob, ws);
  if (s == 8 && r == 2) return launch_sr<8, 2>(compact, grid, st, blob, ws);
  if (s == 16 && r == 1) return launch_sr<16, 1>(compact, grid, st, blob, ws);
  if (s == 32 && r == 1) return launch_sr<32, 1>(compact, grid, st, blob, ws);
  if (s == 64 && r == 1) return launch_sr<64, 1>(compact, grid, st, blob, ws);
  return hipErrorInvalidValue;
}

// Workgroups of the kernel that fit on one CU at once.
int interp_blocks_per_cu(int s, int r, bool compact, int lds_bytes) {
  int blocks = 0;
  hipError_t e = hipErrorInvalidValue;
#define BLZ_OCC(KERNEL) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, KERNEL, kBlock, lds_bytes)
  (void)lds_bytes;
  if (s == 8 && r == 1) { if (compact) BLZ_OCC((interp_kernel<8, 1, true>)); else BLZ_OCC((interp_kernel<8, 1, false>)); }
  else if (s == 8 && r == 2) { if (compact) BLZ_OCC((interp_kernel<8, 2, true>)); else BLZ_OCC((interp_kernel<8, 2, false>)); }
  else if (s == 16 && r == 1) { if (compact) BLZ_OCC((interp_kernel<16, 1, true>)); else BLZ_OCC((interp_kernel<16, 1, false>)); }
  else if (s == 32 && r == 1) { if (compact) BLZ_OCC((interp_kernel<32, 1, true>)); else BLZ_OCC((interp_kernel<32, 1, false>)); }
  else if (s == 64 && r == 1) { if (compact) BLZ_OCC((interp_kernel<64, 1, true>)); else BLZ_OCC((interp_kernel<64, 1, false>)); }
#undef BLZ_OCC
  return (e == hipSuccess && blocks > 0) ? blocks : 1;
}

template <typename T>
static hipError_t launch_gather_t(const void *in, const uint8_t *in_valid, int64_t n_in, const void *idx,
                                  int idx_width, int64_t n_idx, void *out, uint8_t *out_valid,
                                  unsigned long long *null_count, int grid, hipStream_t st) {
  const T *i = (const T *)in;
  T *o = (T *)out;
  switch (idx_width) {
    case 1:
      hipLaunchKernelGGL((gather_kernel<T, int8_t>), dim3(grid), dim3(kBlock), 0, st, i, in_valid, n_in,
                         (const int8_t *)idx, n_idx, o, out_valid, null_count);
      break;
    case 2:
      hipLaunchKernelGGL((gather_kernel<T, int16_t>), dim3(grid), dim3(kBlock), 0, st, i, in_valid, n_in,
                         (const int16_t *)idx, n_idx, o, out_valid, null_count);
      break;
    case 4:
      hipLaunchKernelGGL((gather_kernel<T, int32_t>), dim3(grid), dim3(kBlock), 0, st, i, in_valid, n_in,
                         (const int32_t *)idx, n_idx, o, out_valid, null_count);
      break;
    case 8:
      hipLaunchKernelGGL((gather_kernel<T, int64_t>), dim3(grid), dim3(kBlock), 0, st, i, in_valid, n_in,
                         (const int64_t *)idx, n_idx, o, out_valid, null_count);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
hipError_t launch_gather(int width, const void *in, const uint8_t *in_valid, int64_t n_in, const void *idx,
                         int idx_width, int64_t n_idx, void *out, uint8_t *out_valid,
                         unsigned long long *null_count, int grid, hipStream_t st) {
  switch (width) {
    case 1: return launch_gather_t<uint8_t>(in, in_valid, n_in, idx, idx_width, n_idx, out, out_valid, null_count, grid, st);
    case 2: return launch_gather_t<uint16_t>(in, in_valid, n_in, idx, idx_width, n_idx, out, out_valid, null_count, grid, st);
    case 4: return launch_gather_t<uint32_t>(in, in_valid, n_in, idx, idx_width, n_idx, out, out_valid, null_count, grid, st);
    case 8: return launch_gather_t<uint64_t>(in, in_valid, n_in, idx, idx_width, n_idx, out, out_valid, null_count, grid, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_sequence(int32_t *out, int64_t n, int32_t init, int32_t step, int grid, hipStream_t st) {
  hipLaunchKernelGGL(sequence_kernel, dim3(grid), dim3(kBlock), 0, st, out, n, init, step);
  return hipGetLastError();
}

hipError_t launch_fill_zero(void *p, int64_t bytes, int grid, hipStream_t st) {
  hipLaunchKernelGGL(fill_zero_kernel, dim3(grid), dim3(kBlock), 0, st, (uint64_t *)p, (bytes + 7) / 8);
  return hipGetLastError();
}

hipError_t launch_synth(void *data, uint8_t *valid, int64_t n, int dtype, uint64_t seed, int dist, double a,
                        double b, double null_rate, hipStream_t st) {
  const int64_t blocks = (n + kBlock - 1) / kBlock;
  const int grid = (int)(blocks < 4096 ? (blocks > 0 ? blocks : 1) : 4096);
  hipLaunchKernelGGL(synth_kernel, dim3(grid), dim3(kBlock), 0, st, data, valid, n, dtype, seed, dist, a, b,
                     null_rate);
  return hipGetLastError();
}

hipError_t launch_fill_ranges(void *const *dst, const int64_t *bytes, int n, int grid, hipStream_t st) {
  hipLaunchKernelGGL(fill_ranges_kernel, dim3(grid), dim3(kBlock), 0, st, dst, bytes, n);
  return hipGetLastError();
}

hipError_t launch_null_count(const uint8_t *valid, int64_t n, unsigned long long *out, int grid,
                             hipStream_t st) {
  hipLaunchKernelGGL(null_count_kernel, dim3(grid), dim3(kBlock), 0, st, valid, n, out);
  return hipGetLastError();
}

template <int OP, int LD>
static hipError_t launch_reduce_t(const void *data, const uint8_t *valid, int64_t n, RedPart *parts, int grid,
                                  hipStream_t st) {
  constexpr bool FLT = LD == LD_F32 || LD == LD_F64;
  if (((uintptr_t)data & 15) == 0)
    hipLaunchKernelGGL((reduce_kernel<OP, FLT, LD, true>), dim3(grid), dim3(kBlock), 0, st, data, valid, n, parts);
  else
    hipLaunchKernelGGL((reduce_kernel<OP, FLT, LD, false>), dim3(grid), dim3(kBlock), 0, st, data, valid, n, parts);
  hipLaunchKernelGGL((reduce_final_kernel<OP, FLT>), dim3(1), dim3(kBlock), 0, st, parts, grid, parts + grid);
  return hipGetLastError();
}
template <int OP>
static hipError_t launch_reduce_op(const void *data, int load, const uint8_t *valid, int64_t n, RedPart *parts,
                                   int grid, hipStream_t st) {
  switch (load) {
    case LD_I8: return launch_reduce_t<OP, LD_I8>(data, valid, n, parts, grid, st);
    case LD_U8: return launch_reduce_t<OP, LD_U8>(data, valid, n, parts, grid, st);
    case LD_I16: return launch_reduce_t<OP, LD_I16>(data, valid, n, parts, grid, st);
    case LD_I32: return launch_reduce_t<OP, LD_I32>(data, valid, n, parts, grid, st);
    case LD_I64: return launch_reduce_t<OP, LD_I64>(data, valid, n, parts, grid, st);
    case LD_F32: return launch_reduce_t<OP, LD_F32>(data, valid, n, parts, grid, st);
    case LD_F64: return launch_reduce_t<OP, LD_F64>(data, valid, n, parts, grid, st);
    default: return hipErrorInvalidValue;
  }
}
int reduce_grid(int load, int64_t n) {
  const int w = load == LD_I64 || load == LD_F64 ? 8 : load == LD_I32 || load == LD_F32 ? 4 : load == LD_I16 ? 2 : 1;
  const int64_t step = (int64_t)kBlock * (16 / w) * kRedLoads;
  return (int)std::max<int64_t>(1, std::min<int64_t>(kRedMaxGrid, (n + step - 1) / step));
}
hipError_t launch_reduce(int op, const void *data, int load, const uint8_t *valid, int64_t n, RedPart *parts,
                         int grid, hipStream_t st) {
  switch (op) {
    case 0: return launch_reduce_op<0>(data, load, valid, n, parts, grid, st);
    case 1: return launch_reduce_op<1>(data, load, valid, n, parts, grid, st);
    case 2: return launch_reduce_op<2>(data, load, valid, n, parts, grid, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_copy(void *dst, const void *src, int64_t bytes, int grid, hipStream_t st) {
  hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(kBlock), 0, st, (uint4 *)dst, (const uint4 *)src, bytes / 16);
  return hipGetLastError();
}

// Scratch of the allocation-free entry point (blz_perform_operation_noalloc): the
// launch buffers of one projection program (blob, launch descriptor, null counts,
// counters) with nothing that grows with the rows; sized for the largest blob
// (kLdsBlobBytes) plus descriptor, 64 null counts and counters.
constexpr size_t kNoAllocScratchBytes = 16 << 10;
static_assert(kNoAllocScratchBytes >= kLdsBlobBytes + 256 + 512 + 128, "no-alloc scratch too small");
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

hipError_t noalloc_scratch(uint8_t **ptr, size_t *bytes, std::mutex **lock) {
  static std::mutex mu;
  static std::map<int, std::pair<uint8_t *, std::mutex *>> per_device;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> g(mu);
  auto it = per_device.find(dev);
  if (it == per_device.end()) {
    // the array's name is a host-side stand-in: its device address on this device
    // (the current one) is what kernels and copies take
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    it = per_device.emplace(dev, std::make_pair((uint8_t *)p, new std::mutex())).first;
  }
  *ptr = it->second.first;
  *bytes = kNoAllocScratchBytes;
  *lock = it->second.second;
  return hipSuccess;
}

int blob_capacity() { return kLdsBlobBytes; }
int block_threads() { return kBlock; }

}  // namespace blz

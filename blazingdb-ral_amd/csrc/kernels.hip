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
#pragma unroll
        for (int u = 0; u < R; u++)
          if (row[u] < n) store_narrow(kind, base, row[u], rf[slot * R + u]);
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
          if (lane == 0 && nulls) atomicAdd(ws.null_counts + o, (unsigned long long)nulls);
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
        if (wave_cnt > 0) {
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
          if (lane == 0 && nulls) atomicAdd(ws.null_counts + o, (unsigned long long)nulls);
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
}

// ---------------------------------------------------------------------------
// staged_kernel: the same interpreter, with each tile's input columns (and the
// bitmaps the validity program reads) first copied HBM -> LDS in 4 KiB pieces,
// 16 bytes per thread, all pieces of a tile in flight together. Rows are then
// interpreted from LDS, so the bytes a workgroup keeps in flight no longer depend
// on the register file's size. Tile = T rows (a multiple of 256 * R); wave w owns
// local rows [w*T/4, (w+1)*T/4), processed 64*R rows per pass.
// COMPACT: pass A evaluates the predicate for the whole tile and keeps one
// 64-row keep mask per group in LDS; after the look-back, pass B evaluates the
// projection only for the groups with surviving rows and scatters them.

DEV void stage_tile(const CONSTANT Chunk *chunks, int nch, int64_t tile, int64_t rows, int T,
                    uint8_t *lds) {
  const int tid = threadIdx.x;
  const bool full = rows == T;
  for (int k0 = 0; k0 < nch; k0 += 8) {
    u32x4 v[8];
    int lim[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int k = k0 + j;
      lim[j] = 0;
      if (k < nch) {
        const CONSTANT Chunk &ck = chunks[k];
        const int sh = ck.shift;
        const int64_t slice = full ? ck.tile_bytes : (sh >= 0 ? (rows << sh) : ((rows + 7) >> 3));
        const int64_t off = (int64_t)ck.chunk_off + tid * 16;
        lim[j] = (int)(slice - off < 16 ? slice - off : 16);
        const uint8_t *src = (const uint8_t *)ck.base + tile * (int64_t)ck.tile_bytes + off;
        if (lim[j] >= 16) {
          v[j] = gload<u32x4>(src, 0);
        } else if (lim[j] > 0) {  // tail of the last tile: byte loads, never past the end
          uint32_t w[4] = {0, 0, 0, 0};
          for (int b = 0; b < lim[j]; b++) w[b >> 2] |= (uint32_t)gload<uint8_t>(src, b) << (8 * (b & 3));
          v[j] = u32x4{w[0], w[1], w[2], w[3]};
        }
      }
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int k = k0 + j;
      if (k < nch && lim[j] > 0)
        *(u32x4 *)(lds + chunks[k].lds_off + tid * 16) = v[j];
    }
  }
}

// Inputs of one pass from LDS into the register file (all reads issued first).
template <int S, int R>
DEV void read_pass(const CONSTANT InCol *in, int n_in, const uint8_t *lds, const int (&lr)[R],
                   uint64_t (&rf)[S * R]) {
  constexpr int MI = S < kMaxUnrolledInputs ? S : kMaxUnrolledInputs;
  uint64_t w64[MI][R];
  uint32_t w32[MI][R];
#pragma unroll
  for (int c = 0; c < MI; c++) {
    if (c < n_in) {
      const int kind = in[c].load;
      const uint8_t *p = lds + in[c].lds_off;
      if (is_wide(kind)) {
#pragma unroll
        for (int u = 0; u < R; u++) w64[c][u] = *(const uint64_t *)(p + 8 * lr[u]);
      } else if (kind == LD_I32 || kind == LD_F32) {
#pragma unroll
        for (int u = 0; u < R; u++) w32[c][u] = *(const uint32_t *)(p + 4 * lr[u]);
      } else if (kind == LD_I16) {
#pragma unroll
        for (int u = 0; u < R; u++) w32[c][u] = *(const uint16_t *)(p + 2 * lr[u]);
      } else {
#pragma unroll
        for (int u = 0; u < R; u++) w32[c][u] = p[lr[u]];
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

template <int S, int R, bool COMPACT>
__global__ void __launch_bounds__(kBlock)
staged_kernel(const uint8_t *__restrict__ blob_arg, Workspace ws) {
  extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
  __shared__ int64_t s_tile;
  __shared__ uint64_t s_wave_cnt[kWaves];
  __shared__ uint64_t s_prefix;
  __shared__ uint64_t s_keep[64];  // keep mask per 64-row group (T <= 4096)

  const CONSTANT uint8_t *blob = (const CONSTANT uint8_t *)blob_arg;
  const CONSTANT ProgHeader &h = *(const CONSTANT ProgHeader *)blob;
  const CONSTANT InCol *in = (const CONSTANT InCol *)(blob + h.off_in);
  const CONSTANT OutCol *out = (const CONSTANT OutCol *)(blob + h.off_out);
  const CONSTANT DevOp *ops = (const CONSTANT DevOp *)(blob + h.off_ops);
  const CONSTANT VOp *vops = (const CONSTANT VOp *)(blob + h.off_vops);
  const CONSTANT Chunk *chunks = (const CONSTANT Chunk *)(blob + h.off_chunks);
  const int64_t n = h.n_rows;
  const int n_in = h.n_in, n_out = h.n_out;
  const int n_ops_p = h.n_ops_p, n_ops_q = h.n_ops_q;
  const int n_vops_p = h.n_vops_p, n_vops_q = h.n_vops_q;
  const int any_valid_out = h.any_valid_out;
  const int T = h.tile_rows, nch = h.n_chunks;
  const int passes = T / (kWaves * 64 * R);

  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t n_tiles = (n + T - 1) / T;

  uint64_t rf[S * R];
  uint64_t vrf[S];
#pragma unroll
  for (int i = 0; i < S * R; i++) rf[i] = 0;
#pragma unroll
  for (int i = 0; i < S; i++) vrf[i] = 0;

  int64_t tile = blockIdx.x;
  if (COMPACT) {
    if (threadIdx.x == 0) s_tile = (int64_t)atomicAdd(ws.tile_counter, 1u);
    __syncthreads();
    tile = s_tile;
  }
  while (tile < n_tiles) {
    const int64_t row0 = tile * T;
    const int64_t rows = (n - row0) < T ? (n - row0) : T;
    stage_tile(chunks, nch, tile, rows, T, lds);
    __syncthreads();

    if (!COMPACT) {
      for (int p = 0; p < passes; p++) {
        const int lb = wave * (T / kWaves) + p * 64 * R;
        int lr[R];
        int64_t grow[R], lrow[R];
#pragma unroll
        for (int u = 0; u < R; u++) {
          lr[u] = lb + u * 64 + lane;
          grow[u] = row0 + lr[u];
          lrow[u] = grow[u] < n ? grow[u] : n - 1;
        }
        if (row0 + lb >= n) break;  // wave-uniform: nothing left in this tile
        read_pass<S, R>(in, n_in, lds, lr, rf);
        run_ops<S, R>(ops, n_ops_q, rf, lrow, in);
        if (any_valid_out) {
          read_pass_valid<S, R>(in, n_in, lds, lb >> 6, vrf);
          run_vops<S>(vops, n_vops_q, vrf);
        }
        for (int o = 0; o < n_out; o++) {
          const int kind = out[o].store;
          const int slot = out[o].slot;
          void *base = out[o].data;
#pragma unroll
          for (int u = 0; u < R; u++)
            if (grow[u] < n) store_narrow(kind, base, grow[u], rf[slot * R + u]);
          void *vp = out[o].valid;
          if (vp != nullptr) {
            const uint64_t word = vsrc_word<S>(out[o].vsrc, vrf);
            const int64_t w = ((row0 + lb) >> 6) + lane;
            uint64_t nulls = 0;
            if (lane < R && w * 64 < n) {
              store_vword(vp, w, n, word);
              nulls = __builtin_popcountll(~word & low_mask(n - w * 64));
            }
            for (int off = 32; off > 0; off >>= 1) nulls += __shfl_xor(nulls, off);
            if (lane == 0 && nulls) atomicAdd(ws.null_counts + o, (unsigned long long)nulls);
          }
        }
      }
      __syncthreads();  // the tile buffer is reused by the next tile
      tile += gridDim.x;
      continue;
    }

    // ---- COMPACT, pass A: predicate over the whole tile -> keep masks ----
    const int st_slot = h.stencil_slot, st_width = h.stencil_width, st_vsrc = h.stencil_vsrc;
    const uint64_t st_mask = low_mask(8 * st_width);
    uint64_t wave_cnt = 0;
    for (int p = 0; p < passes; p++) {
      const int lb = wave * (T / kWaves) + p * 64 * R;
      int lr[R];
      int64_t grow[R], lrow[R];
#pragma unroll
      for (int u = 0; u < R; u++) {
        lr[u] = lb + u * 64 + lane;
        grow[u] = row0 + lr[u];
        lrow[u] = grow[u] < n ? grow[u] : n - 1;
      }
      if (row0 + lb >= n) {
#pragma unroll
        for (int u = 0; u < R; u++)
          if (lane == 0) s_keep[(lb >> 6) + u] = 0;
        continue;
      }
      read_pass<S, R>(in, n_in, lds, lr, rf);
      run_ops<S, R>(ops, n_ops_p, rf, lrow, in);
      read_pass_valid<S, R>(in, n_in, lds, lb >> 6, vrf);
      run_vops<S>(vops, n_vops_p, vrf);
      const uint64_t vstencil = vsrc_word<S>(st_vsrc, vrf);
#pragma unroll
      for (int u = 0; u < R; u++) {
        const bool vbit = (readlane64(vstencil, u) >> lane) & 1;
        const bool keep = grow[u] < n && ((rf[st_slot * R + u] & st_mask) != 0) && vbit;
        const uint64_t km = __builtin_amdgcn_ballot_w64(keep);
        if (lane == 0) s_keep[(lb >> 6) + u] = km;
        wave_cnt += __builtin_popcountll(km);
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
          uint64_t st = FP;  // virtual inclusive prefix 0 before tile 0
          if (idx >= 0)
            st = __hip_atomic_load(ws.tile_state + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          const uint64_t flag = st >> 62;
          const uint64_t is_x = __builtin_amdgcn_ballot_w64(flag == 0);
          const uint64_t is_p = __builtin_amdgcn_ballot_w64(flag == 2);
          const uint64_t upto = is_p ? (((is_p & (~is_p + 1)) << 1) - 1) : ~0ull;
          if (is_x & upto) {
            __builtin_amdgcn_s_sleep(1);
            continue;
          }
          uint64_t v = ((upto >> lane) & 1) ? (st & VM) : 0;
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
      // ---- pass B: projection / gather of the surviving rows, in row order ----
      uint64_t *vbits = (uint64_t *)(lds + h.lds_vbits_off);
      const int n_vout = h.n_valid_out;
      const int nw = (int)(((tile_prefix & 63) + tile_cnt + 63) >> 6);
      if (n_vout > 0) {
        for (int i = threadIdx.x; i < n_vout * (T / 64 + 2); i += kBlock) vbits[i] = 0;
        __syncthreads();
      }
      uint64_t run = tile_prefix + wave_prefix;
      uint64_t vnulls[4] = {0, 0, 0, 0};  // per output with validity (first four), this lane
      for (int p = 0; p < passes; p++) {
        const int lb = wave * (T / kWaves) + p * 64 * R;
        if (row0 + lb >= n) break;
        uint64_t km[R], pos[R];
        uint64_t any = 0;
#pragma unroll
        for (int u = 0; u < R; u++) {
          km[u] = s_keep[(lb >> 6) + u];
          const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(km[u] >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)km[u], 0u));
          pos[u] = run + below;
          run += __builtin_popcountll(km[u]);
          any |= km[u];
        }
        if (!any) continue;  // wave-uniform
        int lr[R];
        int64_t lrow[R];
#pragma unroll
        for (int u = 0; u < R; u++) {
          lr[u] = lb + u * 64 + lane;
          const int64_t g = row0 + lr[u];
          lrow[u] = g < n ? g : n - 1;
        }
        read_pass<S, R>(in, n_in, lds, lr, rf);
        run_ops<S, R>(ops + n_ops_p, n_ops_q, rf, lrow, in);
        if (any_valid_out) {
          read_pass_valid<S, R>(in, n_in, lds, lb >> 6, vrf);
          run_vops<S>(vops + n_vops_p, n_vops_q, vrf);
        }
        for (int o = 0; o < n_out; o++) {
          const int kind = out[o].store;
          const int slot = out[o].slot;
          void *base = out[o].data;
#pragma unroll
          for (int u = 0; u < R; u++)
            if ((km[u] >> lane) & 1) store_narrow(kind, base, (int64_t)pos[u], rf[slot * R + u]);
          if (out[o].valid != nullptr) {
            const int vb = out[o].vbuf;
            const uint64_t vword = vsrc_word<S>(out[o].vsrc, vrf);
#pragma unroll
            for (int u = 0; u < R; u++) {
              if ((km[u] >> lane) & 1) {
                if ((readlane64(vword, u) >> lane) & 1) {
                  const uint64_t local = (tile_prefix & 63) + (pos[u] - tile_prefix);
                  atomicOr((unsigned long long *)&vbits[vb * (T / 64 + 2) + (local >> 6)], 1ull << (local & 63));
                } else if (vb < 4) {
                  vnulls[vb]++;
                } else {
                  atomicAdd(ws.null_counts + o, 1ull);
                }
              }
            }
          }
        }
      }
      if (n_vout > 0) {
        __syncthreads();
        const int64_t w0 = (int64_t)(tile_prefix >> 6);
        const uint64_t end = tile_prefix + tile_cnt;
        for (int o = 0; o < n_out; o++) {
          void *vp = out[o].valid;
          if (vp == nullptr) continue;
          const int vb = out[o].vbuf;
          for (int i = threadIdx.x; i < nw; i += kBlock) {
            const bool first_shared = (i == 0) && (tile_prefix & 63);
            const bool last_shared = (i == nw - 1) && (end & 63);
            unsigned long long *gw = (unsigned long long *)((uint8_t *)vp + 8 * (w0 + i));
            const uint64_t word = vbits[vb * (T / 64 + 2) + i];
            if (first_shared || last_shared) atomicOr(gw, (unsigned long long)word);
            else gstore<uint64_t>(gw, 0, word);
          }
          if (vb < 4) {
            uint64_t nl = vnulls[vb];
            for (int off = 32; off > 0; off >>= 1) nl += __shfl_xor(nl, off);
            if (lane == 0 && nl) atomicAdd(ws.null_counts + o, (unsigned long long)nl);
          }
        }
      }
    }
    // Next tile (the LDS tile buffer and keep masks are reused).
    __syncthreads();
    if (threadIdx.x == 0) s_tile = (int64_t)atomicAdd(ws.tile_counter, 1u);
    __syncthreads();
    tile = s_tile;
  }
}

// ---------------------------------------------------------------------------
// acc_kernel: the staged kernel with the program in accumulator form (AccOp).
// Per op one uniform dispatch, at most one operand fetch (immediate, an input
// column read straight from the LDS tile, or one of NT temporaries) and no
// result write-back; EMIT stores the accumulator to an output at once. With few
// temporaries a wave covers 64 * R rows (R up to 8) per dispatch.

// Operand rows lb + lane + 64u (u < R) of an input column staged in LDS, widened.
template <int R>
DEV void lds_operand(int kind, const uint8_t *p, int lb, uint64_t (&B)[R]) {
  const int lane = threadIdx.x & 63;
  const int r0 = lb + lane;
  if (is_wide(kind)) {
#pragma unroll
    for (int u = 0; u < R; u++) B[u] = *(const uint64_t *)(p + 8 * (r0 + 64 * u));
  } else if (kind == LD_I32) {
#pragma unroll
    for (int u = 0; u < R; u++) B[u] = (uint64_t)(int64_t)*(const int32_t *)(p + 4 * (r0 + 64 * u));
  } else if (kind == LD_F32) {
#pragma unroll
    for (int u = 0; u < R; u++) B[u] = d2b((double)*(const float *)(p + 4 * (r0 + 64 * u)));
  } else if (kind == LD_I16) {
#pragma unroll
    for (int u = 0; u < R; u++) B[u] = (uint64_t)(int64_t)*(const int16_t *)(p + 2 * (r0 + 64 * u));
  } else if (kind == LD_I8) {
#pragma unroll
    for (int u = 0; u < R; u++) B[u] = (uint64_t)(int64_t)*(const int8_t *)(p + (r0 + 64 * u));
  } else if (kind == LD_U8) {
#pragma unroll
    for (int u = 0; u < R; u++) B[u] = (uint64_t)p[r0 + 64 * u];
  } else {
#pragma unroll
    for (int u = 0; u < R; u++) B[u] = 0;
  }
}

// acc &= (B CMP imm) for the comparison `sub`.
template <int R>
DEV void cmp_and(int sub, uint64_t (&acc)[R], const uint64_t (&B)[R], uint64_t im) {
  const double fi = b2d(im);
  const int64_t ii = (int64_t)im;
#define BLZ_CMPAND(CODE, EXPR)                      \
  case CODE:                                        \
    _Pragma("unroll") for (int u = 0; u < R; u++) { \
      const uint64_t b = B[u];                      \
      acc[u] &= (uint64_t)(EXPR);                   \
    }                                               \
    break;
  switch (sub) {
    BLZ_CMPAND(OP_LT_I, (int64_t)b < ii)
    BLZ_CMPAND(OP_LE_I, (int64_t)b <= ii)
    BLZ_CMPAND(OP_GT_I, (int64_t)b > ii)
    BLZ_CMPAND(OP_GE_I, (int64_t)b >= ii)
    BLZ_CMPAND(OP_EQ_I, b == im)
    BLZ_CMPAND(OP_NE_I, b != im)
    BLZ_CMPAND(OP_LT_F, b2d(b) < fi)
    BLZ_CMPAND(OP_LE_F, b2d(b) <= fi)
    BLZ_CMPAND(OP_GT_F, b2d(b) > fi)
    BLZ_CMPAND(OP_GE_F, b2d(b) >= fi)
    BLZ_CMPAND(OP_EQ_F, b2d(b) == fi)
    BLZ_CMPAND(OP_NE_F, b2d(b) != fi)
    default: break;
  }
#undef BLZ_CMPAND
}

// Runs accumulator ops over rows lb + lane + 64u of the tile starting at row0.
// EMIT stores row u to output position obase + orel[u] when oput[u].
template <int NT, int R>
DEV void run_acc(const CONSTANT AccOp *ops, int nops, uint64_t (&acc)[R], uint64_t (&tmp)[NT * R],
                 const CONSTANT InCol *in, const CONSTANT OutCol *out, const uint8_t *lds, int lb,
                 int64_t row0, int64_t n, int64_t obase, const int (&orel)[R], const bool (&oput)[R]) {
  for (int k = 0; k < nops; k++) {
    const uint64_t ctl = *(const CONSTANT uint64_t *)(ops + k);
    const int code = (int)(ctl & 0xFF), sub = (int)((ctl >> 8) & 0xFF);
    const int okind = (int)((ctl >> 16) & 0xFF), oidx = (int)((ctl >> 24) & 0xFF);
    const int cv = (int)((ctl >> 32) & 0xFF), fl = (int)((ctl >> 40) & 0xFF);
    const int vcol = (int)((ctl >> 48) & 0xFF);
    if (code == A_STORE) {
#pragma unroll
      for (int u = 0; u < R; u++) tmp[oidx * R + u] = acc[u];
      continue;
    }
    if (code == A_EMIT) {
      const int kind = out[oidx].store;
      void *base = out[oidx].data;
#pragma unroll
      for (int u = 0; u < R; u++)
        if (oput[u]) store_narrow(kind, base, obase + orel[u], acc[u]);
      continue;
    }
    uint64_t B[R];
    if (okind == OK_IMM) {
      const uint64_t im = ops[k].imm;
#pragma unroll
      for (int u = 0; u < R; u++) B[u] = im;
    } else if (okind == OK_INPUT) {
      lds_operand<R>(in[oidx].load, lds + in[oidx].lds_off, lb, B);
    } else if (okind == OK_TEMP) {
#pragma unroll
      for (int u = 0; u < R; u++) B[u] = tmp[oidx * R + u];
    } else {
#pragma unroll
      for (int u = 0; u < R; u++) B[u] = 0;
    }
    const int ca = cv & 3, cb = (cv >> 2) & 3, co = (cv >> 4) & 3;
    if (cb) {
#pragma nounroll
      for (int u = 0; u < R; u++) B[u] = conv_slow(B[u], cb);
    }
    if (code == A_LOAD) {
#pragma unroll
      for (int u = 0; u < R; u++) acc[u] = B[u];
    } else if (code == A_CMPAND) {
      cmp_and<R>(sub, acc, B, ops[k].imm2);
    } else {
      if (ca) {
#pragma nounroll
        for (int u = 0; u < R; u++) acc[u] = conv_slow(acc[u], ca);
      }
      bool vb[R];
      if (vcol != 0xFF) {
        const void *vp = in[vcol].valid;
        const int lane = threadIdx.x & 63;
#pragma unroll
        for (int u = 0; u < R; u++) {
          const int64_t g = row0 + lb + lane + 64 * u;
          vb[u] = valid_bit(vp, g < n ? g : n - 1);
        }
      } else {
#pragma unroll
        for (int u = 0; u < R; u++) vb[u] = true;
      }
      if (code == A_UN) exec_op<R>(sub, acc, acc, vb, acc);
      else if (fl & F_REV) exec_op<R>(sub, B, acc, vb, acc);
      else exec_op<R>(sub, acc, B, vb, acc);
    }
    if (co) {
#pragma nounroll
      for (int u = 0; u < R; u++) acc[u] = conv_slow(acc[u], co);
    }
  }
}

template <int NT, int R, bool COMPACT>
__global__ void __launch_bounds__(kBlock)
acc_kernel(const uint8_t *__restrict__ blob_arg, Workspace ws) {
  extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
  __shared__ int64_t s_tile;
  __shared__ uint64_t s_wave_cnt[kWaves];
  __shared__ uint64_t s_prefix;
  __shared__ uint64_t s_keep[64];  // keep mask per 64-row group (T <= 4096)
  constexpr int VS = 8;            // validity slots

  const CONSTANT uint8_t *blob = (const CONSTANT uint8_t *)blob_arg;
  const CONSTANT ProgHeader &h = *(const CONSTANT ProgHeader *)blob;
  const CONSTANT InCol *in = (const CONSTANT InCol *)(blob + h.off_in);
  const CONSTANT OutCol *out = (const CONSTANT OutCol *)(blob + h.off_out);
  const CONSTANT AccOp *ops = (const CONSTANT AccOp *)(blob + h.off_ops);
  const CONSTANT VOp *vops = (const CONSTANT VOp *)(blob + h.off_vops);
  const CONSTANT Chunk *chunks = (const CONSTANT Chunk *)(blob + h.off_chunks);
  const int64_t n = h.n_rows;
  const int n_in = h.n_in, n_out = h.n_out;
  const int n_ops_p = h.n_ops_p, n_ops_q = h.n_ops_q;
  const int n_vops_p = h.n_vops_p, n_vops_q = h.n_vops_q;
  const int any_valid_out = h.any_valid_out;
  const int T = h.tile_rows, nch = h.n_chunks;
  const int passes = T / (kWaves * 64 * R);

  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t n_tiles = (n + T - 1) / T;

  // Registers are written by the program before they are read (the compiler on
  // the host guarantees it), so none is initialised here: none is live across the
  // staging of a tile.
  uint64_t acc[R];
  uint64_t tmp[NT * R];
  uint64_t vrf[VS];

  int64_t tile = blockIdx.x;
  if (COMPACT) {
    if (threadIdx.x == 0) s_tile = (int64_t)atomicAdd(ws.tile_counter, 1u);
    __syncthreads();
    tile = s_tile;
  }
  while (tile < n_tiles) {
    const int64_t row0 = tile * T;
    const int64_t rows = (n - row0) < T ? (n - row0) : T;
    stage_tile(chunks, nch, tile, rows, T, lds);
    __syncthreads();

    if (!COMPACT) {
      for (int p = 0; p < passes; p++) {
        const int lb = wave * (T / kWaves) + p * 64 * R;
        if (row0 + lb >= n) break;  // wave-uniform: nothing left in this tile
        int orel[R];
        bool put[R];
#pragma unroll
        for (int u = 0; u < R; u++) {
          orel[u] = lb + u * 64 + lane;
          put[u] = row0 + orel[u] < n;
          acc[u] = 0;
        }
        run_acc<NT, R>(ops, n_ops_q, acc, tmp, in, out, lds, lb, row0, n, row0, orel, put);
        if (any_valid_out) {
          read_pass_valid<VS, R>(in, n_in, lds, lb >> 6, vrf);
          run_vops<VS>(vops, n_vops_q, vrf);
          for (int o = 0; o < n_out; o++) {
            void *vp = out[o].valid;
            if (vp == nullptr) continue;
            const uint64_t word = vsrc_word<VS>(out[o].vsrc, vrf);
            const int64_t w = ((row0 + lb) >> 6) + lane;
            uint64_t nulls = 0;
            if (lane < R && w * 64 < n) {
              store_vword(vp, w, n, word);
              nulls = __builtin_popcountll(~word & low_mask(n - w * 64));
            }
            for (int off = 32; off > 0; off >>= 1) nulls += __shfl_xor(nulls, off);
            if (lane == 0 && nulls) atomicAdd(ws.null_counts + o, (unsigned long long)nulls);
          }
        }
      }
      __syncthreads();  // the tile buffer is reused by the next tile
      tile += gridDim.x;
      continue;
    }

    // ---- COMPACT, pass A: predicate over the whole tile -> keep masks ----
    const int st_width = h.stencil_width, st_vsrc = h.stencil_vsrc;
    const uint64_t st_mask = low_mask(8 * st_width);
    uint64_t wave_cnt = 0;
    for (int p = 0; p < passes; p++) {
      const int lb = wave * (T / kWaves) + p * 64 * R;
      if (row0 + lb >= n) {
        if (lane == 0) {
#pragma unroll
          for (int u = 0; u < R; u++) s_keep[(lb >> 6) + u] = 0;
        }
        continue;
      }
      int orel[R];
      bool none[R];
#pragma unroll
      for (int u = 0; u < R; u++) {
        orel[u] = 0;
        none[u] = false;
        acc[u] = 0;
      }
      run_acc<NT, R>(ops, n_ops_p, acc, tmp, in, out, lds, lb, row0, n, 0, orel, none);
      read_pass_valid<VS, R>(in, n_in, lds, lb >> 6, vrf);
      run_vops<VS>(vops, n_vops_p, vrf);
      const uint64_t vstencil = vsrc_word<VS>(st_vsrc, vrf);
#pragma unroll
      for (int u = 0; u < R; u++) {
        const bool vbit = (readlane64(vstencil, u) >> lane) & 1;
        const bool keep = row0 + lb + u * 64 + lane < n && ((acc[u] & st_mask) != 0) && vbit;
        const uint64_t km = __builtin_amdgcn_ballot_w64(keep);
        if (lane == 0) s_keep[(lb >> 6) + u] = km;
        wave_cnt += __builtin_popcountll(km);
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
          uint64_t st = FP;  // virtual inclusive prefix 0 before tile 0
          if (idx >= 0)
            st = __hip_atomic_load(ws.tile_state + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          const uint64_t flag = st >> 62;
          const uint64_t is_x = __builtin_amdgcn_ballot_w64(flag == 0);
          const uint64_t is_p = __builtin_amdgcn_ballot_w64(flag == 2);
          const uint64_t upto = is_p ? (((is_p & (~is_p + 1)) << 1) - 1) : ~0ull;
          if (is_x & upto) {
            __builtin_amdgcn_s_sleep(1);
            continue;
          }
          uint64_t v = ((upto >> lane) & 1) ? (st & VM) : 0;
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
      // ---- pass B: projection / gather of the surviving rows, in row order ----
      uint64_t *vbits = (uint64_t *)(lds + h.lds_vbits_off);
      const int n_vout = h.n_valid_out;
      const int nw = (int)(((tile_prefix & 63) + tile_cnt + 63) >> 6);
      if (n_vout > 0) {
        for (int i = threadIdx.x; i < n_vout * (T / 64 + 2); i += kBlock) vbits[i] = 0;
        __syncthreads();
      }
      uint64_t run = tile_prefix + wave_prefix;
      uint64_t vnulls[4] = {0, 0, 0, 0};
      for (int p = 0; p < passes; p++) {
        const int lb = wave * (T / kWaves) + p * 64 * R;
        if (row0 + lb >= n) break;
        int rel[R];  // output position - tile_prefix
        bool keep[R];
        uint64_t any = 0;
#pragma unroll
        for (int u = 0; u < R; u++) {
          const uint64_t km = s_keep[(lb >> 6) + u];
          const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(km >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)km, 0u));
          rel[u] = (int)(run - tile_prefix) + (int)below;
          keep[u] = (km >> lane) & 1;
          run += __builtin_popcountll(km);
          any |= km;
          acc[u] = 0;
        }
        if (!any) continue;  // wave-uniform
        run_acc<NT, R>(ops + n_ops_p, n_ops_q, acc, tmp, in, out, lds, lb, row0, n, (int64_t)tile_prefix, rel,
                       keep);
        if (any_valid_out) {
          read_pass_valid<VS, R>(in, n_in, lds, lb >> 6, vrf);
          run_vops<VS>(vops + n_vops_p, n_vops_q, vrf);
          for (int o = 0; o < n_out; o++) {
            if (out[o].valid == nullptr) continue;
            const int vb = out[o].vbuf;
            const uint64_t vword = vsrc_word<VS>(out[o].vsrc, vrf);
#pragma unroll
            for (int u = 0; u < R; u++) {
              if (keep[u]) {
                if ((readlane64(vword, u) >> lane) & 1) {
                  const uint64_t local = (tile_prefix & 63) + (uint64_t)rel[u];
                  atomicOr((unsigned long long *)&vbits[vb * (T / 64 + 2) + (local >> 6)], 1ull << (local & 63));
                } else if (vb < 4) {
                  vnulls[vb]++;
                } else {
                  atomicAdd(ws.null_counts + o, 1ull);
                }
              }
            }
          }
        }
      }
      if (n_vout > 0) {
        __syncthreads();
        const int64_t w0 = (int64_t)(tile_prefix >> 6);
        const uint64_t end = tile_prefix + tile_cnt;
        for (int o = 0; o < n_out; o++) {
          void *vp = out[o].valid;
          if (vp == nullptr) continue;
          const int vb = out[o].vbuf;
          for (int i = threadIdx.x; i < nw; i += kBlock) {
            const bool first_shared = (i == 0) && (tile_prefix & 63);
            const bool last_shared = (i == nw - 1) && (end & 63);
            unsigned long long *gw = (unsigned long long *)((uint8_t *)vp + 8 * (w0 + i));
            const uint64_t word = vbits[vb * (T / 64 + 2) + i];
            if (first_shared || last_shared) atomicOr(gw, (unsigned long long)word);
            else gstore<uint64_t>(gw, 0, word);
          }
          if (vb < 4) {
            uint64_t nl = vnulls[vb];
            for (int off = 32; off > 0; off >>= 1) nl += __shfl_xor(nl, off);
            if (lane == 0 && nl) atomicAdd(ws.null_counts + o, (unsigned long long)nl);
          }
        }
      }
    }
    // Next tile (the LDS tile buffer and keep masks are reused).
    __syncthreads();
    if (threadIdx.x == 0) s_tile = (int64_t)atomicAdd(ws.tile_counter, 1u);
    __syncthreads();
    tile = s_tile;
  }
}

// ---------------------------------------------------------------------------
// gather: out[i] = in[idx[i]] (idx >= 0), validity bits gathered, null count.
template <typename T>
__global__ void __launch_bounds__(kBlock)
gather_kernel(const T *__restrict__ in, const uint8_t *__restrict__ in_valid, int64_t n_in,
              const int32_t *__restrict__ idx, int64_t n_idx, T *__restrict__ out,
              uint8_t *__restrict__ out_valid, unsigned long long *null_count) {
  const int lane = threadIdx.x & 63;
  const int64_t n_words = (n_idx + 63) >> 6;
  const int64_t wave_id = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int64_t n_waves = ((int64_t)gridDim.x * kBlock) >> 6;
  uint64_t nulls = 0;
  for (int64_t w = wave_id; w < n_words; w += n_waves) {
    const int64_t i = w * 64 + lane;
    bool vbit = false;
    if (i < n_idx) {
      const int32_t j = idx[i];
      if (j >= 0 && j < n_in) {  // an index past the input is treated like a negative one
        out[i] = in[j];
        vbit = valid_bit(in_valid, j);
      }
    }
    if (out_valid != nullptr) {
      const uint64_t word = __builtin_amdgcn_ballot_w64(vbit);
      if (lane == 0) {
        store_vword(out_valid, w, n_idx, word);
        nulls += (uint64_t)__builtin_popcountll(~word & low_mask(n_idx - w * 64));
      }
    }
  }
  if (out_valid != nullptr && lane == 0 && nulls) atomicAdd(null_count, (unsigned long long)nulls);
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
  if ((threadIdx.x & 63) == 0 && nulls) atomicAdd(nulls_out, (unsigned long long)nulls);
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

template <int S, int R>
static hipError_t launch_staged_sr(bool compact, int grid, int lds, hipStream_t st, const uint8_t *blob,
                                   Workspace ws) {
  if (compact)
    hipLaunchKernelGGL((staged_kernel<S, R, true>), dim3(grid), dim3(kBlock), lds, st, blob, ws);
  else
    hipLaunchKernelGGL((staged_kernel<S, R, false>), dim3(grid), dim3(kBlock), lds, st, blob, ws);
  return hipGetLastError();
}

template <int NT, int R>
static hipError_t launch_acc_tr(bool compact, int grid, int lds, hipStream_t st, const uint8_t *blob,
                                Workspace ws) {
  if (compact)
    hipLaunchKernelGGL((acc_kernel<NT, R, true>), dim3(grid), dim3(kBlock), lds, st, blob, ws);
  else
    hipLaunchKernelGGL((acc_kernel<NT, R, false>), dim3(grid), dim3(kBlock), lds, st, blob, ws);
  return hipGetLastError();
}

// Accumulator kernel shapes: NT temporaries x R rows per lane, NT * R <= 16.
hipError_t launch_acc(int nt, int r, bool compact, int grid, int lds_bytes, hipStream_t st,
                      const uint8_t *blob, Workspace ws) {
  if (nt == 2 && r == 8) return launch_acc_tr<2, 8>(compact, grid, lds_bytes, st, blob, ws);
  if (nt == 4 && r == 4) return launch_acc_tr<4, 4>(compact, grid, lds_bytes, st, blob, ws);
  if (nt == 8 && r == 2) return launch_acc_tr<8, 2>(compact, grid, lds_bytes, st, blob, ws);
  return hipErrorInvalidValue;
}

int acc_blocks_per_cu(int nt, int r, bool compact, int lds_bytes) {
  int blocks = 0;
  hipError_t e = hipErrorInvalidValue;
#define BLZ_OCC(KERNEL) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, KERNEL, kBlock, lds_bytes)
  if (nt == 2 && r == 8) { if (compact) BLZ_OCC((acc_kernel<2, 8, true>)); else BLZ_OCC((acc_kernel<2, 8, false>)); }
  else if (nt == 4 && r == 4) { if (compact) BLZ_OCC((acc_kernel<4, 4, true>)); else BLZ_OCC((acc_kernel<4, 4, false>)); }
  else if (nt == 8 && r == 2) { if (compact) BLZ_OCC((acc_kernel<8, 2, true>)); else BLZ_OCC((acc_kernel<8, 2, false>)); }
#undef BLZ_OCC
  return (e == hipSuccess && blocks > 0) ? blocks : 1;
}

// Register-file shapes: S slots x R rows per lane. S*R <= 16 keeps the file in
// VGPRs; S = 32 / 64 (R = 1) serve large plans with the file in scratch.
// lds_bytes > 0 selects the LDS-staged kernel (shapes (8,1), (8,2), (16,1)).
hipError_t launch_interp(int s, int r, bool compact, int grid, int lds_bytes, hipStream_t st,
                         const uint8_t *blob, Workspace ws) {
  if (lds_bytes > 0) {
    if (s == 8 && r == 1) return launch_staged_sr<8, 1>(compact, grid, lds_bytes, st, blob, ws);
    if (s == 8 && r == 2) return launch_staged_sr<8, 2>(compact, grid, lds_bytes, st, blob, ws);
    if (s == 16 && r == 1) return launch_staged_sr<16, 1>(compact, grid, lds_bytes, st, blob, ws);
    return hipErrorInvalidValue;
  }
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
  if (lds_bytes > 0) {
    if (s == 8 && r == 1) { if (compact) BLZ_OCC((staged_kernel<8, 1, true>)); else BLZ_OCC((staged_kernel<8, 1, false>)); }
    else if (s == 8 && r == 2) { if (compact) BLZ_OCC((staged_kernel<8, 2, true>)); else BLZ_OCC((staged_kernel<8, 2, false>)); }
    else if (s == 16 && r == 1) { if (compact) BLZ_OCC((staged_kernel<16, 1, true>)); else BLZ_OCC((staged_kernel<16, 1, false>)); }
  } else {
    if (s == 8 && r == 1) { if (compact) BLZ_OCC((interp_kernel<8, 1, true>)); else BLZ_OCC((interp_kernel<8, 1, false>)); }
    else if (s == 8 && r == 2) { if (compact) BLZ_OCC((interp_kernel<8, 2, true>)); else BLZ_OCC((interp_kernel<8, 2, false>)); }
    else if (s == 16 && r == 1) { if (compact) BLZ_OCC((interp_kernel<16, 1, true>)); else BLZ_OCC((interp_kernel<16, 1, false>)); }
    else if (s == 32 && r == 1) { if (compact) BLZ_OCC((interp_kernel<32, 1, true>)); else BLZ_OCC((interp_kernel<32, 1, false>)); }
    else if (s == 64 && r == 1) { if (compact) BLZ_OCC((interp_kernel<64, 1, true>)); else BLZ_OCC((interp_kernel<64, 1, false>)); }
  }
#undef BLZ_OCC
  return (e == hipSuccess && blocks > 0) ? blocks : 1;
}

hipError_t launch_gather(int width, const void *in, const uint8_t *in_valid, int64_t n_in, const int32_t *idx,
                         int64_t n_idx, void *out, uint8_t *out_valid, unsigned long long *null_count,
                         int grid, hipStream_t st) {
  switch (width) {
    case 1:
      hipLaunchKernelGGL(gather_kernel<uint8_t>, dim3(grid), dim3(kBlock), 0, st, (const uint8_t *)in,
                         in_valid, n_in, idx, n_idx, (uint8_t *)out, out_valid, null_count);
      break;
    case 2:
      hipLaunchKernelGGL(gather_kernel<uint16_t>, dim3(grid), dim3(kBlock), 0, st, (const uint16_t *)in,
                         in_valid, n_in, idx, n_idx, (uint16_t *)out, out_valid, null_count);
      break;
    case 4:
      hipLaunchKernelGGL(gather_kernel<uint32_t>, dim3(grid), dim3(kBlock), 0, st, (const uint32_t *)in,
                         in_valid, n_in, idx, n_idx, (uint32_t *)out, out_valid, null_count);
      break;
    case 8:
      hipLaunchKernelGGL(gather_kernel<uint64_t>, dim3(grid), dim3(kBlock), 0, st, (const uint64_t *)in,
                         in_valid, n_in, idx, n_idx, (uint64_t *)out, out_valid, null_count);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_sequence(int32_t *out, int64_t n, int32_t init, int32_t step, int grid, hipStream_t st) {
  hipLaunchKernelGGL(sequence_kernel, dim3(grid), dim3(kBlock), 0, st, out, n, init, step);
  return hipGetLastError();
}

hipError_t launch_fill_zero(void *p, int64_t bytes, int grid, hipStream_t st) {
  hipLaunchKernelGGL(fill_zero_kernel, dim3(grid), dim3(kBlock), 0, st, (uint64_t *)p, (bytes + 7) / 8);
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

hipError_t launch_copy(void *dst, const void *src, int64_t bytes, int grid, hipStream_t st) {
  hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(kBlock), 0, st, (uint4 *)dst, (const uint4 *)src, bytes / 16);
  return hipGetLastError();
}

int blob_capacity() { return kLdsBlobBytes; }
int block_threads() { return kBlock; }

}  // namespace blz

/*
 * vm_device.h — the tile VM: the evaluator of src/Interpreter
 * (transformKernel + InterpreterFunctor, interpreter_ops.cuh:1272-1458) and the
 * fused filter of CalciteInterpreter.cpp:509-573, laid out for CDNA4.
 *
 * A workgroup (4 wave64s) owns a tile of T = 256 R rows; wave w owns rows
 * [64 R w, 64 R (w + 1)) of it and lane l holds rows 64 R w + 64 u + l (u < R), so
 * every 64-row group is one validity word and one ballot.
 *
 * Staging. The columns an op reads are moved HBM -> LDS with direct-to-LDS
 * 16-byte loads (no VGPR round trip) into one tile buffer per workgroup
 * (kTileBuffers): a workgroup's loads and its evaluation alternate, and the other
 * workgroups resident on the CU overlap them. Only the last, partial tile of a
 * table is staged with bounded ordinary loads.
 *
 * Evaluation. The plan is in accumulator form (every op combines the running
 * value with at most one operand) pre-decoded on the host into 32-byte VmOps read
 * through the scalar cache, so the per-op cost is one uniform switch amortised
 * over 64 R rows. Operands come from the LDS tile, an immediate or an LDS
 * temporary. Ops outside the hot set run out of line, row by row.
 *
 * Filter (three launches, no inter-workgroup waiting inside a tile loop). The
 * predicate kernel writes one keep word per 64-row group (ballot) and the kept
 * count of every (tile, wave) slot; one scan kernel turns the counts into output
 * offsets with a single-pass decoupled look-back across workgroups; the gather
 * kernel packs each tile's kept rows densely into lanes (output j is the tile's
 * j-th kept row, found by a search over the keep words), stages just those rows
 * of the columns the projection reads, evaluates the projection and stores the
 * outputs and their validity words in row order.
 */
#ifndef BLZ_VM_DEVICE_H
#define BLZ_VM_DEVICE_H
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "program.h"
#include "kernels.h"
#include "device_common.h"

namespace blz {

#define LDSAS __attribute__((address_space(3)))
typedef LDSAS uint8_t *lds_ptr;
typedef const LDSAS uint8_t *clds_ptr;

namespace {

// This wave's share of a program's tile staging (pieces w, w + 4, ...), read from
// the piece table once per program into vector registers (lane i holds the i-th
// piece of the wave), so that staging a tile costs no scalar loads: each piece's
// fields are read across lanes.
struct WaveDma {
  int np;           // pieces of this wave (< 64)
  uint32_t src_lo;  // lane i: slice base + offset of the piece in the slice
  uint32_t src_hi;
  int slice;        // bytes per full tile
  int dst;          // LDS offset (low 20 bits) | bytes << 20
};
DEV void wave_dma_load(WaveDma &d, const CONSTANT Piece *pieces, int np) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  d.np = np > wave ? (np - wave + kWaves - 1) / kWaves : 0;
  d.src_lo = d.src_hi = 0;
  d.slice = d.dst = 0;
  if (lane < d.np) {
    const CONSTANT Piece &pc = pieces[wave + kWaves * lane];
    const uint64_t src = (uint64_t)pc.base + (uint64_t)(int64_t)pc.off;
    d.src_lo = (uint32_t)src;
    d.src_hi = (uint32_t)(src >> 32);
    d.slice = pc.slice;
    d.dst = pc.lds | (pc.bytes << 20);
  }
}
DEV void wave_dma_issue(const WaveDma &d, int64_t tile, lds_ptr tb) {
  const int lane = threadIdx.x & 63;
  for (int i = 0; i < d.np; i++) {
    const uint64_t base = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)d.src_lo, i) |
                          ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)d.src_hi, i) << 32);
    const int slice = __builtin_amdgcn_readlane(d.slice, i);
    const int dst = __builtin_amdgcn_readlane(d.dst, i);
    const uint8_t *src = (const uint8_t *)(base + (uint64_t)(tile * (int64_t)slice)) + lane * 16;
    if (lane * 16 < (dst >> 20))
      __builtin_amdgcn_global_load_lds((const GLOBAL void *)src, (LDSAS void *)(tb + (dst & 0xFFFFF)), 16, 0, 0);
  }
}

// Bounded staging of the last tile (rows < T): whole 16-byte chunks where the
// column has them, bytes at its very end, zeros past it. Never reads beyond
// ceil(rows * width) bytes of a slice (ceil(rows / 8) for a bitmap).
DEV void stage_partial(const CONSTANT Piece *pieces, int np, int64_t tile, int64_t rows, lds_ptr tb) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int j = wave; j < np; j += kWaves) {
    const CONSTANT Piece &p = pieces[j];
    if (lane * 16 >= p.bytes) continue;
    const int64_t have = p.bitmap ? (rows + 7) >> 3 : rows * p.width;
    const int64_t off = (int64_t)p.off + lane * 16;
    const int64_t lim64 = have - off;
    const int lim = lim64 < 0 ? 0 : lim64 > 16 ? 16 : (int)lim64;
    const uint8_t *src = (const uint8_t *)p.base + tile * (int64_t)p.slice + off;
    u32x4 v = u32x4{0u, 0u, 0u, 0u};
    if (lim == 16) {
      v = gload<u32x4>(src, 0);
    } else if (lim > 0) {
      uint32_t w[4] = {0, 0, 0, 0};
      for (int b = 0; b < lim; b++) w[b >> 2] |= (uint32_t)gload<uint8_t>(src, b) << (8 * (b & 3));
      v = u32x4{w[0], w[1], w[2], w[3]};
    }
    *(LDSAS u32x4 *)(tb + p.lds + lane * 16) = v;
  }
}

// Raw element `i` of a column of `width` bytes (no widening; staged as is).
DEV uint64_t load_raw(int width, const void *base, int64_t i) {
  switch (width) {
    case 8: return gload<uint64_t>(base, i);
    case 4: return gload<uint32_t>(base, i);
    case 2: return gload<uint16_t>(base, i);
    default: return gload<uint8_t>(base, i);
  }
}
// The same, non-temporal (rows read once: the gather of a filter's kept rows).
DEV uint64_t load_raw_nt(int width, const void *base, int64_t i) {
  switch (width) {
    case 8: return __builtin_nontemporal_load((const GLOBAL uint64_t *)base + i);
    case 4: return __builtin_nontemporal_load((const GLOBAL uint32_t *)base + i);
    case 2: return __builtin_nontemporal_load((const GLOBAL uint16_t *)base + i);
    default: return __builtin_nontemporal_load((const GLOBAL uint8_t *)base + i);
  }
}
DEV void store_raw(int width, lds_ptr p, uint64_t v) {
  switch (width) {
    case 8: *(LDSAS uint64_t *)p = v; break;
    case 4: *(LDSAS uint32_t *)p = (uint32_t)v; break;
    case 2: *(LDSAS uint16_t *)p = (uint16_t)v; break;
    default: *p = (uint8_t)v; break;
  }
}

// Gather staging: the kept rows grow[u] of every column the projection reads,
// into this lane's rows of the pass buffer (row plrow + 64 u). Up to four
// columns' loads are in flight before the first LDS store waits for them.
template <int R>
DEV void stage_kept(const CONSTANT Piece *gcols, int ng, lds_ptr gb, int plrow, const int64_t (&grow)[R],
                    const bool (&keep)[R]) {
  for (int g0 = 0; g0 < ng; g0 += 4) {
    uint64_t v[4][R];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (g0 + j < ng) {
        const int w = gcols[g0 + j].width;
        const void *base = gcols[g0 + j].base;
#pragma unroll
        for (int u = 0; u < R; u++) v[j][u] = keep[u] ? load_raw_nt(w, base, grow[u]) : 0;
      }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (g0 + j < ng) {
        const int w = gcols[g0 + j].width;
        lds_ptr d = gb + gcols[g0 + j].lds + w * plrow;
#pragma unroll
        for (int u = 0; u < R; u++) store_raw(w, d + w * 64 * u, v[j][u]);
      }
    }
  }
}

// Bare-column outputs: rows grow[u] copied, raw, to outputs obase + orel[u] (the
// kept rows of a filter, or every row of a projection); four outputs' loads in
// flight at a time.
template <int R>
DEV void copy_kept(const CONSTANT CopyCol *cc, int nc, const int64_t (&grow)[R], int64_t obase,
                   const int (&orel)[R], const bool (&live)[R]) {
  for (int g0 = 0; g0 < nc; g0 += 4) {
    uint64_t v[4][R];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (g0 + j < nc) {
        const int w = cc[g0 + j].width;
        const void *src = cc[g0 + j].src;
#pragma unroll
        for (int u = 0; u < R; u++) v[j][u] = live[u] ? load_raw(w, src, grow[u]) : 0;
      }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (g0 + j < nc) {
        void *dst = cc[g0 + j].dst;
        switch (cc[g0 + j].width) {
          case 8:
#pragma unroll
            for (int u = 0; u < R; u++) if (live[u]) gstore<uint64_t>(dst, obase + orel[u], v[j][u]);
            break;
          case 4:
#pragma unroll
            for (int u = 0; u < R; u++) if (live[u]) gstore<uint32_t>(dst, obase + orel[u], (uint32_t)v[j][u]);
            break;
          case 2:
#pragma unroll
            for (int u = 0; u < R; u++) if (live[u]) gstore<uint16_t>(dst, obase + orel[u], (uint16_t)v[j][u]);
            break;
          default:
#pragma unroll
            for (int u = 0; u < R; u++) if (live[u]) gstore<uint8_t>(dst, obase + orel[u], (uint8_t)v[j][u]);
            break;
        }
      }
    }
  }
}

// Operand rows (lane + 64 u) of the op, widened to 8-byte slots.
template <int R>
DEV void fetch(int cls, int off, int arg, int arg2, uint64_t im, clds_ptr tb, int lrow,
               clds_ptr tmpb, int tstride, const CONSTANT InCol *in, const int64_t (&grow)[R],
               const bool (&live)[R], uint64_t (&B)[R]) {
  switch (cls) {
    case VC_NONE:
#pragma unroll
      for (int u = 0; u < R; u++) B[u] = 0;
      break;
    case VC_IMM:
#pragma unroll
      for (int u = 0; u < R; u++) B[u] = im;
      break;
    case VC_TMP: {
      clds_ptr p = tmpb + off;
#pragma unroll
      for (int u = 0; u < R; u++) B[u] = *(const LDSAS uint64_t *)(p + 512 * u);
    } break;
    case VC_L64: {
      clds_ptr p = tb + off + 8 * lrow;
#pragma unroll
      for (int u = 0; u < R; u++) B[u] = *(const LDSAS uint64_t *)(p + 512 * u);
    } break;
    case VC_L32I: {
      clds_ptr p = tb + off + 4 * lrow;
#pragma unroll
      for (int u = 0; u < R; u++) B[u] = (uint64_t)(int64_t)*(const LDSAS int32_t *)(p + 256 * u);
    } break;
    case VC_L32F: {
      clds_ptr p = tb + off + 4 * lrow;
#pragma unroll
      for (int u = 0; u < R; u++) B[u] = d2b((double)*(const LDSAS float *)(p + 256 * u));
    } break;
    case VC_L16I: {
      clds_ptr p = tb + off + 2 * lrow;
#pragma unroll
      for (int u = 0; u < R; u++) B[u] = (uint64_t)(int64_t)*(const LDSAS int16_t *)(p + 128 * u);
    } break;
    case VC_L8I: {
      clds_ptr p = tb + off + lrow;
#pragma unroll
      for (int u = 0; u < R; u++) B[u] = (uint64_t)(int64_t)*(const LDSAS int8_t *)(p + 64 * u);
    } break;
    case VC_L8U: {
      clds_ptr p = tb + off + lrow;
#pragma unroll
      for (int u = 0; u < R; u++) B[u] = (uint64_t)*(const LDSAS uint8_t *)(p + 64 * u);
    } break;
    default:
#pragma unroll
      for (int u = 0; u < R; u++) B[u] = 0;
      break;
  }
}

// acc = acc CMP B (cmp_and = false) or acc &= (B CMP imm) (cmp_and = true).
template <int R>
DEV void compare(int sub, uint64_t (&acc)[R], const uint64_t (&B)[R]) {
#define BLZ_VCMP(CODE, EXPR)                        \
  case CODE:                                        \
    _Pragma("unroll") for (int u = 0; u < R; u++) { \
      const uint64_t a = acc[u], b = B[u];          \
      acc[u] = (uint64_t)(EXPR);                    \
    }                                               \
    break;
  switch (sub) {
    BLZ_VCMP(OP_LT_I, (int64_t)a < (int64_t)b)
    BLZ_VCMP(OP_LE_I, (int64_t)a <= (int64_t)b)
    BLZ_VCMP(OP_GT_I, (int64_t)a > (int64_t)b)
    BLZ_VCMP(OP_GE_I, (int64_t)a >= (int64_t)b)
    BLZ_VCMP(OP_EQ_I, a == b)
    BLZ_VCMP(OP_NE_I, a != b)
    BLZ_VCMP(OP_LT_F, b2d(a) < b2d(b))
    BLZ_VCMP(OP_LE_F, b2d(a) <= b2d(b))
    BLZ_VCMP(OP_GT_F, b2d(a) > b2d(b))
    BLZ_VCMP(OP_GE_F, b2d(a) >= b2d(b))
    BLZ_VCMP(OP_EQ_F, b2d(a) == b2d(b))
    BLZ_VCMP(OP_NE_F, b2d(a) != b2d(b))
    default: break;
  }
#undef BLZ_VCMP
}

template <int R>
DEV void compare_and(int sub, uint64_t (&acc)[R], const uint64_t (&B)[R], uint64_t im) {
  const double fi = b2d(im);
  const int64_t ii = (int64_t)im;
#define BLZ_VCMPAND(CODE, EXPR)                     \
  case CODE:                                        \
    _Pragma("unroll") for (int u = 0; u < R; u++) { \
      const uint64_t b = B[u];                      \
      acc[u] = (EXPR) ? acc[u] : 0ull;              \
    }                                               \
    break;
  switch (sub) {
    BLZ_VCMPAND(OP_LT_I, (int64_t)b < ii)
    BLZ_VCMPAND(OP_LE_I, (int64_t)b <= ii)
    BLZ_VCMPAND(OP_GT_I, (int64_t)b > ii)
    BLZ_VCMPAND(OP_GE_I, (int64_t)b >= ii)
    BLZ_VCMPAND(OP_EQ_I, b == im)
    BLZ_VCMPAND(OP_NE_I, b != im)
    BLZ_VCMPAND(OP_LT_F, b2d(b) < fi)
    BLZ_VCMPAND(OP_LE_F, b2d(b) <= fi)
    BLZ_VCMPAND(OP_GT_F, b2d(b) > fi)
    BLZ_VCMPAND(OP_GE_F, b2d(b) >= fi)
    BLZ_VCMPAND(OP_EQ_F, b2d(b) == fi)
    BLZ_VCMPAND(OP_NE_F, b2d(b) != fi)
    default: break;
  }
#undef BLZ_VCMPAND
}

// One op outside the hot set for one row (interpreter_ops.cuh:473-838 semantics
// with the host-resolved conversions).
__device__ __attribute__((noinline)) uint64_t vm_generic_row(int sub, int gen, uint64_t acc, uint64_t b,
                                                             bool vb) {
  const int cb = gen & 3, ca = (gen >> 2) & 3, co = (gen >> 4) & 3;
  if (cb) b = conv(b, cb);
  if (ca) acc = conv(acc, ca);
  uint64_t x[1], y[1], z[1];
  const bool v[1] = {vb};
  if (gen & VG_UNARY) { x[0] = acc; y[0] = acc; }
  else if (gen & VG_REV) { x[0] = b; y[0] = acc; }
  else { x[0] = acc; y[0] = b; }
  exec_op<1>(sub, x, y, v, z);
  return co ? conv(z[0], co) : z[0];
}

DEV uint64_t vm_idiv(uint64_t a, uint64_t b) { return (uint64_t)idiv((int64_t)a, (int64_t)b); }
DEV uint64_t vm_imod(uint64_t a, uint64_t b) { return (uint64_t)imod((int64_t)a, (int64_t)b); }

// Raw validity bits of input `vcol` for rows grow + 64 u (COALESCE, IS_[NOT_]NULL:
// interpreter_ops.cuh:633-645, :820-829 read the left input's bitmap).
template <int R>
DEV void vcol_bits(const CONSTANT InCol *in, int vcol, const int64_t (&grow)[R], int64_t n, bool (&vb)[R]) {
  const void *vp = vcol != 0xFF ? in[vcol].valid : nullptr;
#pragma unroll
  for (int u = 0; u < R; u++) {
    const int64_t g = grow[u];
    vb[u] = valid_bit(vp, g < n ? g : n - 1);
  }
}

// The ops outside the common set (see run_vm).
template <int R, int OPS>
DEV void run_rare(int code, int arg, uint32_t w3, uint64_t (&acc)[R], const uint64_t (&B)[R],
                  const CONSTANT InCol *in, const int64_t (&grow)[R], int64_t n) {
    switch (code) {
      case VM_F2I:
#pragma unroll
        for (int u = 0; u < R; u++) acc[u] = (uint64_t)f2i(b2d(acc[u]));
        break;
      case VM_FLOOR:
#pragma unroll
        for (int u = 0; u < R; u++) acc[u] = d2b(__builtin_floor(b2d(acc[u])));
        break;
      case VM_CEIL:
#pragma unroll
        for (int u = 0; u < R; u++) acc[u] = d2b(__builtin_ceil(b2d(acc[u])));
        break;
      case VM_ABS_F:
#pragma unroll
        for (int u = 0; u < R; u++) acc[u] = acc[u] & 0x7FFFFFFFFFFFFFFFull;
        break;
      case VM_NOT_I:
#pragma unroll
        for (int u = 0; u < R; u++) acc[u] = (uint64_t)(acc[u] == 0);
        break;
      case VM_NOT_F:
#pragma unroll
        for (int u = 0; u < R; u++) acc[u] = (uint64_t)(!(b2d(acc[u]) != 0.0));
        break;
      case VM_DIV_F:
#pragma unroll
        for (int u = 0; u < R; u++) acc[u] = d2b(b2d(acc[u]) / b2d(B[u]));
        break;
      case VM_RDIV_F:
#pragma unroll
        for (int u = 0; u < R; u++) acc[u] = d2b(b2d(B[u]) / b2d(acc[u]));
        break;
      case VM_DIV_I:
#pragma nounroll
        for (int u = 0; u < R; u++) acc[u] = vm_idiv(acc[u], B[u]);
        break;
      case VM_RDIV_I:
#pragma nounroll
        for (int u = 0; u < R; u++) acc[u] = vm_idiv(B[u], acc[u]);
        break;
      case VM_MOD_I:
#pragma nounroll
        for (int u = 0; u < R; u++) acc[u] = vm_imod(acc[u], B[u]);
        break;
      case VM_RMOD_I:
#pragma nounroll
        for (int u = 0; u < R; u++) acc[u] = vm_imod(B[u], acc[u]);
        break;
      case VM_COALESCE:
      case VM_RCOALESCE: {
        bool vb[R];
        vcol_bits<R>(in, (int)((w3 >> 8) & 0xFF), grow, n, vb);
        const bool rev = code == VM_RCOALESCE;
#pragma unroll
        for (int u = 0; u < R; u++) {
          const uint64_t x = rev ? B[u] : acc[u], y = rev ? acc[u] : B[u];
          acc[u] = vb[u] ? x : y;
        }
      } break;
      case VM_IS_NULL:
      case VM_IS_NOT_NULL: {
        bool vb[R];
        vcol_bits<R>(in, (int)((w3 >> 8) & 0xFF), grow, n, vb);
        const bool want = code == VM_IS_NOT_NULL;
#pragma unroll
        for (int u = 0; u < R; u++) acc[u] = (uint64_t)(vb[u] == want);
      } break;
      default:
        if constexpr (OPS >= 2) {  // VM_GENERIC
          const int gen = (int)(w3 & 0xFF), vcol = (int)((w3 >> 8) & 0xFF);
          const void *vp = vcol != 0xFF ? in[vcol].valid : nullptr;
#pragma nounroll
          for (int u = 0; u < R; u++) {
            const int64_t g = grow[u];
            const bool vb = valid_bit(vp, g < n ? g : n - 1);
            acc[u] = vm_generic_row(arg, gen, acc[u], B[u], vb);
          }
        }
        break;
    }
}

// Runs `nops` VmOps over rows lrow + 64 u of the tile buffer `tb` (global rows
// grow[u]). Temporaries live in LDS: the temporary at byte offset t of row u is at
// tmpb + t + 512 u. VM_EMIT stores row u at output position obase + orel[u] when
// live[u].
//
// Dispatch: one uniform switch over the op's leaf id. A hot leaf (a common op with
// an operand of a common class) is straight-line code: the operand is read where
// it is used, so no operand value crosses a join and the accumulator is updated in
// place. The rest decode their fields in the general leaf. Scalar work per op is
// the switch's compare tree; the next op's descriptor is loaded while this one
// executes. Accumulator moves (a load before an op, a store or an output after
// it) ride on the op beside them as flags (kFusePre, kFuseStore, kFuseEmit), so
// they cost no dispatch of their own.
//
// OPS selects what the general leaf can run: 0 the common set (any operand class
// and conversion of the common ops); 1 adds the rest of the hot set (division,
// modulo, floor / ceil / abs, NOT, binary64 -> int64, COALESCE, IS [NOT] NULL); 2
// adds the out-of-line set (VM_GENERIC: transcendentals, POWER, date parts).
#define BLZ_HOT_OPS(X)                                                                         \
  X(HO_LOAD, acc[u] = b)                                                                       \
  X(HO_ADD_I, acc[u] = acc[u] + b)                                                             \
  X(HO_ADD_F, acc[u] = d2b(b2d(acc[u]) + b2d(b)))                                              \
  X(HO_SUB_I, acc[u] = acc[u] - b)                                                             \
  X(HO_SUB_F, acc[u] = d2b(b2d(acc[u]) - b2d(b)))                                              \
  X(HO_RSUB_I, acc[u] = b - acc[u])                                                            \
  X(HO_RSUB_F, acc[u] = d2b(b2d(b) - b2d(acc[u])))                                             \
  X(HO_MUL_I, acc[u] = acc[u] * b)                                                             \
  X(HO_MUL_F, acc[u] = d2b(b2d(acc[u]) * b2d(b)))                                              \
  X(HO_AND_B, acc[u] = acc[u] & b)                                                             \
  X(HO_CMP + 0, acc[u] = (uint64_t)((int64_t)acc[u] < (int64_t)b))                             \
  X(HO_CMP + 1, acc[u] = (uint64_t)((int64_t)acc[u] <= (int64_t)b))                            \
  X(HO_CMP + 2, acc[u] = (uint64_t)((int64_t)acc[u] > (int64_t)b))                             \
  X(HO_CMP + 3, acc[u] = (uint64_t)((int64_t)acc[u] >= (int64_t)b))                            \
  X(HO_CMP + 4, acc[u] = (uint64_t)(acc[u] == b))                                              \
  X(HO_CMP + 5, acc[u] = (uint64_t)(acc[u] != b))                                              \
  X(HO_CMP + 6, acc[u] = (uint64_t)(b2d(acc[u]) < b2d(b)))                                     \
  X(HO_CMP + 7, acc[u] = (uint64_t)(b2d(acc[u]) <= b2d(b)))                                    \
  X(HO_CMP + 8, acc[u] = (uint64_t)(b2d(acc[u]) > b2d(b)))                                     \
  X(HO_CMP + 9, acc[u] = (uint64_t)(b2d(acc[u]) >= b2d(b)))                                    \
  X(HO_CMP + 10, acc[u] = (uint64_t)(b2d(acc[u]) == b2d(b)))                                   \
  X(HO_CMP + 11, acc[u] = (uint64_t)(b2d(acc[u]) != b2d(b)))                                   \
  X(HO_CMPAND + 0, acc[u] = (int64_t)b < ii ? acc[u] : 0ull)                                   \
  X(HO_CMPAND + 1, acc[u] = (int64_t)b <= ii ? acc[u] : 0ull)                                  \
  X(HO_CMPAND + 2, acc[u] = (int64_t)b > ii ? acc[u] : 0ull)                                   \
  X(HO_CMPAND + 3, acc[u] = (int64_t)b >= ii ? acc[u] : 0ull)                                  \
  X(HO_CMPAND + 4, acc[u] = b == im2 ? acc[u] : 0ull)                                          \
  X(HO_CMPAND + 5, acc[u] = b != im2 ? acc[u] : 0ull)                                          \
  X(HO_CMPAND + 6, acc[u] = b2d(b) < fi ? acc[u] : 0ull)                                       \
  X(HO_CMPAND + 7, acc[u] = b2d(b) <= fi ? acc[u] : 0ull)                                      \
  X(HO_CMPAND + 8, acc[u] = b2d(b) > fi ? acc[u] : 0ull)                                       \
  X(HO_CMPAND + 9, acc[u] = b2d(b) >= fi ? acc[u] : 0ull)                                      \
  X(HO_CMPAND + 10, acc[u] = b2d(b) == fi ? acc[u] : 0ull)                                     \
  X(HO_CMPAND + 11, acc[u] = b2d(b) != fi ? acc[u] : 0ull)                                     \
  X(HO_RANGE_I, acc[u] = ((int64_t)b > ii && (int64_t)b < (int64_t)im) ? acc[u] : 0ull)         \
  X(HO_RANGE_F, acc[u] = ((b2d(b) > fi || ((rf & 1) && b2d(b) == fi)) &&                       \
                          (b2d(b) < hf || ((rf & 2) && b2d(b) == hf))) ? acc[u] : 0ull)
// Operand of row u by hot class.
#define BLZ_B_IMM im
#define BLZ_B_TMP (*(const LDSAS uint64_t *)(tp + 512 * u))
#define BLZ_B_L64 (*(const LDSAS uint64_t *)(lp8 + 512 * u))
#define BLZ_B_L32I ((uint64_t)(int64_t)*(const LDSAS int32_t *)(lp4 + 256 * u))
#define BLZ_B_L32F d2b((double)*(const LDSAS float *)(lp4 + 256 * u))
#define BLZ_B_L32I_F d2b((double)*(const LDSAS int32_t *)(lp4 + 256 * u))
#define BLZ_B_L64_F d2b((double)*(const LDSAS int64_t *)(lp8 + 512 * u))
#define BLZ_LEAF(OP, CLS, FETCH, STMT)                  \
  case (OP) * HC_COUNT_ + (CLS): {                      \
    _Pragma("unroll") for (int u = 0; u < R; u++) {     \
      const uint64_t b = FETCH;                         \
      STMT;                                             \
    }                                                   \
  } break;
#define BLZ_LEAVES(OP, STMT)                            \
  BLZ_LEAF(OP, HC_IMM, BLZ_B_IMM, STMT)                 \
  BLZ_LEAF(OP, HC_TMP, BLZ_B_TMP, STMT)                 \
  BLZ_LEAF(OP, HC_L64, BLZ_B_L64, STMT)                 \
  BLZ_LEAF(OP, HC_L32I, BLZ_B_L32I, STMT)               \
  BLZ_LEAF(OP, HC_L32F, BLZ_B_L32F, STMT)               \
  BLZ_LEAF(OP, HC_L32I_F, BLZ_B_L32I_F, STMT)           \
  BLZ_LEAF(OP, HC_L64_F, BLZ_B_L64_F, STMT)

// Output `arg` <- acc (StoreKind `kind`): rows u with live[u] at obase + orel[u].
template <int R>
DEV void vm_emit(const CONSTANT OutCol *out, int arg, int kind, const uint64_t (&acc)[R], int64_t obase,
                 const int (&orel)[R], const bool (&live)[R]) {
  void *base = out[arg].data;
  switch (kind) {
    case ST_I64: case ST_F64:
#pragma unroll
      for (int u = 0; u < R; u++) if (live[u]) gstore_nt<uint64_t>(base, obase + orel[u], acc[u]);
      break;
    case ST_I32:
#pragma unroll
      for (int u = 0; u < R; u++) if (live[u]) gstore_nt<int32_t>(base, obase + orel[u], (int32_t)acc[u]);
      break;
    case ST_F32:
#pragma unroll
      for (int u = 0; u < R; u++) if (live[u]) gstore_nt<float>(base, obase + orel[u], (float)b2d(acc[u]));
      break;
    case ST_I16:
#pragma unroll
      for (int u = 0; u < R; u++) if (live[u]) gstore_nt<int16_t>(base, obase + orel[u], (int16_t)acc[u]);
      break;
    case ST_I8:
#pragma unroll
      for (int u = 0; u < R; u++) if (live[u]) gstore_nt<int8_t>(base, obase + orel[u], (int8_t)acc[u]);
      break;
    default: break;
  }
}

template <int R, int OPS>
DEV void run_vm(const CONSTANT VmOp *ops, int nops, uint64_t (&acc)[R], lds_ptr tmpb, int tstride,
                const CONSTANT InCol *in, const CONSTANT OutCol *out, clds_ptr tb, int lrow,
                const int64_t (&grow)[R], int64_t n, int64_t obase, const int (&orel)[R],
                const bool (&live)[R]) {
  typedef uint32_t u32x8 __attribute__((ext_vector_type(8)));
  static_assert(sizeof(VmOp) == sizeof(u32x8), "VmOp is one 32-byte descriptor");
  const CONSTANT u32x8 *opw = (const CONSTANT u32x8 *)ops;
  // the op list is followed by one spare descriptor, so op k + 1 is always loadable
  u32x8 w = opw[0];
  clds_ptr b4 = tb + 4 * lrow;  // this lane's row in a 4- / 8-byte column slice
  clds_ptr b8 = tb + 8 * lrow;
  for (int k = 0; k < nops; k++) {
    const u32x8 w_next = opw[k + 1];
    const uint32_t s0 = w.s0;
    const int leaf = (int)(s0 & kLeafMask);
    const int off = (int)w.s1;
    if (s0 & kFusePre) {  // acc = operand (as HO_LOAD), then the op
      const int poff = (int)(w.s2 & 0xFFFFF);
      clds_ptr tp = (clds_ptr)tmpb + poff;
      clds_ptr lp4 = b4 + poff, lp8 = b8 + poff;
      switch ((w.s2 >> 20) & 0xF) {
        case HC_TMP:
#pragma unroll
          for (int u = 0; u < R; u++) acc[u] = BLZ_B_TMP;
          break;
        case HC_L64:
#pragma unroll
          for (int u = 0; u < R; u++) acc[u] = BLZ_B_L64;
          break;
        case HC_L32I:
#pragma unroll
          for (int u = 0; u < R; u++) acc[u] = BLZ_B_L32I;
          break;
        case HC_L32F:
#pragma unroll
          for (int u = 0; u < R; u++) acc[u] = BLZ_B_L32F;
          break;
        case HC_L32I_F:
#pragma unroll
          for (int u = 0; u < R; u++) acc[u] = BLZ_B_L32I_F;
          break;
        default:  // HC_L64_F
#pragma unroll
          for (int u = 0; u < R; u++) acc[u] = BLZ_B_L64_F;
          break;
      }
    }
    const uint64_t im = (uint64_t)w.s4 | ((uint64_t)w.s5 << 32);
    const uint64_t im2 = (uint64_t)w.s6 | ((uint64_t)w.s7 << 32);
    const double fi = b2d(im2), hf = b2d(im);
    const int64_t ii = (int64_t)im2;
    const int rf = (int)(w.s3 & 0xFF);  // VM_RANGE_F: bound inclusivity
    clds_ptr tp = (clds_ptr)tmpb + off;
    clds_ptr lp4 = b4 + off, lp8 = b8 + off;
    switch (leaf) {
      BLZ_HOT_OPS(BLZ_LEAVES)
      case kLeafStore: {
        lds_ptr p = tmpb + off;
#pragma unroll
        for (int u = 0; u < R; u++) *(LDSAS uint64_t *)(p + 512 * u) = acc[u];
      } break;
      case kLeafEmit:
        vm_emit<R>(out, (int)((w.s2 >> 8) & 0xFF), (int)(w.s3 & 0xFF), acc, obase, orel, live);
        break;
      case kLeafI2F:
#pragma unroll
        for (int u = 0; u < R; u++) acc[u] = d2b((double)(int64_t)acc[u]);
        break;
      case kLeafNotI:
#pragma unroll
        for (int u = 0; u < R; u++) acc[u] = (uint64_t)(acc[u] == 0);
        break;
      case kLeafNotF:
#pragma unroll
        for (int u = 0; u < R; u++) acc[u] = (uint64_t)(!(b2d(acc[u]) != 0.0));
        break;
      case kLeafDivK:
      case kLeafModK: {
        // truncating division by the constant im (magic multiplier im2, shift in gen)
        const int64_t d = (int64_t)im, M = (int64_t)im2;
        const int sh = (int)(w.s3 & 0xFF);
        const bool add_n = d > 0 && M < 0, sub_n = d < 0 && M > 0;
        const bool div = leaf == kLeafDivK;
#pragma unroll
        for (int u = 0; u < R; u++) {
          const int64_t x = (int64_t)acc[u];
          uint64_t q = (uint64_t)__mul64hi(M, x);
          if (add_n) q += (uint64_t)x;
          if (sub_n) q -= (uint64_t)x;
          int64_t qs = (int64_t)q >> sh;
          qs += (int64_t)((uint64_t)qs >> 63);
          acc[u] = div ? (uint64_t)qs : (uint64_t)x - (uint64_t)qs * (uint64_t)d;
        }
      } break;
      default: {  // general leaf
        const int code = (int)(w.s2 & 0xFF), arg = (int)((w.s2 >> 8) & 0xFF);
        const int cls = (int)((w.s2 >> 16) & 0xFF), cv = (int)(w.s2 >> 24);
        const int arg2 = (int)((w.s3 >> 16) & 0xFF);
        uint64_t B[R];
        fetch<R>(cls, off, arg, arg2, im, tb, lrow, tmpb, tstride, in, grow, live, B);
        if (cv == CV_I2F) {
#pragma unroll
          for (int u = 0; u < R; u++) B[u] = d2b((double)(int64_t)B[u]);
        } else if (cv == CV_F2I) {
#pragma unroll
          for (int u = 0; u < R; u++) B[u] = (uint64_t)f2i(b2d(B[u]));
        }
        switch (code) {
          case VM_LOAD:
#pragma unroll
            for (int u = 0; u < R; u++) acc[u] = B[u];
            break;
          case VM_ADD_I:
#pragma unroll
            for (int u = 0; u < R; u++) acc[u] = acc[u] + B[u];
            break;
          case VM_ADD_F:
#pragma unroll
            for (int u = 0; u < R; u++) acc[u] = d2b(b2d(acc[u]) + b2d(B[u]));
            break;
          case VM_SUB_I:
#pragma unroll
            for (int u = 0; u < R; u++) acc[u] = acc[u] - B[u];
            break;
          case VM_SUB_F:
#pragma unroll
            for (int u = 0; u < R; u++) acc[u] = d2b(b2d(acc[u]) - b2d(B[u]));
            break;
          case VM_RSUB_I:
#pragma unroll
            for (int u = 0; u < R; u++) acc[u] = B[u] - acc[u];
            break;
          case VM_RSUB_F:
#pragma unroll
            for (int u = 0; u < R; u++) acc[u] = d2b(b2d(B[u]) - b2d(acc[u]));
            break;
          case VM_MUL_I:
#pragma unroll
            for (int u = 0; u < R; u++) acc[u] = acc[u] * B[u];
            break;
          case VM_MUL_F:
#pragma unroll
            for (int u = 0; u < R; u++) acc[u] = d2b(b2d(acc[u]) * b2d(B[u]));
            break;
          case VM_AND_B:
#pragma unroll
            for (int u = 0; u < R; u++) acc[u] = acc[u] & B[u];
            break;
          case VM_CMP:
            compare<R>(arg, acc, B);
            break;
          case VM_CMPAND:
            compare_and<R>(arg, acc, B, im2);
            break;
          case VM_RANGE_I:
#pragma unroll
            for (int u = 0; u < R; u++) acc[u] = ((int64_t)B[u] > ii && (int64_t)B[u] < (int64_t)im) ? acc[u] : 0ull;
            break;
          case VM_RANGE_F:
#pragma unroll
            for (int u = 0; u < R; u++) {
              const double x = b2d(B[u]);
              acc[u] = ((x > fi || ((rf & 1) && x == fi)) && (x < hf || ((rf & 2) && x == hf))) ? acc[u] : 0ull;
            }
            break;
          default:
            if constexpr (OPS >= 1) run_rare<R, OPS>(code, arg, w.s3, acc, B, in, grow, n);
            break;
        }
      } break;
    }
    if (s0 & (kFuseStore | kFuseEmit)) {
      if (s0 & kFuseStore) {
        lds_ptr p = tmpb + (int)(w.s2 >> 24) * tstride;
#pragma unroll
        for (int u = 0; u < R; u++) *(LDSAS uint64_t *)(p + 512 * u) = acc[u];
      }
      if (s0 & kFuseEmit) vm_emit<R>(out, (int)((w.s3 >> 8) & 0xFF), (int)((w.s3 >> 16) & 0xFF), acc, obase, orel, live);
    }
    w = w_next;
  }
}
#undef BLZ_LEAVES
#undef BLZ_LEAF
#undef BLZ_HOT_OPS

// Validity registers live in LDS: word u (64-row group u of the pass) of register
// s of a wave is vr[s * R + u]. Lane u < R computes group u; a register is read
// back by any lane at a wave-uniform address. LDS operations of one wave complete
// in order, so a lane reads what another lane of its wave wrote before it.

// Input validity words of a pass from the tile buffer (groups group0 .. + R).
template <int R>
DEV void vm_load_valid(const CONSTANT InCol *in, int n_in, clds_ptr tb, int group0, LDSAS uint64_t *vr) {
  const int lane = threadIdx.x & 63;
  for (int c = 0; c < n_in; c++) {
    const int vs = in[c].vslot;
    if (vs >= 0 && lane < R) vr[vs * R + lane] = *(const LDSAS uint64_t *)(tb + in[c].lds_voff + 8 * (group0 + lane));
  }
  __builtin_amdgcn_wave_barrier();
}

// The validity program (AND / OR of 64-bit words).
template <int R>
DEV void vm_run_vops(const CONSTANT VOp *vops, int nvops, LDSAS uint64_t *vr) {
  const int lane = threadIdx.x & 63;
  for (int k = 0; k < nvops; k++) {
    const uint32_t w = *(const CONSTANT uint32_t *)(vops + k);
    const int code = w & 0xFF, d = (w >> 8) & 0xFF, a = (w >> 16) & 0xFF, b = (w >> 24) & 0xFF;
    if (lane < R) {
      const uint64_t x = vr[a * R + lane], y = vr[b * R + lane];
      vr[d * R + lane] = code == VOP_OR ? (x | y) : (x & y);
    }
  }
  __builtin_amdgcn_wave_barrier();
}

// Word of group u of a validity source: a register, all valid (-1) or none (-2).
template <int R>
DEV uint64_t vm_vword(int vsrc, const LDSAS uint64_t *vr, int u) {
  if (vsrc >= 0) return vr[vsrc * R + u];
  return vsrc == -2 ? 0ull : ~0ull;
}

}  // namespace

// Kernel modes.
//   VM_PROJECT  evaluate the projection of every row (no filter);
//   VM_PRED     evaluate the predicate: one keep word per 64-row group and one
//               kept-row count per wave slot (tile, wave);
//   VM_GATHER   from the kept rows' indices in output order (written by the slot
//               scan), evaluate the projection of the kept rows only, packed densely
//               into lanes, and store them, and their validity words, in row order.
enum VmMode { VM_PROJECT = 0, VM_PRED = 1, VM_GATHER = 2 };

// Position of the r-th (0-based) set bit of m (r < popcount(m)).
DEV int select_bit(uint64_t m, int r) {
  int p = 0;
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) {
    const int lo = __builtin_popcountll(m & ((1ull << s) - 1));
    if (r >= lo) {
      r -= lo;
      m >>= s;
      p += s;
    }
  }
  return p;
}

// Program of a global id: largest i with first(i) <= id (first(0) == 0).
template <typename F>
DEV int find_plan(int n_plans, int64_t id, F first) {
  int lo = 0, hi = n_plans - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first(mid) <= id) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// A program's workspace pointers, read through the constant cache.
DEV Workspace load_ws(const CONSTANT Workspace &w) {
  Workspace r;
  r.null_counts = w.null_counts;
  r.tile_counter = w.tile_counter;
  r.tile_state = w.tile_state;
  r.n_sel = w.n_sel;
  r.keep = w.keep;
  r.wave_count = w.wave_count;
  r.idx = w.idx;
  return r;
}

// The fields of one program the kernels read (scalar loads through the constant
// cache, once per program change).
struct VmCtx {
  const CONSTANT InCol *in;
  const CONSTANT OutCol *out;
  const CONSTANT VmOp *ops;
  const CONSTANT VOp *vops;
  const CONSTANT Piece *pieces;
  const CONSTANT Piece *gcols;
  const CONSTANT CopyCol *copies;
  int64_t n;
  int n_in, n_out, n_ops_p, n_ops_q, n_vops_p, n_vops_q, any_valid_out, np, n_gcols, n_copy, T, n_vout,
      st_width, st_vsrc;
};
DEV VmCtx load_ctx(const uint8_t *blob_arg) {
  const CONSTANT uint8_t *blob = (const CONSTANT uint8_t *)blob_arg;
  const CONSTANT ProgHeader &h = *(const CONSTANT ProgHeader *)blob;
  VmCtx c;
  c.in = (const CONSTANT InCol *)(blob + h.off_in);
  c.out = (const CONSTANT OutCol *)(blob + h.off_out);
  c.ops = (const CONSTANT VmOp *)(blob + h.off_ops);
  c.vops = (const CONSTANT VOp *)(blob + h.off_vops);
  c.pieces = (const CONSTANT Piece *)(blob + h.off_pieces);
  c.gcols = (const CONSTANT Piece *)(blob + h.off_gcols);
  c.copies = (const CONSTANT CopyCol *)(blob + h.off_copy);
  c.n = h.n_rows;
  c.n_in = h.n_in;
  c.n_out = h.n_out;
  c.n_ops_p = h.n_ops_p;
  c.n_ops_q = h.n_ops_q;
  c.n_vops_p = h.n_vops_p;
  c.n_vops_q = h.n_vops_q;
  c.any_valid_out = h.any_valid_out;
  c.np = h.n_pieces;
  c.n_gcols = h.n_gcols;
  c.n_copy = h.n_copy;
  c.T = h.tile_rows;
  c.n_vout = h.n_valid_out;
  c.st_width = h.stencil_width;
  c.st_vsrc = h.stencil_vsrc;
  return c;
}

// Minimum waves per SIMD the register allocator must allow (launch bounds): with
// the common op set, R <= 4 programs fit 8 waves (projection, predicate) or 4
// (gather) without spills, which hides more of the staging and gather latency.
constexpr int vm_min_waves(int r, int mode, int ops) {
  return ops != 0 || r > 4 ? 1 : mode == 2 ? 4 : 8;
}

template <int R, int MODE, int OPS>
__global__ void __launch_bounds__(kBlock, vm_min_waves(R, MODE, OPS))
vm_kernel(const VmBatch *__restrict__ bd_arg) {
  extern __shared__ __attribute__((aligned(16))) uint8_t lds_raw[];
  constexpr bool STAGED = MODE != VM_GATHER;

  lds_ptr lds = (lds_ptr)lds_raw;
  const CONSTANT VmBatch &bd = *(const CONSTANT VmBatch *)bd_arg;
  const CONSTANT VmPlanRef *plans = (const CONSTANT VmPlanRef *)bd.plans;
  const int n_plans = bd.n_plans;
  const int bufb = bd.bufb;

  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  // Temporaries: LDS rows of this lane (one pass of the workgroup: 256 R rows).
  const int tstride = 8 * kBlock * R;
  lds_ptr tmpb = lds + (MODE == VM_GATHER ? bd.g_tmp_off : bd.tmp_off) + 8 * (wave * 64 * R + lane);

  // Every register below is written by the program before it is read (the host
  // compiler guarantees it); none is live across passes.
  uint64_t acc[R];
  // this wave's validity registers
  LDSAS uint64_t *vr = (LDSAS uint64_t *)(lds + (MODE == VM_GATHER ? bd.g_vr_off : bd.vr_off)) + wave * bd.vr_wave_words;

  if (STAGED) {
    const int64_t n_tiles = bd.n_tiles;
    auto tile_first = [&](int i) { return plans[i].tile0; };
    auto tile_end = [&](int i) { return i + 1 < n_plans ? plans[i + 1].tile0 : n_tiles; };
    int64_t cur = blockIdx.x;
    // The tiles a workgroup visits increase, so the program of its current tile,
    // and of the tile being staged, change only at program boundaries.
    int64_t p_first = 0, p_end = -1;  // global tiles of the current program
    VmCtx cx;
    unsigned long long *null_counts = nullptr;  // the current program's workspace
    // nulls of output `lane` (< 64) counted by this wave since the program began,
    // added to the program's counts once, when the wave leaves it
    unsigned long long nacc = 0;
    uint64_t *keep = nullptr;
    uint32_t *wave_count = nullptr;
    int64_t d_first = 0, d_end = -1, d_full = 0;  // the staged program: its tiles, full tiles
    WaveDma dma;
    int buf = 0;
    // staging of global tile g (full tiles only; the last tile of a program is
    // staged with bounded loads when it is current)
    auto issue = [&](int64_t g, lds_ptr dst) {
      if (g >= n_tiles) return;
      if (g >= d_end) {
        const int q = find_plan(n_plans, g, tile_first);
        const CONSTANT uint8_t *qb = (const CONSTANT uint8_t *)plans[q].blob;
        const CONSTANT ProgHeader &qh = *(const CONSTANT ProgHeader *)qb;
        d_first = plans[q].tile0;
        d_end = tile_end(q);
        d_full = qh.n_rows / qh.tile_rows;
        wave_dma_load(dma, (const CONSTANT Piece *)(qb + qh.off_pieces), qh.n_pieces);
      }
      const int64_t t = g - d_first;
      if (t < d_full) wave_dma_issue(dma, t, dst);
    };
    if (kTileBuffers == 2) issue(cur, lds);
    for (; cur < n_tiles; cur += gridDim.x, buf ^= (kTileBuffers == 2)) {
      if (cur >= p_end) {
        if (MODE == VM_PROJECT && null_counts != nullptr && nacc) atomicAdd(null_counts + lane, nacc);
        nacc = 0;
        const int p = find_plan(n_plans, cur, tile_first);
        p_first = plans[p].tile0;
        p_end = tile_end(p);
        cx = load_ctx(plans[p].blob);
        const Workspace ws = load_ws(plans[p].ws);
        null_counts = ws.null_counts;
        keep = ws.keep;
        wave_count = ws.wave_count;
      }
      const int64_t tile = cur - p_first;
      const int T = cx.T;
      const int P = T / (kBlock * R);  // passes per tile
      const int GW = P * R;            // 64-row groups per wave slot
      const int lbw = wave * 64 * GW;  // first row of this wave's slot in the tile
      const int64_t n = cx.n;
      const int64_t row0 = tile * T;
      const int64_t rows = (n - row0) < T ? (n - row0) : T;
      lds_ptr tb = lds + buf * bufb;
      if (kTileBuffers == 1) issue(cur, tb);
      // This wave's staging loads of the tile are complete; the barrier makes every
      // wave's visible and ends all reads of the other buffer.
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (rows < T) {
        stage_partial(cx.pieces, cx.np, tile, rows, tb);
        __syncthreads();
      }
      if (kTileBuffers == 2) issue(cur + gridDim.x, lds + (buf ^ 1) * bufb);

      if (MODE == VM_PROJECT) {
        for (int ps = 0; ps < P; ps++) {
          const int lb = lbw + ps * 64 * R;
          if (lb >= rows) break;  // wave-uniform
          const int lrow = lb + lane;
          int orel[R];
          bool put[R];
          int64_t grow[R];
#pragma unroll
          for (int u = 0; u < R; u++) {
            orel[u] = lrow + 64 * u;
            put[u] = orel[u] < rows;
            grow[u] = row0 + orel[u];
            acc[u] = 0;
          }
          run_vm<R, OPS>(cx.ops, cx.n_ops_q, acc, tmpb, tstride, cx.in, cx.out, tb, lrow, grow, n, row0, orel,
                          put);
          if (cx.n_copy) copy_kept<R>(cx.copies, cx.n_copy, grow, row0, orel, put);
          if (cx.any_valid_out) {
            vm_load_valid<R>(cx.in, cx.n_in, tb, lb >> 6, vr);
            vm_run_vops<R>(cx.vops, cx.n_vops_q, vr);
            for (int o = 0; o < cx.n_out; o++) {
              void *vp = cx.out[o].valid;
              if (vp == nullptr) continue;
              const uint64_t word = vm_vword<R>(cx.out[o].vsrc, vr, lane < R ? lane : 0);
              const int64_t wi = ((row0 + lb) >> 6) + lane;
              uint64_t nulls = 0;
              if (lane < R && wi * 64 < n) {
                store_vword(vp, wi, n, word);
                nulls = __builtin_popcountll(~word & low_mask(n - wi * 64));
              }
              for (int s = 32; s > 0; s >>= 1) nulls += __shfl_xor(nulls, s);
              if (o < 64) nacc += lane == o ? nulls : 0;
              else if (lane == 0 && nulls) atomicAdd(null_counts + o, (unsigned long long)nulls);
            }
          }
        }
      } else {
        // ---- predicate -> keep word per 64-row group, kept rows per wave slot ----
        const uint64_t st_mask = low_mask(8 * cx.st_width);
        uint64_t myword = 0, cnt = 0;
        for (int ps = 0; ps < P; ps++) {
          const int lb = lbw + ps * 64 * R;
          if (lb >= rows) break;  // wave-uniform
          const int lrow = lb + lane;
          int orel[R];
          bool none[R];
          int64_t grow[R];
#pragma unroll
          for (int u = 0; u < R; u++) {
            orel[u] = 0;
            none[u] = false;
            grow[u] = row0 + lrow + 64 * u;
            acc[u] = 1;  // a predicate may open with VM_CMPAND (conjunction of comparisons)
          }
          run_vm<R, OPS>(cx.ops, cx.n_ops_p, acc, tmpb, tstride, cx.in, cx.out, tb, lrow, grow, n, 0, orel, none);
          if (cx.st_vsrc >= 0) {  // the predicate's validity is computed
            vm_load_valid<R>(cx.in, cx.n_in, tb, lb >> 6, vr);
            vm_run_vops<R>(cx.vops, cx.n_vops_p, vr);
          }
#pragma unroll
          for (int u = 0; u < R; u++) {
            const bool vbit = (vm_vword<R>(cx.st_vsrc, vr, u) >> lane) & 1;
            const bool kp = lrow + 64 * u < rows && ((acc[u] & st_mask) != 0) && vbit;
            const uint64_t km = __builtin_amdgcn_ballot_w64(kp);
            if (lane == ps * R + u) myword = km;
            cnt += __builtin_popcountll(km);
          }
        }
        const int64_t g0 = tile * (T / 64) + wave * GW;  // this wave's first group
        if (lane < GW) gstore<uint64_t>(keep, g0 + lane, myword);
        if (lane == 0) gstore<uint32_t>(wave_count, tile * kWaves + wave, (uint32_t)cnt);
      }
      if (kTileBuffers == 1) __syncthreads();  // every wave is done with the buffer
    }
    if (MODE == VM_PROJECT && null_counts != nullptr && nacc) atomicAdd(null_counts + lane, nacc);
    return;
  }

  // ---- VM_GATHER: output-driven. The kept rows of all programs, in output order,
  // are cut into chunks of 64 R outputs; wave chunk c of the batch belongs to the
  // program whose chunk range holds c (per-program chunk counts from n_sel, their
  // prefix computed once per workgroup into LDS). Lane l's rows are outputs
  // j0 + 64 u + l, whose source rows are idx[j] (written by the slot scan). Chunks
  // start at multiples of 64 outputs, so each validity word of an output is
  // written by exactly one wave, whole.
  constexpr int CR = 64 * R;  // outputs per chunk
  LDSAS int64_t *cpre = (LDSAS int64_t *)(lds + bd.g_cpre_off);  // [n_plans + 1] chunk prefix
  if (wave == 0) {
    int64_t run = 0;
    for (int i0 = 0; i0 < n_plans; i0 += 64) {
      const int i = i0 + lane;
      int64_t c = 0;
      if (i < n_plans) {
        const Workspace w = load_ws(plans[i].ws);
        c = ((int64_t)*(volatile unsigned long long *)w.n_sel + CR - 1) / CR;
      }
      int64_t inc = c;
      for (int off = 1; off < 64; off <<= 1) {
        const int64_t y = __shfl_up(inc, off);
        if (lane >= off) inc += y;
      }
      if (i < n_plans) cpre[i] = run + inc - c;
      run += __shfl(inc, 63);
    }
    if (lane == 0) cpre[n_plans] = run;
  }
  __syncthreads();
  const int64_t total_chunks = cpre[n_plans];
  lds_ptr gbuf = lds + bd.g_buf_off + wave * bd.g_wave_bytes;  // this wave's pass buffer
  const int plrow = lane;  // this lane's row in the pass buffer
  int64_t p_first = 0, p_end = -1;  // chunks of the current program
  VmCtx cx;
  const uint32_t *idx = nullptr;
  int64_t n_sel = 0;
  // nulls of output `lane` (< 64) counted by this wave since it entered the
  // current program, added to the program's counts when the wave leaves it
  unsigned long long *null_counts = nullptr;
  unsigned long long nacc = 0;
  // the next chunk's source rows are loaded while this one is gathered (when it
  // belongs to the same program)
  const int64_t cstride = (int64_t)gridDim.x * kWaves;
  uint32_t nidx[R];
  bool have_next = false;
  for (int64_t ch = (int64_t)blockIdx.x * kWaves + wave; ch < total_chunks; ch += cstride) {
    if (ch >= p_end) {
      have_next = false;
      if (null_counts != nullptr && nacc) atomicAdd(null_counts + lane, nacc);
      nacc = 0;
      int lo = 0, hi = n_plans - 1;  // the program with cpre[p] <= ch < cpre[p + 1]
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cpre[mid] <= ch) lo = mid;
        else hi = mid - 1;
      }
      p_first = cpre[lo];
      p_end = cpre[lo + 1];
      cx = load_ctx(plans[lo].blob);
      const Workspace w = load_ws(plans[lo].ws);
      null_counts = w.null_counts;
      idx = w.idx;
      n_sel = (int64_t)*(volatile unsigned long long *)w.n_sel;
    }
    const int64_t j0 = (ch - p_first) * CR;  // first output of the chunk
    const int64_t n = cx.n;
    int orel[R];
    bool live[R];
    int64_t grow[R];
#pragma unroll
    for (int u = 0; u < R; u++) {
      orel[u] = 64 * u + lane;
      live[u] = j0 + orel[u] < n_sel;
    }
#pragma unroll
    for (int u = 0; u < R; u++)
      grow[u] = !live[u] ? 0 : have_next ? (int64_t)nidx[u] : (int64_t)__builtin_nontemporal_load(idx + j0 + orel[u]);
    have_next = ch + cstride < p_end;
    if (have_next) {
      const int64_t jn = j0 + cstride * CR;
#pragma unroll
      for (int u = 0; u < R; u++) nidx[u] = jn + orel[u] < n_sel ? __builtin_nontemporal_load(idx + jn + orel[u]) : 0u;
    }
#pragma unroll
    for (int u = 0; u < R; u++) acc[u] = 0;
    stage_kept<R>(cx.gcols, cx.n_gcols, gbuf, plrow, grow, live);
    run_vm<R, OPS>(cx.ops + cx.n_ops_p, cx.n_ops_q, acc, tmpb, tstride, cx.in, cx.out, gbuf, plrow, grow, n, j0,
                    orel, live);
    if (cx.n_copy) copy_kept<R>(cx.copies, cx.n_copy, grow, j0, orel, live);
    if (cx.any_valid_out) {
      // input validity of the chunk's rows: word u (outputs j0 + 64 u ..) to lane 0;
      // a column's R bit loads are all issued before the first ballot waits
      // (rows of lanes that are not live are row 0: in range, masked after)
      for (int ci = 0; ci < cx.n_in; ci++) {
        const int vs = cx.in[ci].vslot;
        if (vs < 0) continue;
        const void *vp = cx.in[ci].valid;
        bool vb[R];
#pragma unroll
        for (int u = 0; u < R; u++) vb[u] = valid_bit(vp, grow[u]);
#pragma unroll
        for (int u = 0; u < R; u++) {
          const uint64_t word = __builtin_amdgcn_ballot_w64(live[u] & vb[u]);
          if (lane == 0) vr[vs * R + u] = word;
        }
      }
      __builtin_amdgcn_wave_barrier();
      // The projection's validity program may reuse words the predicate's
      // computed (shared subexpressions): run both.
      vm_run_vops<R>(cx.vops, cx.n_vops_p + cx.n_vops_q, vr);
      for (int o = 0; o < cx.n_out; o++) {
        void *vp = cx.out[o].valid;
        if (vp == nullptr) continue;
        const int vsrc = cx.out[o].vsrc;
#pragma unroll
        for (int u = 0; u < R; u++) {
          const int64_t jw = j0 + 64 * u;  // first output of word u
          if (jw >= n_sel) break;          // wave-uniform
          const uint64_t lm = low_mask(n_sel - jw);
          const uint64_t bits = vm_vword<R>(vsrc, vr, u) & lm;
          const unsigned long long nn = __builtin_popcountll(~bits & lm);
          if (o < 64) nacc += lane == o ? nn : 0;
          else if (lane == 0) atomicAdd(null_counts + o, nn);
          if (lane == u) gstore<uint64_t>(vp, jw >> 6, bits);
        }
      }
    }
  }
  if (null_counts != nullptr && nacc) atomicAdd(null_counts + lane, nacc);
}

// ---------------------------------------------------------------------------
// launch templates

template <int RR, int MODE>
hipError_t launch_vm_mode(int ops, int grid, int lds_bytes, hipStream_t st, const VmBatch *bd) {
  if (ops >= 2) hipLaunchKernelGGL((vm_kernel<RR, MODE, 2>), dim3(grid), dim3(kBlock), lds_bytes, st, bd);
  else if (ops == 1) hipLaunchKernelGGL((vm_kernel<RR, MODE, 1>), dim3(grid), dim3(kBlock), lds_bytes, st, bd);
  else hipLaunchKernelGGL((vm_kernel<RR, MODE, 0>), dim3(grid), dim3(kBlock), lds_bytes, st, bd);
  return hipGetLastError();
}

template <int RR, int MODE>
int vm_occupancy_mode(int ops, int lds_bytes) {
  int b = 0;
  if (ops >= 2) (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, vm_kernel<RR, MODE, 2>, kBlock, lds_bytes);
  else if (ops == 1) (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, vm_kernel<RR, MODE, 1>, kBlock, lds_bytes);
  else (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, vm_kernel<RR, MODE, 0>, kBlock, lds_bytes);
  return b;
}

template <int RR>
hipError_t launch_vm_shape(int mode, int slow, int grid, int lds_bytes, hipStream_t st, const VmBatch *bd) {
  switch (mode) {
    case VM_PROJECT: return launch_vm_mode<RR, VM_PROJECT>(slow, grid, lds_bytes, st, bd);
    case VM_PRED: return launch_vm_mode<RR, VM_PRED>(slow, grid, lds_bytes, st, bd);
    case VM_GATHER: return launch_vm_mode<RR, VM_GATHER>(slow, grid, lds_bytes, st, bd);
    default: return hipErrorInvalidValue;
  }
}

template <int RR>
int vm_occupancy(int mode, int slow, int lds_bytes) {
  switch (mode) {
    case VM_PROJECT: return vm_occupancy_mode<RR, VM_PROJECT>(slow, lds_bytes);
    case VM_PRED: return vm_occupancy_mode<RR, VM_PRED>(slow, lds_bytes);
    case VM_GATHER: return vm_occupancy_mode<RR, VM_GATHER>(slow, lds_bytes);
    default: return 0;
  }
}


// Per-shape entry points: vm_r2.hip, vm_r4.hip and vm_r8.hip each instantiate the
// kernels of one rows-per-lane value (separate translation units compile in
// parallel).
#define BLZ_VM_SHAPE_ENTRY(RR)                                                                    \
  hipError_t launch_vm_r##RR(int mode, int slow, int grid, int lds_bytes, hipStream_t st,         \
                             const VmBatch *bd) {                                                 \
    return launch_vm_shape<RR>(mode, slow, grid, lds_bytes, st, bd);                              \
  }                                                                                               \
  int vm_occupancy_r##RR(int mode, int slow, int lds_bytes) { return vm_occupancy<RR>(mode, slow, lds_bytes); }

}  // namespace blz
#endif

/*
 * vm_kernel.hip — the tile VM: the evaluator of src/Interpreter
 * (transformKernel + InterpreterFunctor, interpreter_ops.cuh:1272-1458) and the
 * fused filter of CalciteInterpreter.cpp:509-573, laid out for CDNA4.
 *
 * A workgroup (4 wave64s) owns a tile of T = 256 R rows; wave w owns rows
 * [64 R w, 64 R (w + 1)) of it and lane l holds rows 64 R w + 64 u + l (u < R), so
 * every 64-row group is one validity word and one ballot.
 *
 * Staging. The columns an op reads are moved HBM -> LDS with direct-to-LDS
 * 16-byte loads (no VGPR round trip), double buffered: while the waves evaluate
 * tile i from one buffer, the loads of their next tile fill the other. Only the
 * last, partial tile of a table is staged with bounded ordinary loads.
 *
 * Evaluation. The plan is in accumulator form (every op combines the running
 * value with at most one operand) pre-decoded on the host into 32-byte VmOps read
 * through the scalar cache, so the per-op cost is one uniform switch amortised
 * over 64 R rows. Operands come from the LDS tile, an immediate, one of NT
 * temporaries or, in the projection pass of a filter, straight from HBM for the
 * rows the filter keeps (columns the predicate does not read are never staged).
 * Ops outside the hot set run out of line, row by row.
 *
 * Filter. Pass A evaluates the predicate into per-group keep masks (ballot);
 * the tile's count is combined across workgroups by a single-pass decoupled
 * look-back over dynamically assigned tile ids; pass B evaluates the projection
 * and scatters the kept rows, in row order, and their validity bits.
 */
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

// Direct HBM -> LDS staging of full tile `tile` into the buffer at `tb`: wave w
// moves pieces w, w + 4, ...; lane l carries bytes [16 l, 16 l + 16) of a piece.
DEV void dma_tile(const CONSTANT Piece *pieces, int np, int64_t tile, lds_ptr tb) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int j = wave; j < np; j += kWaves) {
    const CONSTANT Piece &p = pieces[j];
    const int slice = p.slice, off = p.off, dst = p.lds, bytes = p.bytes;
    const uint8_t *src = (const uint8_t *)p.base + tile * (int64_t)slice + off + lane * 16;
    if (lane * 16 < bytes)
      __builtin_amdgcn_global_load_lds((const GLOBAL void *)src, (LDSAS void *)(tb + dst), 16, 0, 0);
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

// Operand rows (lane + 64 u) of the op, widened to 8-byte slots.
template <int R, int NT>
DEV void fetch(int cls, int off, int arg, int arg2, const CONSTANT VmOp &op, clds_ptr tb, int lrow,
               const uint64_t (&tmp)[NT][R], const CONSTANT InCol *in, int64_t grow, const bool (&live)[R],
               uint64_t (&B)[R]) {
  switch (cls) {
    case VC_IMM: {
      const uint64_t im = op.imm;
#pragma unroll
      for (int u = 0; u < R; u++) B[u] = im;
    } break;
    case VC_TMP:
#pragma unroll
      for (int t = 0; t < NT; t++)
        if (t == arg2) {
#pragma unroll
          for (int u = 0; u < R; u++) B[u] = tmp[t][u];
        }
      break;
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
    default: {
      // VC_G*: the column in HBM, read only for live rows.
      const void *base = in[arg2].data;
#pragma unroll
      for (int u = 0; u < R; u++) {
        uint64_t v = 0;
        if (live[u]) {
          const int64_t g = grow + 64 * u;
          switch (cls) {
            case VC_G64: v = gload<uint64_t>(base, g); break;
            case VC_G32I: v = (uint64_t)(int64_t)gload<int32_t>(base, g); break;
            case VC_G32F: v = d2b((double)gload<float>(base, g)); break;
            case VC_G16I: v = (uint64_t)(int64_t)gload<int16_t>(base, g); break;
            case VC_G8I: v = (uint64_t)(int64_t)gload<int8_t>(base, g); break;
            case VC_G8U: v = (uint64_t)gload<uint8_t>(base, g); break;
            default: break;
          }
        }
        B[u] = v;
      }
    } break;
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
DEV void vcol_bits(const CONSTANT InCol *in, int vcol, int64_t grow, int64_t n, bool (&vb)[R]) {
  const void *vp = vcol != 0xFF ? in[vcol].valid : nullptr;
#pragma unroll
  for (int u = 0; u < R; u++) {
    const int64_t g = grow + 64 * u;
    vb[u] = valid_bit(vp, g < n ? g : n - 1);
  }
}

// Runs `nops` VmOps over rows lrow + 64 u of the tile buffer `tb` (global row
// grow + 64 u). VM_EMIT stores row u at output position obase + orel[u] when
// live[u]. SLOW: the program has ops of the out-of-line set (VM_GENERIC).
template <int R, int NT, bool SLOW>
DEV void run_vm(const CONSTANT VmOp *ops, int nops, uint64_t (&acc)[R], uint64_t (&tmp)[NT][R],
                const CONSTANT InCol *in, const CONSTANT OutCol *out, clds_ptr tb, int lrow, int64_t grow,
                int64_t n, int64_t obase, const int (&orel)[R], const bool (&live)[R]) {
  for (int k = 0; k < nops; k++) {
    const CONSTANT VmOp &op = ops[k];
    const u32x4 w = *(const CONSTANT u32x4 *)&op;
    const int code = (int)(w.x & 0xFF), arg = (int)((w.x >> 8) & 0xFF);
    const int cls = (int)((w.x >> 16) & 0xFF), cv = (int)(w.x >> 24);
    switch (code) {
      case VM_STORE:
#pragma unroll
        for (int t = 0; t < NT; t++)
          if (t == arg) {
#pragma unroll
            for (int u = 0; u < R; u++) tmp[t][u] = acc[u];
          }
        continue;
      case VM_EMIT: {
        void *base = out[arg].data;
        switch (cls) {
          case ST_I64: case ST_F64:
#pragma unroll
            for (int u = 0; u < R; u++) if (live[u]) gstore<uint64_t>(base, obase + orel[u], acc[u]);
            break;
          case ST_I32:
#pragma unroll
            for (int u = 0; u < R; u++) if (live[u]) gstore<int32_t>(base, obase + orel[u], (int32_t)acc[u]);
            break;
          case ST_F32:
#pragma unroll
            for (int u = 0; u < R; u++) if (live[u]) gstore<float>(base, obase + orel[u], (float)b2d(acc[u]));
            break;
          case ST_I16:
#pragma unroll
            for (int u = 0; u < R; u++) if (live[u]) gstore<int16_t>(base, obase + orel[u], (int16_t)acc[u]);
            break;
          case ST_I8:
#pragma unroll
            for (int u = 0; u < R; u++) if (live[u]) gstore<int8_t>(base, obase + orel[u], (int8_t)acc[u]);
            break;
          default: break;
        }
      }
        continue;
      case VM_I2F:
#pragma unroll
        for (int u = 0; u < R; u++) acc[u] = d2b((double)(int64_t)acc[u]);
        continue;
      case VM_F2I:
#pragma unroll
        for (int u = 0; u < R; u++) acc[u] = (uint64_t)f2i(b2d(acc[u]));
        continue;
      case VM_FLOOR:
#pragma unroll
        for (int u = 0; u < R; u++) acc[u] = d2b(__builtin_floor(b2d(acc[u])));
        continue;
      case VM_CEIL:
#pragma unroll
        for (int u = 0; u < R; u++) acc[u] = d2b(__builtin_ceil(b2d(acc[u])));
        continue;
      case VM_ABS_F:
#pragma unroll
        for (int u = 0; u < R; u++) acc[u] = acc[u] & 0x7FFFFFFFFFFFFFFFull;
        continue;
      case VM_NOT_I:
#pragma unroll
        for (int u = 0; u < R; u++) acc[u] = (uint64_t)(acc[u] == 0);
        continue;
      case VM_NOT_F:
#pragma unroll
        for (int u = 0; u < R; u++) acc[u] = (uint64_t)(!(b2d(acc[u]) != 0.0));
        continue;
      case VM_IS_NULL:
      case VM_IS_NOT_NULL: {
        bool vb[R];
        vcol_bits<R>(in, (int)((w.y >> 8) & 0xFF), grow, n, vb);
        const bool want = code == VM_IS_NOT_NULL;
#pragma unroll
        for (int u = 0; u < R; u++) acc[u] = (uint64_t)(vb[u] == want);
      }
        continue;
      default:
        break;
    }
    const int off = (int)w.z;
    const int arg2 = (int)((w.y >> 16) & 0xFF);
    uint64_t B[R];
    fetch<R, NT>(cls, off, arg, arg2, op, tb, lrow, tmp, in, grow, live, B);
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
        compare_and<R>(arg, acc, B, op.imm2);
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
        vcol_bits<R>(in, (int)((w.y >> 8) & 0xFF), grow, n, vb);
        const bool rev = code == VM_RCOALESCE;
#pragma unroll
        for (int u = 0; u < R; u++) {
          const uint64_t x = rev ? B[u] : acc[u], y = rev ? acc[u] : B[u];
          acc[u] = vb[u] ? x : y;
        }
      } break;
      default:
        if (SLOW) {  // VM_GENERIC
          const int gen = (int)(w.y & 0xFF), vcol = (int)((w.y >> 8) & 0xFF);
          const void *vp = vcol != 0xFF ? in[vcol].valid : nullptr;
#pragma nounroll
          for (int u = 0; u < R; u++) {
            const int64_t g = grow + 64 * u;
            const bool vb = valid_bit(vp, g < n ? g : n - 1);
            acc[u] = vm_generic_row(arg, gen, acc[u], B[u], vb);
          }
        }
        break;
    }
  }
}

// Validity words of a pass from the tile buffer: lane u < R loads group u's word
// of every staged bitmap.
template <int VS, int R>
DEV void vm_read_valid(const CONSTANT InCol *in, int n_in, clds_ptr tb, int group0, uint64_t (&vrf)[VS]) {
  const int lane = threadIdx.x & 63;
  const int g = group0 + (lane < R ? lane : 0);
  for (int c = 0; c < n_in; c++) {
    const int vs = in[c].vslot;
    if (vs >= 0) vrf[vs] = *(const LDSAS uint64_t *)(tb + in[c].lds_voff + 8 * g);
  }
}

}  // namespace

// Kernel modes.
//   VM_PROJECT  evaluate the projection of every row (no filter);
//   VM_PRED     evaluate the predicate: one keep word per 64-row group and one
//               kept-row count per wave slot (tile, wave);
//   VM_GATHER   from the keep words and the exclusive prefix of the counts, evaluate
//               the projection of the kept rows only (operands read from HBM for those
//               rows) and store them, and their validity bits, in row order.
enum VmMode { VM_PROJECT = 0, VM_PRED = 1, VM_GATHER = 2 };

template <int R, int NT, int MODE, bool SLOW>
__global__ void __launch_bounds__(kBlock)
vm_kernel(const uint8_t *__restrict__ blob_arg, Workspace ws) {
  extern __shared__ __attribute__((aligned(16))) uint8_t lds_raw[];
  constexpr int VS = 8;  // validity registers
  constexpr int T = 256 * R;
  constexpr bool STAGED = MODE != VM_GATHER;

  lds_ptr lds = (lds_ptr)lds_raw;
  const CONSTANT uint8_t *blob = (const CONSTANT uint8_t *)blob_arg;
  const CONSTANT ProgHeader &h = *(const CONSTANT ProgHeader *)blob;
  const CONSTANT InCol *in = (const CONSTANT InCol *)(blob + h.off_in);
  const CONSTANT OutCol *out = (const CONSTANT OutCol *)(blob + h.off_out);
  const CONSTANT VmOp *ops = (const CONSTANT VmOp *)(blob + h.off_ops);
  const CONSTANT VOp *vops = (const CONSTANT VOp *)(blob + h.off_vops);
  const CONSTANT Piece *pieces = (const CONSTANT Piece *)(blob + h.off_pieces);
  const int64_t n = h.n_rows;
  const int n_in = h.n_in, n_out = h.n_out;
  const int n_ops_p = h.n_ops_p, n_ops_q = h.n_ops_q;
  const int n_vops_p = h.n_vops_p, n_vops_q = h.n_vops_q;
  const int any_valid_out = h.any_valid_out;
  const int np = h.n_pieces, bufb = h.buf_bytes;

  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int lb = wave * 64 * R;  // first row of this wave in the tile
  const int lrow = lb + lane;
  const int64_t n_tiles = (n + T - 1) / T;

  // Every register below is written by the program before it is read (the host
  // compiler guarantees it); none is live across tiles.
  uint64_t acc[R];
  uint64_t tmp[NT][R];
  uint64_t vrf[VS];

  int64_t cur = blockIdx.x;
  int buf = 0;
  if (STAGED && cur < n_tiles && (cur + 1) * T <= n) dma_tile(pieces, np, cur, lds);

  for (; cur < n_tiles; cur += gridDim.x, buf ^= 1) {
    const int64_t row0 = cur * T;
    const int64_t rows = (n - row0) < T ? (n - row0) : T;
    lds_ptr tb = lds + buf * bufb;
    if (STAGED) {
      // This wave's staging loads of the tile are complete; the barrier makes every
      // wave's visible and ends all reads of the other buffer.
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (rows < T) {
        stage_partial(pieces, np, cur, rows, tb);
        __syncthreads();
      }
      const int64_t nxt = cur + gridDim.x;
      if (nxt < n_tiles && (nxt + 1) * T <= n) dma_tile(pieces, np, nxt, lds + (buf ^ 1) * bufb);
    }

    if (MODE == VM_PROJECT) {
      if (lb >= rows) continue;  // wave-uniform
      int orel[R];
      bool put[R];
#pragma unroll
      for (int u = 0; u < R; u++) {
        orel[u] = lrow + 64 * u;
        put[u] = orel[u] < rows;
        acc[u] = 0;
      }
      run_vm<R, NT, SLOW>(ops, n_ops_q, acc, tmp, in, out, tb, lrow, row0 + lrow, n, row0, orel, put);
      if (any_valid_out) {
        vm_read_valid<VS, R>(in, n_in, tb, lb >> 6, vrf);
        run_vops<VS>(vops, n_vops_q, vrf);
        for (int o = 0; o < n_out; o++) {
          void *vp = out[o].valid;
          if (vp == nullptr) continue;
          const uint64_t word = vsrc_word<VS>(out[o].vsrc, vrf);
          const int64_t wi = ((row0 + lb) >> 6) + lane;
          uint64_t nulls = 0;
          if (lane < R && wi * 64 < n) {
            store_vword(vp, wi, n, word);
            nulls = __builtin_popcountll(~word & low_mask(n - wi * 64));
          }
          for (int s = 32; s > 0; s >>= 1) nulls += __shfl_xor(nulls, s);
          if (lane == 0 && nulls) atomicAdd(ws.null_counts + o, (unsigned long long)nulls);
        }
      }
      continue;
    }

    if (MODE == VM_PRED) {
      // ---- predicate -> keep word per 64-row group, kept rows per wave slot ----
      const uint64_t st_mask = low_mask(8 * h.stencil_width);
      uint64_t myword = 0, cnt = 0;
      if (lb < rows) {
        int orel[R];
        bool none[R];
#pragma unroll
        for (int u = 0; u < R; u++) {
          orel[u] = 0;
          none[u] = false;
          acc[u] = 0;
        }
        run_vm<R, NT, SLOW>(ops, n_ops_p, acc, tmp, in, out, tb, lrow, row0 + lrow, n, 0, orel, none);
        vm_read_valid<VS, R>(in, n_in, tb, lb >> 6, vrf);
        run_vops<VS>(vops, n_vops_p, vrf);
        const uint64_t vstencil = vsrc_word<VS>(h.stencil_vsrc, vrf);
#pragma unroll
        for (int u = 0; u < R; u++) {
          const bool vbit = (readlane64(vstencil, u) >> lane) & 1;
          const bool keep = lrow + 64 * u < rows && ((acc[u] & st_mask) != 0) && vbit;
          const uint64_t km = __builtin_amdgcn_ballot_w64(keep);
          if (lane == u) myword = km;
          cnt += __builtin_popcountll(km);
        }
      }
      const int64_t g0 = cur * (T / 64) + wave * R;  // this wave's first group
      if (lane < R) gstore<uint64_t>(ws.keep, g0 + lane, myword);
      if (lane == 0) gstore<uint32_t>(ws.wave_count, cur * kWaves + wave, (uint32_t)cnt);
      continue;
    }

    // ---- VM_GATHER: the kept rows of this wave slot, in row order ----
    if (lb >= rows) continue;
    const int64_t slot = cur * kWaves + wave;
    const uint32_t wcnt = ws.wave_count[slot];
    if (wcnt == 0) continue;  // wave-uniform
    const uint64_t obase = ws.wave_prefix[slot];
    const int64_t g0 = cur * (T / 64) + wave * R;
    int rel[R];
    bool keep[R];
    uint64_t run = 0;
#pragma unroll
    for (int u = 0; u < R; u++) {
      const uint64_t m = ((const CONSTANT uint64_t *)ws.keep)[g0 + u];
      const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                 __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
      rel[u] = (int)run + (int)below;
      keep[u] = (m >> lane) & 1;
      run += __builtin_popcountll(m);
      acc[u] = 0;
    }
    run_vm<R, NT, SLOW>(ops + n_ops_p, n_ops_q, acc, tmp, in, out, tb, lrow, row0 + lrow, n, (int64_t)obase,
                        rel, keep);
    if (any_valid_out) {
      // Validity words of the wave's groups straight from the input bitmaps.
      const int64_t gw = ((row0 + lb) >> 6) + (lane < R ? lane : 0);
      for (int c = 0; c < n_in; c++) {
        const int vs = in[c].vslot;
        if (vs >= 0) vrf[vs] = load_vword(in[c].valid, gw, n);
      }
      run_vops<VS>(vops + n_vops_p, n_vops_q, vrf);
      // The wave's kept rows are output positions [obase, obase + wcnt): their bits
      // are gathered in this wave's LDS words, then stored; the two end words may
      // be shared with the neighbouring wave slots and are merged atomically.
      constexpr int VW = R + 2;
      LDSAS uint64_t *vbits = (LDSAS uint64_t *)(lds + h.lds_vbits_off) + wave * VW;
      const int nw = (int)(((obase & 63) + wcnt + 63) >> 6);
      const uint64_t end = obase + wcnt;
      const int64_t w0 = (int64_t)(obase >> 6);
      for (int o = 0; o < n_out; o++) {
        void *vp = out[o].valid;
        if (vp == nullptr) continue;
        if (lane < VW) vbits[lane] = 0;
        __builtin_amdgcn_wave_barrier();
        const uint64_t vword = vsrc_word<VS>(out[o].vsrc, vrf);
        unsigned long long nl = 0;
#pragma unroll
        for (int u = 0; u < R; u++) {
          if (keep[u]) {
            if ((readlane64(vword, u) >> lane) & 1) {
              const uint64_t local = (obase & 63) + (uint64_t)rel[u];
              __hip_atomic_fetch_or(vbits + (local >> 6), 1ull << (local & 63), __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_WORKGROUP);
            } else {
              nl++;
            }
          }
        }
        __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): the wave's LDS atomics are done
        __builtin_amdgcn_wave_barrier();
        if (lane < nw) {
          unsigned long long *gp = (unsigned long long *)((uint8_t *)vp + 8 * (w0 + lane));
          const uint64_t word = vbits[lane];
          const bool shared = (lane == 0 && (obase & 63)) || (lane == nw - 1 && (end & 63));
          if (shared) atomicOr(gp, (unsigned long long)word);
          else gstore<uint64_t>(gp, 0, word);
        }
        for (int s = 32; s > 0; s >>= 1) nl += __shfl_xor(nl, s);
        if (lane == 0 && nl) atomicAdd(ws.null_counts + o, nl);
        __builtin_amdgcn_wave_barrier();
      }
    }
  }
}

// Exclusive prefix of the per-wave-slot kept counts (VM_PRED) in one pass:
// each workgroup scans 4096 counts and combines its total with the workgroups
// before it by a single-pass decoupled look-back. The last one writes the total.
constexpr int kScanItems = 16;
__global__ void __launch_bounds__(kBlock)
slot_scan_kernel(const uint32_t *__restrict__ counts, uint32_t *__restrict__ prefix, int64_t n_slots,
                 Workspace ws) {
  __shared__ uint64_t s_wave[kWaves];
  __shared__ uint64_t s_prefix;
  __shared__ int64_t s_block;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // Chunks are taken in the order workgroups start, so every chunk a look-back
  // waits for belongs to a workgroup that is already running.
  if (threadIdx.x == 0) s_block = (int64_t)atomicAdd(ws.tile_counter, 1u);
  __syncthreads();
  const int64_t b = s_block;
  const int64_t base = b * kBlock * kScanItems + (int64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint64_t sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    v[i] = base + i < n_slots ? counts[base + i] : 0u;
    sum += v[i];
  }
  // inclusive scan of the thread sums across the wave
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
    constexpr uint64_t FA = 1ull << 62, FP = 2ull << 62, VM = (1ull << 62) - 1;
    uint64_t pre = 0;
    if (b == 0) {
      if (lane == 0) __hip_atomic_store(ws.tile_state, FP | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      if (lane == 0)
        __hip_atomic_store(ws.tile_state + b, FA | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      int64_t look = b - 1;
      while (true) {
        const int64_t idx = look - lane;
        uint64_t st = FP;
        if (idx >= 0) st = __hip_atomic_load(ws.tile_state + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t flag = st >> 62;
        const uint64_t is_x = __builtin_amdgcn_ballot_w64(flag == 0);
        const uint64_t is_p = __builtin_amdgcn_ballot_w64(flag == 2);
        const uint64_t upto = is_p ? (((is_p & (~is_p + 1)) << 1) - 1) : ~0ull;
        if (is_x & upto) {
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        uint64_t x = ((upto >> lane) & 1) ? (st & VM) : 0;
        for (int s = 32; s > 0; s >>= 1) x += __shfl_xor(x, s);
        pre += x;
        if (is_p) break;
        look -= 64;
      }
      if (lane == 0)
        __hip_atomic_store(ws.tile_state + b, FP | (pre + total), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (lane == 0) {
      s_prefix = pre;
      if (b == (int64_t)gridDim.x - 1) *ws.n_sel = pre + total;
    }
  }
  __syncthreads();
  uint64_t run = s_prefix + wbase + inc - sum;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    if (base + i < n_slots) prefix[base + i] = (uint32_t)run;
    run += v[i];
  }
}

// ---------------------------------------------------------------------------
// launchers

#define BLZ_VM_SHAPES(X) X(8, 2) X(8, 4) X(4, 2) X(4, 4) X(4, 8) X(2, 4) X(2, 8)

template <int RR, int NN, int MODE>
hipError_t launch_vm_mode(bool slow, int grid, int lds_bytes, hipStream_t st, const uint8_t *blob, Workspace ws) {
  if (slow) hipLaunchKernelGGL((vm_kernel<RR, NN, MODE, true>), dim3(grid), dim3(kBlock), lds_bytes, st, blob, ws);
  else hipLaunchKernelGGL((vm_kernel<RR, NN, MODE, false>), dim3(grid), dim3(kBlock), lds_bytes, st, blob, ws);
  return hipGetLastError();
}

template <int RR, int NN, int MODE>
int vm_occupancy_mode(bool slow, int lds_bytes) {
  int b = 0;
  if (slow) (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, vm_kernel<RR, NN, MODE, true>, kBlock, lds_bytes);
  else (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, vm_kernel<RR, NN, MODE, false>, kBlock, lds_bytes);
  return b;
}

template <int RR, int NN>
hipError_t launch_vm_shape(int mode, bool slow, int grid, int lds_bytes, hipStream_t st, const uint8_t *blob,
                           Workspace ws) {
  switch (mode) {
    case VM_PROJECT: return launch_vm_mode<RR, NN, VM_PROJECT>(slow, grid, lds_bytes, st, blob, ws);
    case VM_PRED: return launch_vm_mode<RR, NN, VM_PRED>(slow, grid, lds_bytes, st, blob, ws);
    case VM_GATHER: return launch_vm_mode<RR, NN, VM_GATHER>(slow, grid, lds_bytes, st, blob, ws);
    default: return hipErrorInvalidValue;
  }
}

template <int RR, int NN>
int vm_occupancy(int mode, bool slow, int lds_bytes) {
  switch (mode) {
    case VM_PROJECT: return vm_occupancy_mode<RR, NN, VM_PROJECT>(slow, lds_bytes);
    case VM_PRED: return vm_occupancy_mode<RR, NN, VM_PRED>(slow, lds_bytes);
    case VM_GATHER: return vm_occupancy_mode<RR, NN, VM_GATHER>(slow, lds_bytes);
    default: return 0;
  }
}

hipError_t launch_vm(int r, int nt, int mode, bool slow, int grid, int lds_bytes, hipStream_t st,
                     const uint8_t *blob, Workspace ws) {
#define BLZ_VM_LAUNCH(RR, NN) \
  if (r == RR && nt == NN) return launch_vm_shape<RR, NN>(mode, slow, grid, lds_bytes, st, blob, ws);
  BLZ_VM_SHAPES(BLZ_VM_LAUNCH)
#undef BLZ_VM_LAUNCH
  return hipErrorInvalidValue;
}

int vm_blocks_per_cu(int r, int nt, int mode, bool slow, int lds_bytes) {
#define BLZ_VM_OCC(RR, NN) \
  if (r == RR && nt == NN) return vm_occupancy<RR, NN>(mode, slow, lds_bytes);
  BLZ_VM_SHAPES(BLZ_VM_OCC)
#undef BLZ_VM_OCC
  return 0;
}

bool vm_shape_ok(int r, int nt) {
#define BLZ_VM_OK(RR, NN) if (r == RR && nt == NN) return true;
  BLZ_VM_SHAPES(BLZ_VM_OK)
#undef BLZ_VM_OK
  return false;
}

hipError_t launch_slot_scan(const uint32_t *counts, uint32_t *prefix, int64_t n_slots, hipStream_t st,
                            Workspace ws) {
  const int64_t per = (int64_t)kBlock * kScanItems;
  const int grid = (int)(n_slots > 0 ? (n_slots + per - 1) / per : 1);
  hipLaunchKernelGGL(slot_scan_kernel, dim3(grid), dim3(kBlock), 0, st, counts, prefix, n_slots, ws);
  return hipGetLastError();
}
int64_t slot_scan_blocks(int64_t n_slots) {
  const int64_t per = (int64_t)kBlock * kScanItems;
  return n_slots > 0 ? (n_slots + per - 1) / per : 1;
}

}  // namespace blz

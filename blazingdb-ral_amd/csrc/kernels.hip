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

namespace blz {

#define DEV __device__ __forceinline__

DEV double b2d(uint64_t b) { return __builtin_bit_cast(double, b); }
DEV uint64_t d2b(double d) { return __builtin_bit_cast(uint64_t, d); }

// double -> int64, saturating, NaN -> 0 (the reference leaves out-of-range undefined).
DEV int64_t f2i(double d) {
  if (d != d) return 0;
  if (d >= 9223372036854775807.0) return INT64_MAX;
  if (d <= -9223372036854775808.0) return INT64_MIN;
  return (int64_t)d;
}
DEV uint64_t conv(uint64_t x, int cv) {
  if (cv == CV_I2F) return d2b((double)(int64_t)x);
  if (cv == CV_F2I) return (uint64_t)f2i(b2d(x));
  return x;
}

// interpreter_ops.cuh:114-153 (days since epoch; `int z` truncates as there).
DEV void civil(int64_t days, int &y, unsigned &doy, unsigned &mp) {
  const int z = (int)(days + 719468);
  const int era = (z >= 0 ? z : z - 146096) / 146097;
  const unsigned doe = (unsigned)(z - era * 146097);
  const unsigned yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
  y = (int)yoe + era * 400;
  doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
  mp = (5 * doy + 2) / 153;
}
DEV int64_t year_d(int64_t d) {
  int y; unsigned doy, mp; civil(d, y, doy, mp);
  const unsigned m = mp + (mp < 10 ? 3 : -9);
  return m <= 2 ? y + 1 : y;
}
DEV int64_t month_d(int64_t d) {
  int y; unsigned doy, mp; civil(d, y, doy, mp);
  return (int64_t)(unsigned)(mp + (mp < 10 ? 3 : -9));
}
DEV int64_t day_d(int64_t d) {
  int y; unsigned doy, mp; civil(d, y, doy, mp);
  return (int64_t)(unsigned)(doy - (153 * mp + 2) / 5 + 1);
}
// interpreter_ops.cuh:61-112 (milliseconds since epoch).
constexpr int64_t MS_DAY = 86400000, MS_HOUR = 3600000, MS_MIN = 60000, MS_SEC = 1000;
DEV int64_t ms_days(int64_t t) { return (t >= 0 ? t : t - (MS_DAY - 1)) / MS_DAY; }
DEV int64_t hour_ms(int64_t t) {
  return t >= 0 ? ((t % MS_DAY) / MS_HOUR) : ((MS_DAY + (t % MS_DAY)) / MS_HOUR);
}
DEV int64_t minute_ms(int64_t t) {
  return t >= 0 ? ((t % MS_HOUR) / MS_MIN) : ((MS_HOUR + (t % MS_HOUR)) / MS_MIN);
}
DEV int64_t second_ms(int64_t t) {
  return t >= 0 ? ((t % MS_MIN) / MS_SEC) : ((MS_MIN + (t % MS_MIN)) / MS_SEC);
}
DEV int64_t idiv(int64_t a, int64_t b) {
  return (b == 0 || (a == INT64_MIN && b == -1)) ? 0 : a / b;
}
DEV int64_t imod(int64_t a, int64_t b) {
  return (b == 0 || (a == INT64_MIN && b == -1)) ? 0 : a % b;
}
// interpreter_ops.cuh:667-672: `data = 1; for (i < R) data *= L` == L^R mod 2^64.
DEV uint64_t ipow(uint64_t base, int64_t e) {
  uint64_t r = 1;
  if (e <= 0) return 1;
  uint64_t n = (uint64_t)e;
  while (n) {
    if (n & 1) r *= base;
    base *= base;
    n >>= 1;
  }
  return r;
}

DEV bool valid_bit(const uint8_t *v, int64_t row) {
  return v == nullptr ? true : ((v[row >> 3] >> (row & 7)) & 1);
}

// Uniform value from LDS.
template <typename T>
DEV T uni(T v) {
  return (T)__builtin_amdgcn_readfirstlane((int)v);
}
DEV uint64_t uni64(uint64_t v) {
  uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}
DEV const void *uniptr(const void *p) { return (const void *)uni64((uint64_t)p); }

// Validity word `w` of a bitmap covering n rows; the last partial word is read
// byte by byte so a buffer of exactly ceil(n/8) bytes is never over-read.
DEV uint64_t load_vword(const uint8_t *v, int64_t w, int64_t n) {
  if ((w + 1) * 64 <= n) return *(const uint64_t *)(v + 8 * w);
  uint64_t x = 0;
  const int64_t nb = (n - w * 64 + 7) >> 3;
  for (int64_t b = 0; b < nb; b++) x |= (uint64_t)v[8 * w + b] << (8 * b);
  return x;
}
DEV void store_vword(uint8_t *v, int64_t w, int64_t n, uint64_t x) {
  if ((w + 1) * 64 <= n) { *(uint64_t *)(v + 8 * w) = x; return; }
  const int64_t nb = (n - w * 64 + 7) >> 3;
  for (int64_t b = 0; b < nb; b++) v[8 * w + b] = (uint8_t)(x >> (8 * b));
}
DEV uint64_t low_mask(int64_t bits) { return bits >= 64 ? ~0ull : ((1ull << bits) - 1); }

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kLdsBlobBytes = 12288;
constexpr int kMaxUnrolledInputs = 16;

// Raw load of one element (width class only; widened later).
template <int R>
DEV void load_raw(int kind, const void *base, const int64_t (&row)[R], uint64_t (&dst)[R]) {
  switch (kind) {
    case LD_I8:
#pragma unroll
      for (int u = 0; u < R; u++) dst[u] = (uint64_t)(int64_t)((const int8_t *)base)[row[u]];
      break;
    case LD_U8:
#pragma unroll
      for (int u = 0; u < R; u++) dst[u] = (uint64_t)((const uint8_t *)base)[row[u]];
      break;
    case LD_I16:
#pragma unroll
      for (int u = 0; u < R; u++) dst[u] = (uint64_t)(int64_t)((const int16_t *)base)[row[u]];
      break;
    case LD_I32: case LD_F32:
#pragma unroll
      for (int u = 0; u < R; u++) dst[u] = (uint64_t)((const uint32_t *)base)[row[u]];
      break;
    case LD_I64: case LD_F64:
#pragma unroll
      for (int u = 0; u < R; u++) dst[u] = ((const uint64_t *)base)[row[u]];
      break;
    default:
#pragma unroll
      for (int u = 0; u < R; u++) dst[u] = 0;
      break;
  }
}
DEV uint64_t widen(int kind, uint64_t raw) {
  if (kind == LD_I32) return (uint64_t)(int64_t)(int32_t)(uint32_t)raw;
  if (kind == LD_F32) return d2b((double)__builtin_bit_cast(float, (uint32_t)raw));
  return raw;
}

template <int R>
DEV void store_narrow(int kind, void *base, int64_t row, uint64_t v) {
  switch (kind) {
    case ST_I8: ((int8_t *)base)[row] = (int8_t)v; break;
    case ST_I16: ((int16_t *)base)[row] = (int16_t)v; break;
    case ST_I32: ((int32_t *)base)[row] = (int32_t)v; break;
    case ST_I64: ((uint64_t *)base)[row] = v; break;
    case ST_F32: ((float *)base)[row] = (float)b2d(v); break;
    case ST_F64: ((uint64_t *)base)[row] = v; break;
    default: break;
  }
}

// Transcendental and pow ops: out of line and one row at a time, so their
// register needs do not inflate the interpreter's common path.
__device__ __attribute__((noinline)) uint64_t heavy_op(int code, uint64_t a, uint64_t b) {
  const double x = b2d(a), y = b2d(b);
  switch (code) {
    case OP_POW_F: return d2b(pow(x, y));
    case OP_SIN: return d2b(sin(x));
    case OP_COS: return d2b(cos(x));
    case OP_ASIN: return d2b(asin(x));
    case OP_ACOS: return d2b(acos(x));
    case OP_TAN: return d2b(tan(x));
    case OP_COTAN: return d2b(cos(x) / sin(x));
    case OP_ATAN: return d2b(atan(x));
    case OP_LN: return d2b(log(x));
    default: return d2b(log10(x));
  }
}

// One fully typed op over R rows. `vb` carries the raw validity bit of the op's
// validity column per row (COALESCE / IS_[NOT_]NULL).
template <int R>
DEV void exec_op(int code, const uint64_t (&A)[R], const uint64_t (&B)[R], const bool (&vb)[R],
                 uint64_t (&Z)[R]) {
#define BLZ_CASE(CODE, EXPR)                     \
  case CODE: {                                   \
    _Pragma("unroll") for (int u = 0; u < R; u++) { \
      const uint64_t a = A[u], b = B[u];         \
      (void)a; (void)b;                          \
      Z[u] = (EXPR);                             \
    }                                            \
  } break;
  switch (code) {
    BLZ_CASE(OP_ADD_I, a + b)
    BLZ_CASE(OP_ADD_F, d2b(b2d(a) + b2d(b)))
    BLZ_CASE(OP_SUB_I, a - b)
    BLZ_CASE(OP_SUB_F, d2b(b2d(a) - b2d(b)))
    BLZ_CASE(OP_MUL_I, a * b)
    BLZ_CASE(OP_MUL_F, d2b(b2d(a) * b2d(b)))
    BLZ_CASE(OP_DIV_I, (uint64_t)idiv((int64_t)a, (int64_t)b))
    BLZ_CASE(OP_DIV_F, d2b(b2d(a) / b2d(b)))
    BLZ_CASE(OP_MOD_I, (uint64_t)imod((int64_t)a, (int64_t)b))
    BLZ_CASE(OP_POW_I, ipow(a, (int64_t)b))
    case OP_POW_F:
    BLZ_CASE(OP_EQ_I, (uint64_t)(a == b))
    BLZ_CASE(OP_EQ_F, (uint64_t)(b2d(a) == b2d(b)))
    BLZ_CASE(OP_NE_I, (uint64_t)(a != b))
    BLZ_CASE(OP_NE_F, (uint64_t)(b2d(a) != b2d(b)))
    BLZ_CASE(OP_LT_I, (uint64_t)((int64_t)a < (int64_t)b))
    BLZ_CASE(OP_LT_F, (uint64_t)(b2d(a) < b2d(b)))
    BLZ_CASE(OP_GT_I, (uint64_t)((int64_t)a > (int64_t)b))
    BLZ_CASE(OP_GT_F, (uint64_t)(b2d(a) > b2d(b)))
    BLZ_CASE(OP_LE_I, (uint64_t)((int64_t)a <= (int64_t)b))
    BLZ_CASE(OP_LE_F, (uint64_t)(b2d(a) <= b2d(b)))
    BLZ_CASE(OP_GE_I, (uint64_t)((int64_t)a >= (int64_t)b))
    BLZ_CASE(OP_GE_F, (uint64_t)(b2d(a) >= b2d(b)))
    BLZ_CASE(OP_COALESCE, vb[u] ? a : b)
    BLZ_CASE(OP_AND_B, a & b)
    BLZ_CASE(OP_FLOOR, d2b(floor(b2d(a))))
    BLZ_CASE(OP_CEIL, d2b(ceil(b2d(a))))
    case OP_SIN:
    case OP_COS:
    case OP_ASIN:
    case OP_ACOS:
    case OP_TAN:
    case OP_COTAN:
    case OP_ATAN:
    BLZ_CASE(OP_ABS, d2b(fabs(b2d(a))))
    case OP_LN:
    case OP_LOG10:
#pragma nounroll
      for (int u = 0; u < R; u++) Z[u] = heavy_op(code, A[u], B[u]);
      break;
    BLZ_CASE(OP_NOT_I, (uint64_t)(a == 0))
    BLZ_CASE(OP_NOT_F, (uint64_t)(!(b2d(a) != 0.0)))
    BLZ_CASE(OP_YEAR32, (uint64_t)year_d((int64_t)a))
    BLZ_CASE(OP_MONTH32, (uint64_t)month_d((int64_t)a))
    BLZ_CASE(OP_DAY32, (uint64_t)day_d((int64_t)a))
    BLZ_CASE(OP_YEAR64, (uint64_t)year_d(ms_days((int64_t)a)))
    BLZ_CASE(OP_MONTH64, (uint64_t)month_d(ms_days((int64_t)a)))
    BLZ_CASE(OP_DAY64, (uint64_t)day_d(ms_days((int64_t)a)))
    BLZ_CASE(OP_HOUR64, (uint64_t)hour_ms((int64_t)a))
    BLZ_CASE(OP_MINUTE64, (uint64_t)minute_ms((int64_t)a))
    BLZ_CASE(OP_SECOND64, (uint64_t)second_ms((int64_t)a))
    BLZ_CASE(OP_ZERO, 0ull)
    BLZ_CASE(OP_IS_NULL, (uint64_t)(!vb[u]))
    BLZ_CASE(OP_IS_NOT_NULL, (uint64_t)(vb[u]))
    default:  // OP_IDENT
#pragma unroll
      for (int u = 0; u < R; u++) Z[u] = A[u];
      break;
  }
#undef BLZ_CASE
}

struct LdsProg {
  ProgHeader h;
  const InCol *in;
  const OutCol *out;
  const DevOp *ops;
  const VOp *vops;
};

// Runs `nops` device ops starting at `ops` over the register file.
template <int S, int R>
DEV void run_ops(const DevOp *ops, int nops, uint64_t (&rf)[S * R], const int64_t (&row)[R],
                 const InCol *in) {
  for (int k = 0; k < nops; k++) {
    const DevOp &op = ops[k];
    // code | dst | a | b | flags | conv | vcol | pad, one LDS read into SGPRs
    const uint64_t ctl = uni64(*(const uint64_t *)&op);
    const int code = (int)(ctl & 0xFF), dst = (int)((ctl >> 8) & 0xFF);
    const int fa = (int)((ctl >> 32) & 0xFF), cv = (int)((ctl >> 40) & 0xFF);
    const int vcol = (int)((ctl >> 48) & 0xFF);
    uint64_t A[R], B[R], Z[R];
    bool vb[R];
    if (fa & F_A_IMM) {
      const uint64_t ia = uni64(op.imm_a);
#pragma unroll
      for (int u = 0; u < R; u++) A[u] = ia;
    } else {
      const int a = (int)((ctl >> 16) & 0xFF);
#pragma unroll
      for (int u = 0; u < R; u++) A[u] = rf[a * R + u];
    }
    if (fa & F_B_IMM) {
      const uint64_t ib = uni64(op.imm_b);
#pragma unroll
      for (int u = 0; u < R; u++) B[u] = ib;
    } else {
      const int b = (int)((ctl >> 24) & 0xFF);
#pragma unroll
      for (int u = 0; u < R; u++) B[u] = rf[b * R + u];
    }
    const int ca = cv & 3, cb = (cv >> 2) & 3, co = (cv >> 4) & 3;
    if (ca) {
#pragma unroll
      for (int u = 0; u < R; u++) A[u] = conv(A[u], ca);
    }
    if (cb) {
#pragma unroll
      for (int u = 0; u < R; u++) B[u] = conv(B[u], cb);
    }
    if (vcol != 0xFF) {
      const uint8_t *vp = (const uint8_t *)uniptr(in[vcol].valid);
#pragma unroll
      for (int u = 0; u < R; u++) vb[u] = valid_bit(vp, row[u]);
    } else {
#pragma unroll
      for (int u = 0; u < R; u++) vb[u] = true;
    }
    exec_op<R>(code, A, B, vb, Z);
    if (co) {
#pragma unroll
      for (int u = 0; u < R; u++) Z[u] = conv(Z[u], co);
    }
#pragma unroll
    for (int u = 0; u < R; u++) rf[dst * R + u] = Z[u];
  }
}

// Validity program: one 64-row word per lane (lanes [0, R) of each wave).
template <int S>
DEV void run_vops(const VOp *vops, int nvops, uint64_t (&vrf)[S]) {
  for (int k = 0; k < nvops; k++) {
    const VOp &op = vops[k];
    const int code = uni(op.code), d = uni(op.dst), a = uni(op.a), b = uni(op.b);
    const uint64_t x = vrf[a], y = vrf[b];
    vrf[d] = code == VOP_OR ? (x | y) : (x & y);
  }
}

// Value of a validity source for the word held by lane u.
template <int S>
DEV uint64_t vsrc_word(int vsrc, const uint64_t (&vrf)[S]) {
  if (vsrc >= 0) return vrf[vsrc];
  if (vsrc == -2) return 0ull;
  return ~0ull;
}

template <int S, int R, bool COMPACT>
__global__ void __launch_bounds__(kBlock)
interp_kernel(const uint8_t *__restrict__ blob, Workspace ws) {
  __shared__ __attribute__((aligned(16))) uint8_t s_blob[kLdsBlobBytes];
  __shared__ int64_t s_tile;
  __shared__ uint64_t s_wave_cnt[kWaves];
  __shared__ uint64_t s_prefix;
  __shared__ uint64_t s_vbits[kWaves * R + 2];

  // Stage the program in LDS.
  const int blob_bytes = *(const int32_t *)(blob + offsetof(ProgHeader, blob_bytes));
  for (int i = threadIdx.x * 16; i < blob_bytes; i += kBlock * 16)
    *(uint4 *)(s_blob + i) = *(const uint4 *)(blob + i);
  __syncthreads();
  const ProgHeader &h = *(const ProgHeader *)s_blob;
  const InCol *in = (const InCol *)(s_blob + uni(h.off_in));
  const OutCol *out = (const OutCol *)(s_blob + uni(h.off_out));
  const DevOp *ops = (const DevOp *)(s_blob + uni(h.off_ops));
  const VOp *vops = (const VOp *)(s_blob + uni(h.off_vops));
  const int64_t n = (int64_t)uni64((uint64_t)h.n_rows);
  const int n_in = uni(h.n_in), n_out = uni(h.n_out);
  const int n_ops_p = uni(h.n_ops_p), n_ops_q = uni(h.n_ops_q);
  const int n_vops_p = uni(h.n_vops_p), n_vops_q = uni(h.n_vops_q);
  const int any_valid_out = uni(h.any_valid_out);

  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t rows_per_tile = (int64_t)kBlock * R;
  const int64_t n_tiles = (n + rows_per_tile - 1) / rows_per_tile;

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

    // 1. Raw loads of every input column, all issued before any is widened.
#pragma unroll
    for (int c = 0; c < (S < kMaxUnrolledInputs ? S : kMaxUnrolledInputs); c++) {
      if (c < n_in) {
        const int kind = uni(in[c].load);
        const void *base = uniptr(in[c].data);
        uint64_t tmp[R];
        load_raw<R>(kind, base, lrow, tmp);
#pragma unroll
        for (int u = 0; u < R; u++) rf[c * R + u] = tmp[u];
      }
    }
    for (int c = kMaxUnrolledInputs; c < n_in; c++) {
      const int kind = uni(in[c].load);
      const void *base = uniptr(in[c].data);
      uint64_t tmp[R];
      load_raw<R>(kind, base, lrow, tmp);
#pragma unroll
      for (int u = 0; u < R; u++) rf[c * R + u] = tmp[u];
    }
    // Validity words of the wave's R groups: lane u < R holds word of group u.
    const int64_t wave_word = wave_base >> 6;
    if (any_valid_out || COMPACT) {
      const int64_t w = wave_word + (lane < R ? lane : 0);
      for (int c = 0; c < n_in; c++) {
        const int vs = uni(in[c].vslot);
        if (vs >= 0) {
          const uint8_t *vp = (const uint8_t *)uniptr(in[c].valid);
          vrf[vs] = (lane < R && w * 64 < n) ? load_vword(vp, w, n) : 0ull;
        }
      }
    }
    // 2. Widen.
#pragma unroll
    for (int c = 0; c < (S < kMaxUnrolledInputs ? S : kMaxUnrolledInputs); c++) {
      if (c < n_in) {
        const int kind = uni(in[c].load);
        if (kind == LD_I32 || kind == LD_F32) {
#pragma unroll
          for (int u = 0; u < R; u++) rf[c * R + u] = widen(kind, rf[c * R + u]);
        }
      }
    }
    for (int c = kMaxUnrolledInputs; c < n_in; c++) {
      const int kind = uni(in[c].load);
#pragma unroll
      for (int u = 0; u < R; u++) rf[c * R + u] = widen(kind, rf[c * R + u]);
    }

    if (!COMPACT) {
      // 3. Projection program, then narrowing stores of every output.
      run_ops<S, R>(ops, n_ops_q, rf, lrow, in);
      if (any_valid_out) run_vops<S>(vops, n_vops_q, vrf);
      for (int o = 0; o < n_out; o++) {
        const int kind = uni(out[o].store);
        const int slot = uni(out[o].slot);
        void *base = (void *)uniptr(out[o].data);
#pragma unroll
        for (int u = 0; u < R; u++)
          if (row[u] < n) store_narrow<R>(kind, base, row[u], rf[slot * R + u]);
        uint8_t *vp = (uint8_t *)uniptr(out[o].valid);
        if (vp != nullptr) {
          const int vsrc = uni(out[o].vsrc);
          const uint64_t word = vsrc_word<S>(vsrc, vrf);
          const int64_t w = wave_word + lane;
          uint64_t nulls = 0;
          if (lane < R && w * 64 < n) {
            store_vword(vp, w, n, word);
            nulls = __builtin_popcountll(~word & low_mask(n - w * 64));
          }
          // wave reduction of the null count, one atomic per wave
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
    const int st_slot = uni(h.stencil_slot), st_width = uni(h.stencil_width);
    const int st_vsrc = uni(h.stencil_vsrc);
    const uint64_t st_mask = low_mask(8 * st_width);
    bool keep[R];
    uint64_t kmask[R];
    uint64_t wave_cnt = 0;
    {
      // Predicate validity: a validity slot, a constant, or (bare-column
      // predicate) the raw bitmap of that column.
      const uint64_t vstencil = vsrc_word<S>(st_vsrc == -3 ? -1 : st_vsrc, vrf);
      const uint8_t *rawv = st_vsrc == -3 ? (const uint8_t *)uniptr(in[st_slot].valid) : nullptr;
#pragma unroll
      for (int u = 0; u < R; u++) {
        const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)vstencil, u);
        const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(vstencil >> 32), u);
        const uint64_t vw = ((uint64_t)hi << 32) | lo;
        const bool vbit = st_vsrc == -3 ? valid_bit(rawv, lrow[u]) : ((vw >> lane) & 1);
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
      if (wave_cnt > 0) {
        run_ops<S, R>(ops + n_ops_p, n_ops_q, rf, lrow, in);
      }
      if (any_valid_out) run_vops<S>(vops + n_vops_p, n_vops_q, vrf);
      for (int o = 0; o < n_out; o++) {
        const int kind = uni(out[o].store);
        const int slot = uni(out[o].slot);
        void *base = (void *)uniptr(out[o].data);
        if (wave_cnt > 0) {
#pragma unroll
          for (int u = 0; u < R; u++)
            if (keep[u]) store_narrow<R>(kind, base, (int64_t)pos[u], rf[slot * R + u]);
        }
        uint8_t *vp = (uint8_t *)uniptr(out[o].valid);
        if (vp != nullptr) {
          // Compacted validity bits: tile-local bit buffer in LDS, then words.
          const int vsrc = uni(out[o].vsrc);
          const uint8_t *rawv = vsrc == -3 ? (const uint8_t *)uniptr(in[slot].valid) : nullptr;
          const uint64_t vword = vsrc_word<S>(vsrc >= 0 ? vsrc : (vsrc == -3 ? -1 : vsrc), vrf);
          const int nw = (int)(((tile_prefix & 63) + tile_cnt + 63) >> 6);
          for (int i = threadIdx.x; i < nw; i += kBlock) s_vbits[i] = 0;
          __syncthreads();
          uint64_t nulls = 0;
#pragma unroll
          for (int u = 0; u < R; u++) {
            const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)vword, u);
            const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(vword >> 32), u);
            const uint64_t vw = ((uint64_t)hi << 32) | lo;
            const bool vbit = vsrc == -3 ? valid_bit(rawv, lrow[u]) : ((vw >> lane) & 1);
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
            unsigned long long *gw = (unsigned long long *)(vp + 8 * w);
            if (first_shared || last_shared) atomicOr(gw, (unsigned long long)s_vbits[i]);
            else *gw = s_vbits[i];
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

// Register-file shapes: S slots x R rows per lane. S*R <= 16 keeps the file in
// VGPRs; S = 32 / 64 (R = 1) serve large plans with the file in scratch.
hipError_t launch_interp(int s, int r, bool compact, int grid, hipStream_t st, const uint8_t *blob,
                         Workspace ws) {
  if (s == 8 && r == 1) return launch_sr<8, 1>(compact, grid, st, blob, ws);
  if (s == 8 && r == 2) return launch_sr<8, 2>(compact, grid, st, blob, ws);
  if (s == 16 && r == 1) return launch_sr<16, 1>(compact, grid, st, blob, ws);
  if (s == 32 && r == 1) return launch_sr<32, 1>(compact, grid, st, blob, ws);
  if (s == 64 && r == 1) return launch_sr<64, 1>(compact, grid, st, blob, ws);
  return hipErrorInvalidValue;
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

hipError_t launch_null_count(const uint8_t *valid, int64_t n, unsigned long long *out, int grid,
                             hipStream_t st) {
  hipLaunchKernelGGL(null_count_kernel, dim3(grid), dim3(kBlock), 0, st, valid, n, out);
  return hipGetLastError();
}

int blob_capacity() { return kLdsBlobBytes; }
int block_threads() { return kBlock; }

}  // namespace blz

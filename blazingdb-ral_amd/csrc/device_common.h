/*
 * device_common.h — device helpers shared by the kernels of kernels.hip and
 * vm_device.h: slot reinterpretation, conversions, date parts, wide global
 * loads/stores through explicit address spaces, validity words, the out-of-line
 * slow op set and the validity program.
 *
 * Included by exactly one place per translation unit; everything is inline or
 * static so the two code objects do not share symbols.
 */
#ifndef BLZ_DEVICE_COMMON_H_
#define BLZ_DEVICE_COMMON_H_

// Run-time compiled kernels (hiprtc) get the HIP runtime and fixed-width types
// from the compiler itself.
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#include "program.h"

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

// Explicit address spaces: column data is global memory (vector loads that do not
// share a wait counter with LDS), the program blob is read through the scalar
// cache (constant address space: uniform values land in SGPRs).
#define GLOBAL __attribute__((address_space(1)))
#define CONSTANT __attribute__((address_space(4)))

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <typename T>
DEV T gload(const void *base, int64_t i) { return ((const GLOBAL T *)base)[i]; }
template <typename T>
DEV void gstore(void *base, int64_t i, T v) { ((GLOBAL T *)base)[i] = v; }
// A store of data written once and not read back by the kernel (output columns).
template <typename T>
DEV void gstore_nt(void *base, int64_t i, T v) { __builtin_nontemporal_store(v, (GLOBAL T *)base + i); }

DEV bool valid_bit(const void *v, int64_t row) {
  return v == nullptr ? true : ((gload<uint8_t>(v, row >> 3) >> (row & 7)) & 1);
}

// Validity word `w` of a bitmap covering n rows; the last partial word is read
// byte by byte so a buffer of exactly ceil(n/8) bytes is never over-read.
DEV uint64_t load_vword(const void *v, int64_t w, int64_t n) {
  if ((w + 1) * 64 <= n) return gload<uint64_t>(v, w);
  uint64_t x = 0;
  const int64_t nb = (n - w * 64 + 7) >> 3;
  for (int64_t b = 0; b < nb; b++) x |= (uint64_t)gload<uint8_t>(v, 8 * w + b) << (8 * b);
  return x;
}
DEV void store_vword(void *v, int64_t w, int64_t n, uint64_t x) {
  if ((w + 1) * 64 <= n) { gstore<uint64_t>(v, w, x); return; }
  const int64_t nb = (n - w * 64 + 7) >> 3;
  for (int64_t b = 0; b < nb; b++) gstore<uint8_t>(v, 8 * w + b, (uint8_t)(x >> (8 * b)));
}
DEV uint64_t low_mask(int64_t bits) { return bits >= 64 ? ~0ull : ((1ull << bits) - 1); }
DEV uint64_t readlane64(uint64_t v, int lane) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, lane);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), lane);
  return ((uint64_t)hi << 32) | lo;
}

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kLdsBlobBytes = 12288;  // program size limit (bytes)
constexpr int kMaxUnrolledInputs = 16;

DEV bool is_wide(int kind) { return kind == LD_I64 || kind == LD_F64; }
// Widens a 32-bit raw load (interpreter_ops.cuh:386-452).
DEV uint64_t widen32(int kind, uint32_t w) {
  switch (kind) {
    case LD_I8: return (uint64_t)(int64_t)(int8_t)(uint8_t)w;
    case LD_U8: return (uint64_t)(uint8_t)w;
    case LD_I16: return (uint64_t)(int64_t)(int16_t)(uint16_t)w;
    case LD_I32: return (uint64_t)(int64_t)(int32_t)w;
    case LD_F32: return d2b((double)__builtin_bit_cast(float, w));
    default: return 0;
  }
}
// Raw load of a narrow column element into 32 bits (no conversion: the widening
// happens after every load of the tile is issued).
DEV uint32_t load_narrow(int kind, const void *base, int64_t row) {
  switch (kind) {
    case LD_I32: case LD_F32: return gload<uint32_t>(base, row);
    case LD_I16: return gload<uint16_t>(base, row);
    case LD_I8: case LD_U8: return gload<uint8_t>(base, row);
    default: return 0;
  }
}

// A bare-column output keeps its bytes (materialize_column copies by width): row
// irow of the input column to row orow of the output, raw (a float32 is not
// widened, so its NaN payloads stay as they are).
DEV void copy_alias(int kind, void *dst, int64_t orow, const void *src, int64_t irow) {
  switch (kind) {
    case ST_I8: gstore<uint8_t>(dst, orow, gload<uint8_t>(src, irow)); break;
    case ST_I16: gstore<uint16_t>(dst, orow, gload<uint16_t>(src, irow)); break;
    case ST_I32: case ST_F32: gstore<uint32_t>(dst, orow, gload<uint32_t>(src, irow)); break;
    case ST_I64: case ST_F64: gstore<uint64_t>(dst, orow, gload<uint64_t>(src, irow)); break;
    default: break;
  }
}

DEV void store_narrow(int kind, void *base, int64_t row, uint64_t v) {
  switch (kind) {
    case ST_I8: gstore<int8_t>(base, row, (int8_t)v); break;
    case ST_I16: gstore<int16_t>(base, row, (int16_t)v); break;
    case ST_I32: gstore<int32_t>(base, row, (int32_t)v); break;
    case ST_I64: gstore<uint64_t>(base, row, v); break;
    case ST_F32: gstore<float>(base, row, (float)b2d(v)); break;
    case ST_F64: gstore<uint64_t>(base, row, v); break;
    default: break;
  }
}

// Single-pass decoupled look-back, run by one wave: publishes `total` for block b
// of a sequence whose state words (bits 62-63: 0 nothing yet, 1 aggregate, 2
// inclusive prefix; bits 0-61 the value) are zero before the pass, and returns the
// exclusive prefix of block b (in every lane). Blocks are taken in the order
// workgroups start, so each block it waits for belongs to a running workgroup.
DEV uint64_t lookback_prefix(unsigned long long *state, int64_t b, uint64_t total, int lane) {
  constexpr uint64_t FA = 1ull << 62, FP = 2ull << 62, VMASK = (1ull << 62) - 1;
  if (b == 0) {
    if (lane == 0) __hip_atomic_store(state, FP | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return 0;
  }
  if (lane == 0) __hip_atomic_store(state + b, FA | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint64_t pre = 0;
  int64_t look = b - 1;
  while (true) {
    const int64_t i = look - lane;
    uint64_t st = FP;
    if (i >= 0) st = __hip_atomic_load(state + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t flag = st >> 62;
    const uint64_t is_x = __builtin_amdgcn_ballot_w64(flag == 0);
    const uint64_t is_p = __builtin_amdgcn_ballot_w64(flag == 2);
    // lanes up to and including the nearest inclusive prefix
    const uint64_t upto = is_p ? (((is_p & (~is_p + 1)) << 1) - 1) : ~0ull;
    if (is_x & upto) {
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    uint64_t x = ((upto >> lane) & 1) ? (st & VMASK) : 0;
    for (int s = 32; s > 0; s >>= 1) x += __shfl_xor(x, s);
    pre += x;
    if (is_p) break;
    look -= 64;
  }
  if (lane == 0) __hip_atomic_store(state + b, FP | (pre + total), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return pre;
}

// Every op outside the hot set (division, floor/ceil/abs, integer power and
// modulo, date parts, transcendentals): out of line and one row at a time, so
// their long instruction sequences do not set the interpreter's register budget.
static __device__ __attribute__((noinline)) uint64_t slow_op(int code, uint64_t a, uint64_t b) {
  const double x = b2d(a), y = b2d(b);
  switch (code) {
    case OP_DIV_I: return (uint64_t)idiv((int64_t)a, (int64_t)b);
    case OP_DIV_F: return d2b(x / y);
    case OP_MOD_I: return (uint64_t)imod((int64_t)a, (int64_t)b);
    case OP_POW_I: return ipow(a, (int64_t)b);
    case OP_POW_F: return d2b(pow(x, y));
    case OP_FLOOR: return d2b(floor(x));
    case OP_CEIL: return d2b(ceil(x));
    case OP_ABS: return d2b(fabs(x));
    case OP_SIN: return d2b(sin(x));
    case OP_COS: return d2b(cos(x));
    case OP_ASIN: return d2b(asin(x));
    case OP_ACOS: return d2b(acos(x));
    case OP_TAN: return d2b(tan(x));
    case OP_COTAN: return d2b(cos(x) / sin(x));
    case OP_ATAN: return d2b(atan(x));
    case OP_LN: return d2b(log(x));
    case OP_LOG10: return d2b(log10(x));
    case OP_YEAR32: return (uint64_t)year_d((int64_t)a);
    case OP_MONTH32: return (uint64_t)month_d((int64_t)a);
    case OP_DAY32: return (uint64_t)day_d((int64_t)a);
    case OP_YEAR64: return (uint64_t)year_d(ms_days((int64_t)a));
    case OP_MONTH64: return (uint64_t)month_d(ms_days((int64_t)a));
    case OP_DAY64: return (uint64_t)day_d(ms_days((int64_t)a));
    case OP_HOUR64: return (uint64_t)hour_ms((int64_t)a);
    case OP_MINUTE64: return (uint64_t)minute_ms((int64_t)a);
    case OP_SECOND64: return (uint64_t)second_ms((int64_t)a);
    default: return a;
  }
}
// Operand / result conversion, out of line (int64 <-> binary64 sequences).
static __device__ __attribute__((noinline)) uint64_t conv_slow(uint64_t x, int cv) { return conv(x, cv); }

// One fully typed op over R rows. `vb` carries the raw validity bit of the op's
// validity column per row (COALESCE / IS_[NOT_]NULL).
template <int R>
DEV void exec_op(int code, const uint64_t (&A)[R], const uint64_t (&B)[R], const bool (&vb)[R],
                 uint64_t (&Z)[R]) {
#define BLZ_CASE(CODE, EXPR)                        \
  case CODE: {                                      \
    _Pragma("unroll") for (int u = 0; u < R; u++) { \
      const uint64_t a = A[u], b = B[u];            \
      (void)a; (void)b;                             \
      Z[u] = (EXPR);                                \
    }                                               \
  } break;
  switch (code) {
    BLZ_CASE(OP_ADD_I, a + b)
    BLZ_CASE(OP_ADD_F, d2b(b2d(a) + b2d(b)))
    BLZ_CASE(OP_SUB_I, a - b)
    BLZ_CASE(OP_SUB_F, d2b(b2d(a) - b2d(b)))
    BLZ_CASE(OP_MUL_I, a * b)
    BLZ_CASE(OP_MUL_F, d2b(b2d(a) * b2d(b)))
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
    BLZ_CASE(OP_NOT_I, (uint64_t)(a == 0))
    BLZ_CASE(OP_NOT_F, (uint64_t)(!(b2d(a) != 0.0)))
    BLZ_CASE(OP_ZERO, 0ull)
    BLZ_CASE(OP_IS_NULL, (uint64_t)(!vb[u]))
    BLZ_CASE(OP_IS_NOT_NULL, (uint64_t)(vb[u]))
    case OP_IDENT:
#pragma unroll
      for (int u = 0; u < R; u++) Z[u] = A[u];
      break;
    default:  // the slow set
#pragma nounroll
      for (int u = 0; u < R; u++) Z[u] = slow_op(code, A[u], B[u]);
      break;
  }
#undef BLZ_CASE
}

// One fully typed op for one row, its code a template argument: the run-time
// compiled kernels instantiate the ops they use, each folded to that op's
// expression (the same expressions as exec_op and slow_op).
template <int C>
DEV uint64_t op_t(uint64_t a, uint64_t b, bool vb) {
  const double x = b2d(a), y = b2d(b);
  (void)x; (void)y; (void)vb;
  if constexpr (C == OP_ADD_I) return a + b;
  else if constexpr (C == OP_ADD_F) return d2b(x + y);
  else if constexpr (C == OP_SUB_I) return a - b;
  else if constexpr (C == OP_SUB_F) return d2b(x - y);
  else if constexpr (C == OP_MUL_I) return a * b;
  else if constexpr (C == OP_MUL_F) return d2b(x * y);
  else if constexpr (C == OP_DIV_I) return (uint64_t)idiv((int64_t)a, (int64_t)b);
  else if constexpr (C == OP_DIV_F) return d2b(x / y);
  else if constexpr (C == OP_MOD_I) return (uint64_t)imod((int64_t)a, (int64_t)b);
  else if constexpr (C == OP_POW_I) return ipow(a, (int64_t)b);
  else if constexpr (C == OP_POW_F) return d2b(pow(x, y));
  else if constexpr (C == OP_EQ_I) return (uint64_t)(a == b);
  else if constexpr (C == OP_EQ_F) return (uint64_t)(x == y);
  else if constexpr (C == OP_NE_I) return (uint64_t)(a != b);
  else if constexpr (C == OP_NE_F) return (uint64_t)(x != y);
  else if constexpr (C == OP_LT_I) return (uint64_t)((int64_t)a < (int64_t)b);
  else if constexpr (C == OP_LT_F) return (uint64_t)(x < y);
  else if constexpr (C == OP_GT_I) return (uint64_t)((int64_t)a > (int64_t)b);
  else if constexpr (C == OP_GT_F) return (uint64_t)(x > y);
  else if constexpr (C == OP_LE_I) return (uint64_t)((int64_t)a <= (int64_t)b);
  else if constexpr (C == OP_LE_F) return (uint64_t)(x <= y);
  else if constexpr (C == OP_GE_I) return (uint64_t)((int64_t)a >= (int64_t)b);
  else if constexpr (C == OP_GE_F) return (uint64_t)(x >= y);
  else if constexpr (C == OP_COALESCE) return vb ? a : b;
  else if constexpr (C == OP_AND_B) return a & b;
  else if constexpr (C == OP_FLOOR) return d2b(floor(x));
  else if constexpr (C == OP_CEIL) return d2b(ceil(x));
  else if constexpr (C == OP_SIN) return d2b(sin(x));
  else if constexpr (C == OP_COS) return d2b(cos(x));
  else if constexpr (C == OP_ASIN) return d2b(asin(x));
  else if constexpr (C == OP_ACOS) return d2b(acos(x));
  else if constexpr (C == OP_TAN) return d2b(tan(x));
  else if constexpr (C == OP_COTAN) return d2b(cos(x) / sin(x));
  else if constexpr (C == OP_ATAN) return d2b(atan(x));
  else if constexpr (C == OP_ABS) return d2b(fabs(x));
  else if constexpr (C == OP_LN) return d2b(log(x));
  else if constexpr (C == OP_LOG10) return d2b(log10(x));
  else if constexpr (C == OP_NOT_I) return (uint64_t)(a == 0);
  else if constexpr (C == OP_NOT_F) return (uint64_t)(!(x != 0.0));
  else if constexpr (C == OP_YEAR32) return (uint64_t)year_d((int64_t)a);
  else if constexpr (C == OP_MONTH32) return (uint64_t)month_d((int64_t)a);
  else if constexpr (C == OP_DAY32) return (uint64_t)day_d((int64_t)a);
  else if constexpr (C == OP_YEAR64) return (uint64_t)year_d(ms_days((int64_t)a));
  else if constexpr (C == OP_MONTH64) return (uint64_t)month_d(ms_days((int64_t)a));
  else if constexpr (C == OP_DAY64) return (uint64_t)day_d(ms_days((int64_t)a));
  else if constexpr (C == OP_HOUR64) return (uint64_t)hour_ms((int64_t)a);
  else if constexpr (C == OP_MINUTE64) return (uint64_t)minute_ms((int64_t)a);
  else if constexpr (C == OP_SECOND64) return (uint64_t)second_ms((int64_t)a);
  else if constexpr (C == OP_ZERO) return 0ull;
  else if constexpr (C == OP_IS_NULL) return (uint64_t)(!vb);
  else if constexpr (C == OP_IS_NOT_NULL) return (uint64_t)vb;
  else return a;  // OP_IDENT
}

// Validity program: one 64-row word per lane (lanes [0, R) of each wave).
template <int S>
DEV void run_vops(const CONSTANT VOp *vops, int nvops, uint64_t (&vrf)[S]) {
  for (int k = 0; k < nvops; k++) {
    const uint32_t w = *(const CONSTANT uint32_t *)(vops + k);
    const int code = w & 0xFF, d = (w >> 8) & 0xFF, a = (w >> 16) & 0xFF, b = (w >> 24) & 0xFF;
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

// Validity words of a pass: lane u < R loads group u's word of every staged bitmap.
template <int S, int R>
DEV void read_pass_valid(const CONSTANT InCol *in, int n_in, const uint8_t *lds, int group0,
                         uint64_t (&vrf)[S]) {
  const int lane = threadIdx.x & 63;
  const int g = group0 + (lane < R ? lane : 0);
  for (int c = 0; c < n_in; c++) {
    const int vs = in[c].vslot;
    if (vs >= 0) vrf[vs] = *(const uint64_t *)(lds + in[c].lds_voff + 8 * g);
  }
}


}  // namespace blz
#endif

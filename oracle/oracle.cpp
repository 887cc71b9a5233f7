/*
 * oracle.cpp — CPU restatement of the BlazingSQL RAL expression/filter path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker the parity tests, smoke() and
 * the CPU-baseline leg of bench.py compare the HIP path against. The product
 * (libblzral.so) never links, loads or calls it.
 *
 * Scalar, single thread, row at a time, compiled with -O2 -ffp-contract=off so
 * every floating-point operation rounds once, exactly as written.
 *
 * What it restates (reference file:line):
 *   - per-op kernel type resolution on the host   interpreter_ops.cuh:1095-1178
 *   - plan output dtypes                           CalciteExpressionParsing.cpp:177-270
 *   - scalar widening (scalar_mode)                interpreter_ops.cuh:156-175
 *   - input widening into 8-byte slots             interpreter_ops.cuh:386-452
 *   - the operator table                           interpreter_ops.cuh:473-838
 *   - narrowing writes of final outputs            interpreter_ops.cuh:317-381
 *   - 64-row validity words                        interpreter_ops.cuh:507-571, 1301-1374
 *   - date part extraction                         interpreter_ops.cuh:52-153
 *   - stencil -> ascending row indices             CalciteInterpreter.cpp:520-548
 *   - gather of data and validity, null count      ColumnManipulation.cu:26-150
 *
 * Where the reference is undefined the oracle fixes one behaviour, and the HIP
 * path implements the same one:
 *   - a slot read before any write this row holds 0 (data) and 0 (validity);
 *   - a BOOL8 input (never read by the reference) holds its byte value;
 *   - a null-scalar operand (-3) reads data 0;
 *   - integer / and % by 0, and INT64_MIN / -1, give 0;
 *   - double -> int64 conversion saturates, NaN -> 0;
 *   - IS_NULL / IS_NOT_NULL / COALESCE whose left operand is not an input column
 *     treat the row as valid;
 *   - the validity "force" rule indexes the unary-op list with the column index
 *     (as the reference does) only for column indices < number of ops.
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

#include "../include/blz/gdf_types.h"

namespace {

enum ScalarMode { SCALAR_INTENDED = 0, SCALAR_SHIPPED = 1 };

// interpreter_ops.cuh:180-190 (isInt) — the dtypes the kernel treats as integers.
bool is_int_kernel(int t) {
  return t == GDF_INT32 || t == GDF_INT64 || t == GDF_INT16 || t == GDF_INT8 ||
         t == GDF_DATE32 || t == GDF_DATE64 || t == GDF_TIMESTAMP ||
         t == GDF_STRING_CATEGORY;
}
// interpreter_ops.cuh:197-201 (isFloat)
bool is_float(int t) { return t == GDF_FLOAT32 || t == GDF_FLOAT64; }
// CalciteExpressionParsing.cpp:21-31 (is_type_signed; also is_numeric_type :53-56)
bool is_signed(int t) {
  return t == GDF_INT8 || t == GDF_INT16 || t == GDF_INT32 || t == GDF_INT64 ||
         t == GDF_FLOAT32 || t == GDF_FLOAT64 || t == GDF_DATE32 ||
         t == GDF_DATE64 || t == GDF_TIMESTAMP;
}
bool is_date(int t) { return t == GDF_DATE32 || t == GDF_DATE64 || t == GDF_TIMESTAMP; }
int width(int t) { return blz_dtype_width((gdf_dtype)t); }

bool is_trig_or_log(int u) {
  return u == BLZ_SIN || u == BLZ_COS || u == BLZ_ASIN || u == BLZ_ACOS ||
         u == BLZ_TAN || u == BLZ_COTAN || u == BLZ_ATAN || u == BLZ_LOG ||
         u == BLZ_LN;
}

// CalciteExpressionParsing.cpp:177-191 (unary output dtype). The CAST_FLOAT64
// extension yields FLOAT64.
int unary_out_type(int l, int u) {
  if (u == BLZ_CAST_FLOAT64) return GDF_FLOAT64;
  if (is_date(l)) return GDF_INT16;
  if (is_trig_or_log(u)) return is_float(l) ? l : GDF_FLOAT64;
  if (u == BLZ_IS_NULL || u == BLZ_IS_NOT_NULL) return GDF_INT8;
  return l;
}

// CalciteExpressionParsing.cpp:194-270 (binary output dtype). Every numeric dtype
// counts as signed, so only the "both signed" arm of the integer rule is live.
int binary_out_type(int l, int r, int op) {
  if (!is_signed(l) || !is_signed(r)) return GDF_invalid;
  switch (op) {
    case GDF_ADD: case GDF_SUB: case GDF_MUL: case GDF_DIV: case GDF_MOD:
      if (is_float(l) || is_float(r)) {
        if (is_float(l) && is_float(r)) return width(l) >= width(r) ? l : r;
        return is_float(l) ? l : r;
      }
      return width(l) >= width(r) ? l : r;
    case GDF_EQUAL: case GDF_NOT_EQUAL: case GDF_GREATER: case GDF_GREATER_EQUAL:
    case GDF_LESS: case GDF_LESS_EQUAL:
      return GDF_INT8;
    case GDF_POW:
      return (is_float(l) || is_float(r)) ? GDF_FLOAT64 : GDF_INT64;
    case GDF_COALESCE:
      return l;
    default:
      return GDF_invalid;
  }
}

// interpreter_ops.cuh:156-175 (scale_to_64_bit_return_bytes). SHIPPED truncates
// float scalars to int64 (its bits are later read as a double); INTENDED keeps the
// value's own binary64 bits (the commented-out form on :167-168).
int64_t scalar_bits(const gdf_scalar& s, int mode) {
  switch (s.dtype) {
    case GDF_INT8: return s.data.si08;
    case GDF_INT16: return s.data.si16;
    case GDF_INT32: case GDF_STRING_CATEGORY: return s.data.si32;
    case GDF_INT64: return s.data.si64;
    case GDF_DATE32: return s.data.dt32;
    case GDF_DATE64: return s.data.dt64;
    case GDF_TIMESTAMP: return s.data.tmst;
    case GDF_FLOAT32: {
      if (mode == SCALAR_SHIPPED) return (int64_t)s.data.fp32;
      double d = (double)s.data.fp32;
      int64_t b; std::memcpy(&b, &d, 8); return b;
    }
    case GDF_FLOAT64: {
      if (mode == SCALAR_SHIPPED) return (int64_t)s.data.fp64;
      int64_t b; std::memcpy(&b, &s.data.fp64, 8); return b;
    }
    default: return 0;
  }
}

inline double as_f64(int64_t b) { double d; std::memcpy(&d, &b, 8); return d; }
inline int64_t as_i64(double d) { int64_t b; std::memcpy(&b, &d, 8); return b; }

// double -> int64: saturating, NaN -> 0 (the reference leaves this undefined).
inline int64_t f2i(double d) {
  if (d != d) return 0;
  if (d >= 9223372036854775807.0) return INT64_MAX;
  if (d <= -9223372036854775808.0) return INT64_MIN;
  return (int64_t)d;
}
inline int64_t wrap_add(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }
inline int64_t wrap_sub(int64_t a, int64_t b) { return (int64_t)((uint64_t)a - (uint64_t)b); }
inline int64_t wrap_mul(int64_t a, int64_t b) { return (int64_t)((uint64_t)a * (uint64_t)b); }
inline int64_t idiv(int64_t a, int64_t b) {
  if (b == 0 || (a == INT64_MIN && b == -1)) return 0;
  return a / b;
}
inline int64_t imod(int64_t a, int64_t b) {
  if (b == 0 || (a == INT64_MIN && b == -1)) return 0;
  return a % b;
}
// interpreter_ops.cuh:657-676, integer arm: `data = 1; for (i < R) data *= L`,
// i.e. L^R modulo 2^64 for R > 0 and 1 otherwise.
inline int64_t ipow(int64_t base, int64_t e) {
  uint64_t r = 1, b = (uint64_t)base;
  if (e <= 0) return 1;
  uint64_t n = (uint64_t)e;
  while (n) { if (n & 1) r *= b; b *= b; n >>= 1; }
  return (int64_t)r;
}

// interpreter_ops.cuh:114-153 (days-based date parts; `int z` truncates).
int64_t year32(int64_t d) {
  const int z = (int)(d + 719468);
  const int era = (z >= 0 ? z : z - 146096) / 146097;
  const unsigned doe = (unsigned)(z - era * 146097);
  const unsigned yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
  const int y = (int)yoe + era * 400;
  const unsigned doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
  const unsigned mp = (5 * doy + 2) / 153;
  const unsigned m = mp + (mp < 10 ? 3 : -9);
  return m <= 2 ? y + 1 : y;
}
int64_t month32(int64_t d) {
  const int z = (int)(d + 719468);
  const int era = (z >= 0 ? z : z - 146096) / 146097;
  const unsigned doe = (unsigned)(z - era * 146097);
  const unsigned yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
  const unsigned doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
  const unsigned mp = (5 * doy + 2) / 153;
  return (int64_t)(unsigned)(mp + (mp < 10 ? 3 : -9));
}
int64_t day32(int64_t d) {
  const int z = (int)(d + 719468);
  const int era = (z >= 0 ? z : z - 146096) / 146097;
  const unsigned doe = (unsigned)(z - era * 146097);
  const unsigned yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
  const unsigned doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
  const unsigned mp = (5 * doy + 2) / 153;
  return (int64_t)(unsigned)(doy - (153 * mp + 2) / 5 + 1);
}
// interpreter_ops.cuh:61-112 (millisecond-based date parts).
const int64_t MS_DAY = 86400000, MS_HOUR = 3600000, MS_MIN = 60000, MS_SEC = 1000;
int64_t ms_to_days(int64_t t) { return (t >= 0 ? t : t - (MS_DAY - 1)) / MS_DAY; }
int64_t year64(int64_t t) { return year32(ms_to_days(t)); }
int64_t month64(int64_t t) { return month32(ms_to_days(t)); }
int64_t day64(int64_t t) { return day32(ms_to_days(t)); }
int64_t hour64(int64_t t) {
  return t >= 0 ? ((t % MS_DAY) / MS_HOUR) : ((MS_DAY + (t % MS_DAY)) / MS_HOUR);
}
int64_t minute64(int64_t t) {
  return t >= 0 ? ((t % MS_HOUR) / MS_MIN) : ((MS_HOUR + (t % MS_HOUR)) / MS_MIN);
}
int64_t second64(int64_t t) {
  return t >= 0 ? ((t % MS_MIN) / MS_SEC) : ((MS_MIN + (t % MS_MIN)) / MS_SEC);
}

inline bool bit_at(const uint8_t* v, int64_t row) {
  return v == nullptr ? true : ((v[row >> 3] >> (row & 7)) & 1);
}
inline uint64_t word_at(const uint8_t* v, int64_t w) {
  uint64_t x; std::memcpy(&x, v + 8 * w, 8); return x;
}

struct TypedOp {
  int lk, rk, ok;   // kernel types: 0 = int64, 1 = double
  int lt, rt;       // resolved plan dtypes (for POW's float test and date parts)
};

}  // namespace

extern "C" {

/*
 * perform_operation restated (interpreter_cpp.cu:219-286 with the functor ctor at
 * interpreter_ops.cuh:959-1266 and transformKernel :1417-1458). Buffers are HOST
 * memory. Validity buffers must span ceil(size/64) whole 64-bit words. Writes
 * `null_counts_out[k]` = size - popcount(valid bits < size) for outputs that have
 * validity (0 otherwise), which the reference leaves to its callers.
 * Returns 0, or -1 for a plan that references a slot outside [0, 64).
 */
int oracle_perform_operation(const gdf_column* outs, int n_out, const gdf_column* ins,
                             int n_in, const int16_t* left, const int16_t* right,
                             const int16_t* outpos, int n_ops, const int16_t* final_pos,
                             const int32_t* binop, const int32_t* unop,
                             const gdf_scalar* lsc, const gdf_scalar* rsc, int scalar_mode,
                             int32_t* null_counts_out) {
  const int NS = 64;
  for (int k = 0; k < n_ops; k++) {
    if (outpos[k] < 0 || outpos[k] >= NS || left[k] >= NS || right[k] >= NS) return -1;
  }
  for (int k = 0; k < n_out; k++)
    if (final_pos[k] < 0 || final_pos[k] >= NS) return -1;
  const int64_t n = n_in > 0 ? ins[0].size : 0;

  // Host type resolution, interpreter_ops.cuh:1095-1178.
  std::vector<TypedOp> ty(n_ops);
  std::vector<int64_t> lbits(n_ops, 0), rbits(n_ops, 0);
  std::map<int, int> out_map;
  for (int k = 0; k < n_ops; k++) {
    int L = left[k], R = right[k];
    int lt = GDF_invalid, rt = GDF_invalid;
    if (L >= 0 && L < n_in) lt = ins[L].dtype;
    else if (L == SCALAR_NULL_INDEX) lt = lsc[k].dtype;
    else if (L == SCALAR_INDEX) {
      lt = is_int_kernel(lsc[k].dtype) ? GDF_INT64 : GDF_FLOAT64;
      lbits[k] = scalar_bits(lsc[k], scalar_mode);
    } else if (L >= n_in) lt = out_map.count(L) ? out_map[L] : GDF_invalid;
    if (R >= 0 && R < n_in) rt = ins[R].dtype;
    else if (R == SCALAR_NULL_INDEX) rt = rsc[k].dtype;
    else if (R == SCALAR_INDEX) {
      rt = is_int_kernel(rsc[k].dtype) ? GDF_INT64 : GDF_FLOAT64;
      rbits[k] = scalar_bits(rsc[k], scalar_mode);
    } else if (R >= n_in) rt = out_map.count(R) ? out_map[R] : GDF_invalid;
    int t = (R == BLZ_UNARY_INDEX) ? unary_out_type(lt, unop[k]) : binary_out_type(lt, rt, binop[k]);
    int ot = GDF_invalid;
    if (is_signed(t) && !is_float(t)) ot = GDF_INT64;
    else if (is_float(t)) ot = GDF_FLOAT64;
    out_map[outpos[k]] = ot;
    ty[k] = TypedOp{is_float(lt) ? 1 : 0, is_float(rt) ? 1 : 0, is_float(ot) ? 1 : 0, lt, rt};
  }

  int64_t slot[NS];
  for (int64_t row = 0; row < n; row++) {
    std::memset(slot, 0, sizeof(slot));
    // read_data, interpreter_ops.cuh:386-452
    for (int c = 0; c < n_in; c++) {
      const void* d = ins[c].data;
      switch (ins[c].dtype) {
        case GDF_INT8: slot[c] = ((const int8_t*)d)[row]; break;
        case GDF_INT16: slot[c] = ((const int16_t*)d)[row]; break;
        case GDF_INT32: case GDF_DATE32: case GDF_STRING_CATEGORY:
          slot[c] = ((const int32_t*)d)[row]; break;
        case GDF_INT64: case GDF_DATE64: case GDF_TIMESTAMP:
          slot[c] = ((const int64_t*)d)[row]; break;
        case GDF_FLOAT32: slot[c] = as_i64((double)((const float*)d)[row]); break;
        case GDF_FLOAT64: slot[c] = as_i64(((const double*)d)[row]); break;
        case GDF_BOOL8: slot[c] = ((const uint8_t*)d)[row]; break;  // unread in the reference
        default: break;
      }
    }
    // process_operator -> _3, interpreter_ops.cuh:473-505, 574-838
    for (int k = 0; k < n_ops; k++) {
      const int L = left[k], R = right[k];
      const TypedOp& T = ty[k];
      int64_t lb = 0, rb = 0;
      if (L == SCALAR_INDEX) lb = lbits[k]; else if (L >= 0) lb = slot[L];
      if (R == SCALAR_INDEX) rb = rbits[k]; else if (R >= 0) rb = slot[R];
      int64_t res = 0;
      auto store_i = [&](int64_t v) { res = T.ok ? as_i64((double)v) : v; };
      auto store_d = [&](double v) { res = T.ok ? as_i64(v) : f2i(v); };
      auto store_b = [&](bool v) { store_i(v ? 1 : 0); };
      if (R != BLZ_UNARY_INDEX) {
        const bool fp = T.lk || T.rk;  // usual arithmetic conversions
        const double ld = T.lk ? as_f64(lb) : (double)lb;
        const double rd = T.rk ? as_f64(rb) : (double)rb;
        switch (binop[k]) {
          case GDF_ADD: if (fp) store_d(ld + rd); else store_i(wrap_add(lb, rb)); break;
          case GDF_SUB: if (fp) store_d(ld - rd); else store_i(wrap_sub(lb, rb)); break;
          case GDF_MUL: if (fp) store_d(ld * rd); else store_i(wrap_mul(lb, rb)); break;
          case GDF_DIV: case GDF_FLOOR_DIV:
            if (fp) store_d(ld / rd); else store_i(idiv(lb, rb)); break;
          case GDF_COALESCE: {
            const uint8_t* v = (L >= 0 && L < n_in) ? ins[L].valid : nullptr;
            if (bit_at(v, row)) { if (T.lk) store_d(ld); else store_i(lb); }
            else { if (T.rk) store_d(rd); else store_i(rb); }
            break;
          }
          case GDF_MOD: {
            int64_t a = T.lk ? f2i(ld) : lb, b = T.rk ? f2i(rd) : rb;
            store_i(imod(a, b));
            break;
          }
          case GDF_POW:
            if (is_float(T.lt) || is_float(T.rt)) store_d(std::pow(ld, rd));
            else store_i(ipow(lb, rb));
            break;
          case GDF_EQUAL: store_b(fp ? ld == rd : lb == rb); break;
          case GDF_NOT_EQUAL: store_b(fp ? ld != rd : lb != rb); break;
          case GDF_LESS: store_b(fp ? ld < rd : lb < rb); break;
          case GDF_GREATER: store_b(fp ? ld > rd : lb > rb); break;
          case GDF_LESS_EQUAL: store_b(fp ? ld <= rd : lb <= rb); break;
          case GDF_GREATER_EQUAL: store_b(fp ? ld >= rd : lb >= rb); break;
          default: return -2;  // no kernel branch in the reference (slot unwritten)
        }
      } else {
        const double ld = T.lk ? as_f64(lb) : (double)lb;
        const int64_t li = T.lk ? f2i(ld) : lb;  // int64 view for the date parts
        const bool d32 = T.lt == GDF_DATE32;
        switch (unop[k]) {
          case BLZ_FLOOR: store_d(std::floor(ld)); break;
          case BLZ_CEIL: store_d(std::ceil(ld)); break;
          case BLZ_SIN: store_d(std::sin(ld)); break;
          case BLZ_COS: store_d(std::cos(ld)); break;
          case BLZ_ASIN: store_d(std::asin(ld)); break;
          case BLZ_ACOS: store_d(std::acos(ld)); break;
          case BLZ_TAN: store_d(std::tan(ld)); break;
          case BLZ_COTAN: store_d(std::cos(ld) / std::sin(ld)); break;
          case BLZ_ATAN: store_d(std::atan(ld)); break;
          case BLZ_ABS: store_d(std::fabs(ld)); break;
          case BLZ_NOT: store_b(T.lk ? !(ld != 0.0) : !lb); break;
          case BLZ_LN: store_d(std::log(ld)); break;
          case BLZ_LOG: store_d(std::log10(ld)); break;
          case BLZ_YEAR: store_i(d32 ? year32(li) : year64(li)); break;
          case BLZ_MONTH: store_i(d32 ? month32(li) : month64(li)); break;
          case BLZ_DAY: store_i(d32 ? day32(li) : day64(li)); break;
          case BLZ_HOUR: store_i(d32 ? 0 : hour64(li)); break;
          case BLZ_MINUTE: store_i(d32 ? 0 : minute64(li)); break;
          case BLZ_SECOND: store_i(d32 ? 0 : second64(li)); break;
          case BLZ_IS_NULL: case BLZ_IS_NOT_NULL: {
            const uint8_t* v = (L >= 0 && L < n_in) ? ins[L].valid : nullptr;
            bool valid = bit_at(v, row);
            store_b(unop[k] == BLZ_IS_NULL ? !valid : valid);
            break;
          }
          case BLZ_CAST_FLOAT64: store_d(ld); break;
          default:  // `computed = left_value` with no further branch
            if (T.lk) store_d(ld); else store_i(lb);
            break;
        }
      }
      slot[outpos[k]] = res;
    }
    // write_data, interpreter_ops.cuh:317-381
    for (int o = 0; o < n_out; o++) {
      int64_t b = slot[final_pos[o]];
      void* d = outs[o].data;
      switch (outs[o].dtype) {
        case GDF_INT8: ((int8_t*)d)[row] = (int8_t)b; break;
        case GDF_INT16: ((int16_t*)d)[row] = (int16_t)b; break;
        case GDF_INT32: case GDF_DATE32: case GDF_STRING_CATEGORY:
          ((int32_t*)d)[row] = (int32_t)b; break;
        case GDF_INT64: case GDF_DATE64: case GDF_TIMESTAMP: ((int64_t*)d)[row] = b; break;
        case GDF_FLOAT32: ((float*)d)[row] = (float)as_f64(b); break;
        case GDF_FLOAT64: ((double*)d)[row] = as_f64(b); break;
        default: break;
      }
    }
  }

  // valid_operator, interpreter_ops.cuh:1301-1374 (with process_valids :507-571)
  bool any_valid_out = false;
  for (int o = 0; o < n_out; o++) any_valid_out |= outs[o].valid != nullptr;
  const int64_t nwords = (n + 63) / 64;
  if (any_valid_out) {
    std::vector<bool> forced(n_in);
    for (int c = 0; c < n_in; c++)
      forced[c] = ins[c].valid == nullptr || ins[c].null_count == 0 ||
                  (c < n_ops && (unop[c] == BLZ_IS_NULL || unop[c] == BLZ_IS_NOT_NULL));
    uint64_t vs[NS];
    for (int64_t w = 0; w < nwords; w++) {
      std::memset(vs, 0, sizeof(vs));
      for (int c = 0; c < n_in; c++) vs[c] = forced[c] ? ~0ull : word_at(ins[c].valid, w);
      for (int k = 0; k < n_ops; k++) {
        const int L = left[k], R = right[k];
        uint64_t lv = L == SCALAR_INDEX ? ~0ull : L == SCALAR_NULL_INDEX ? 0ull : (L >= 0 ? vs[L] : 0ull);
        if (R != BLZ_UNARY_INDEX) {
          uint64_t rv = R == SCALAR_INDEX ? ~0ull : R == SCALAR_NULL_INDEX ? 0ull : (R >= 0 ? vs[R] : 0ull);
          vs[outpos[k]] = binop[k] == GDF_COALESCE ? (lv | rv) : (lv & rv);
        } else {
          vs[outpos[k]] = lv;
        }
      }
      for (int o = 0; o < n_out; o++)
        if (outs[o].valid) std::memcpy(outs[o].valid + 8 * w, &vs[final_pos[o]], 8);
    }
  }
  if (null_counts_out) {
    for (int o = 0; o < n_out; o++) {
      int64_t cnt = 0;
      if (outs[o].valid && n > 0) {
        for (int64_t r = 0; r < n; r++) cnt += bit_at(outs[o].valid, r);
        null_counts_out[o] = (int32_t)(n - cnt);
      } else {
        null_counts_out[o] = 0;
      }
    }
  }
  return 0;
}

/*
 * Stencil -> ascending row indices: process_filter's gdf_sequence +
 * cudf::apply_boolean_mask (CalciteInterpreter.cpp:532-543; Utils.cu:4-7).
 * Row i is kept when its stencil value is non-zero and its validity bit is set
 * (cuDF's documented contract; the reference's own tests only hold 0/1 stencils
 * without nulls). Returns the number of rows kept.
 */
int64_t oracle_stencil_select(const gdf_column* stencil, int32_t* idx_out) {
  int64_t m = 0;
  const int w = width(stencil->dtype);
  for (int64_t i = 0; i < stencil->size; i++) {
    bool nz = false;
    const uint8_t* p = (const uint8_t*)stencil->data + (size_t)i * w;
    for (int b = 0; b < w; b++) nz |= p[b] != 0;
    if (nz && bit_at(stencil->valid, i)) idx_out[m++] = (int32_t)i;
  }
  return m;
}

/*
 * materialize_column (ColumnManipulation.cu:83-150): out[i] = in[idx[i]] for
 * idx[i] >= 0; validity bit i = validity bit idx[i] (gather_bits, :26-68; 0 for a
 * negative index; an index past the input, which the reference would read out of
 * bounds, is treated like a negative one); null_count = n - popcount(valid bits < n) (:105-113).
 * Output padding bits beyond n are written as 0.
 */
int32_t oracle_gather(const gdf_column* in, gdf_column* out, const int32_t* idx, int32_t n_idx) {
  const int w = width(in->dtype);
  for (int32_t i = 0; i < n_idx; i++) {
    if (idx[i] >= 0 && idx[i] < in->size)
      std::memcpy((uint8_t*)out->data + (size_t)i * w, (const uint8_t*)in->data + (size_t)idx[i] * w, w);
  }
  int32_t nulls = 0;
  if (out->valid != nullptr) {
    std::memset(out->valid, 0, ((size_t)n_idx + 7) / 8);
    for (int32_t i = 0; i < n_idx; i++) {
      bool b = idx[i] >= 0 && idx[i] < in->size && bit_at(in->valid, idx[i]);
      if (b) out->valid[i >> 3] |= (uint8_t)(1u << (i & 7));
      else nulls++;
    }
  }
  out->null_count = (n_idx == 0 || out->valid == nullptr) ? 0 : nulls;
  return out->null_count;
}

/* gdf_column_cpp::update_null_count (GDFColumn.cu:519-528). */
int32_t oracle_null_count(const uint8_t* valid, int32_t n) {
  if (n == 0 || valid == nullptr) return 0;
  int32_t c = 0;
  for (int32_t i = 0; i < n; i++) c += bit_at(valid, i);
  return n - c;
}

}  // extern "C"

/*
 * program.cpp — compile reference-format plans into one device program blob.
 *
 * Host work that the reference does per call inside the InterpreterFunctor
 * constructor (interpreter_ops.cuh:959-1266) and per row inside the kernel
 * (interpreter_ops.cuh:473-505, 1301-1374), done once here:
 *  1. per-op type resolution (plan dtypes -> int64 / binary64 kernel types), with
 *     the same rules as interpreter_ops.cuh:1095-1178 and
 *     CalciteExpressionParsing.cpp:177-270;
 *  2. lowering each op to a fully typed device op with explicit operand and
 *     result conversions (C++ usual arithmetic conversions made explicit);
 *     scalar operands become immediates (scalar_mode picks the reference's shipped
 *     float-literal truncation, interpreter_ops.cuh:167-168, or the intended
 *     value-preserving widening);
 *  3. SSA over the plan's register slots, Sethi-Ullman ordering of each output's
 *     expression tree and linear-scan allocation of device slots, so a C1-style
 *     predicate needs 2 temporaries instead of the planner's 5;
 *  4. the validity program (AND / OR / copy of 64-row words, process_valids
 *     :507-571) constant-folded: forced-valid inputs (the Q3 rule of :1312-1327)
 *     and scalars fold away, so null-free plans run no validity ops at all.
 */
#include "program_compiler.h"

#include <algorithm>
#include <cstring>
#include <functional>
#include <map>
#include <stdexcept>

#include "codegen.h"
#include "kernels.h"
#include "ssa.h"

namespace blz {
namespace {

bool is_int_kernel(int t) {  // interpreter_ops.cuh:180-190
  return t == GDF_INT32 || t == GDF_INT64 || t == GDF_INT16 || t == GDF_INT8 ||
         t == GDF_DATE32 || t == GDF_DATE64 || t == GDF_TIMESTAMP || t == GDF_STRING_CATEGORY;
}
bool is_float(int t) { return t == GDF_FLOAT32 || t == GDF_FLOAT64; }
bool is_signed_t(int t) {  // CalciteExpressionParsing.cpp:21-31
  return t == GDF_INT8 || t == GDF_INT16 || t == GDF_INT32 || t == GDF_INT64 ||
         t == GDF_FLOAT32 || t == GDF_FLOAT64 || t == GDF_DATE32 || t == GDF_DATE64 ||
         t == GDF_TIMESTAMP;
}
bool is_date(int t) { return t == GDF_DATE32 || t == GDF_DATE64 || t == GDF_TIMESTAMP; }
int wid(int t) { return blz_dtype_width((gdf_dtype)t); }

}  // namespace

int unary_output_dtype(int l, int u) {  // CalciteExpressionParsing.cpp:177-191
  if (u == BLZ_CAST_FLOAT64) return GDF_FLOAT64;
  if (is_date(l)) return GDF_INT16;
  if (u == BLZ_SIN || u == BLZ_COS || u == BLZ_ASIN || u == BLZ_ACOS || u == BLZ_TAN ||
      u == BLZ_COTAN || u == BLZ_ATAN || u == BLZ_LOG || u == BLZ_LN)
    return is_float(l) ? l : GDF_FLOAT64;
  if (u == BLZ_IS_NULL || u == BLZ_IS_NOT_NULL) return GDF_INT8;
  return l;
}

int binary_output_dtype(int l, int r, int op) {  // CalciteExpressionParsing.cpp:194-270
  if (!is_signed_t(l) || !is_signed_t(r)) return GDF_invalid;
  switch (op) {
    case GDF_ADD: case GDF_SUB: case GDF_MUL: case GDF_DIV: case GDF_MOD:
      if (is_float(l) || is_float(r)) {
        if (is_float(l) && is_float(r)) return wid(l) >= wid(r) ? l : r;
        return is_float(l) ? l : r;
      }
      return wid(l) >= wid(r) ? l : r;
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

// interpreter_ops.cuh:156-175 (scale_to_64_bit_return_bytes), see scalar_mode.
uint64_t scalar_bits(const gdf_scalar &s, int mode) {
  switch (s.dtype) {
    case GDF_INT8: return (uint64_t)(int64_t)s.data.si08;
    case GDF_INT16: return (uint64_t)(int64_t)s.data.si16;
    case GDF_INT32: case GDF_STRING_CATEGORY: return (uint64_t)(int64_t)s.data.si32;
    case GDF_INT64: return (uint64_t)s.data.si64;
    case GDF_DATE32: return (uint64_t)(int64_t)s.data.dt32;
    case GDF_DATE64: return (uint64_t)s.data.dt64;
    case GDF_TIMESTAMP: return (uint64_t)s.data.tmst;
    case GDF_FLOAT32: {
      if (mode == BLZ_SCALAR_SHIPPED) return (uint64_t)(int64_t)s.data.fp32;
      double d = (double)s.data.fp32;
      uint64_t b; std::memcpy(&b, &d, 8); return b;
    }
    case GDF_FLOAT64: {
      if (mode == BLZ_SCALAR_SHIPPED) return (uint64_t)(int64_t)s.data.fp64;
      uint64_t b; std::memcpy(&b, &s.data.fp64, 8); return b;
    }
    default: return 0;
  }
}

namespace {

int64_t host_f2i(double d) {
  if (d != d) return 0;
  if (d >= 9223372036854775807.0) return INT64_MAX;
  if (d <= -9223372036854775808.0) return INT64_MIN;
  return (int64_t)d;
}
uint64_t host_conv(uint64_t x, int cv) {
  if (cv == CV_I2F) { double d = (double)(int64_t)x; uint64_t b; std::memcpy(&b, &d, 8); return b; }
  if (cv == CV_F2I) { double d; std::memcpy(&d, &x, 8); return (uint64_t)host_f2i(d); }
  return x;
}

int load_kind(int dtype) {
  switch (dtype) {
    case GDF_INT8: return LD_I8;
    case GDF_INT16: return LD_I16;
    case GDF_INT32: case GDF_DATE32: case GDF_STRING_CATEGORY: return LD_I32;
    case GDF_INT64: case GDF_DATE64: case GDF_TIMESTAMP: return LD_I64;
    case GDF_FLOAT32: return LD_F32;
    case GDF_FLOAT64: return LD_F64;
    case GDF_BOOL8: return LD_U8;
    default: return LD_NONE;
  }
}

// write_data (interpreter_ops.cuh:317-381): BOOL8 and others are not written.
int store_kind_program(int dtype) {
  switch (dtype) {
    case GDF_INT8: return ST_I8;
    case GDF_INT16: return ST_I16;
    case GDF_INT32: case GDF_DATE32: case GDF_STRING_CATEGORY: return ST_I32;
    case GDF_INT64: case GDF_DATE64: case GDF_TIMESTAMP: return ST_I64;
    case GDF_FLOAT32: return ST_F32;
    case GDF_FLOAT64: return ST_F64;
    default: return ST_NONE;
  }
}
// A gathered column keeps its bytes (materialize_column copies by width).
int store_kind_alias(int dtype) {
  if (dtype == GDF_BOOL8) return ST_I8;
  return store_kind_program(dtype);
}

struct PlanResult {
  std::vector<int> finals;   // value node per final output
  std::vector<int> vfinals;  // validity value per final output
};

void fail(const std::string &m) { throw std::runtime_error(m); }

// Builds the SSA graph of one reference-format plan.
PlanResult build_plan(Builder &B, const PlanSpec &p, const std::vector<DevInputInfo> &dev_in,
                      int program, int scalar_mode, bool after_filter = false) {
  const int n_ops = (int)p.left.size();
  const int n_in = (int)p.in_dev.size();
  if ((int)p.right.size() != n_ops || (int)p.outs.size() != n_ops || (int)p.binops.size() != n_ops ||
      (int)p.unops.size() != n_ops || (int)p.lsc.size() != n_ops || (int)p.rsc.size() != n_ops)
    fail("plan vectors have different lengths");
  if (n_ops > kMaxOps) fail("plan has more than 256 operations");
  auto slot_ok = [](int s) { return s >= 0 && s < kMaxSlots; };
  for (int k = 0; k < n_ops; k++) {
    if (!slot_ok(p.outs[k])) fail("output register outside [0, 64)");
    if (p.left[k] >= kMaxSlots || p.right[k] >= kMaxSlots) fail("operand register outside [0, 64)");
  }
  for (int f : p.final_pos)
    if (!slot_ok(f)) fail("final output register outside [0, 64)");

  // Slot -> current value; plan input slot c holds dev input in_dev[c].
  std::vector<int> cur(kMaxSlots, -1), vcur(kMaxSlots, V_NONE);
  for (int c = 0; c < n_in; c++) {
    const DevInputInfo &d = dev_in[p.in_dev[c]];
    cur[c] = B.in_node(p.in_dev[c]);
    // process_valids input rule (interpreter_ops.cuh:1312-1327): forced all-valid
    // when there is no bitmap, null_count == 0, or the op with the column's index
    // is IS_NULL / IS_NOT_NULL. A projection after a filter reads the filtered
    // columns, whose null counts materialize_column recomputes from the kept rows'
    // bits (ColumnManipulation.cu:105-113): there the bitmap counts even when the
    // unfiltered column's null_count is 0 (forcing would change only rows whose
    // bit is clear, and those make the recomputed count non-zero).
    bool forced = d.valid == nullptr || (d.null_count == 0 && !after_filter) ||
                  (c < n_ops && (p.unops[c] == BLZ_IS_NULL || p.unops[c] == BLZ_IS_NOT_NULL));
    vcur[c] = forced ? V_ALL : B.vraw(p.in_dev[c]);
  }
  // Host type resolution (interpreter_ops.cuh:1095-1178).
  std::map<int, int> out_map;
  for (int k = 0; k < n_ops; k++) {
    const int L = p.left[k], R = p.right[k];
    int lt = GDF_invalid, rt = GDF_invalid;
    uint64_t lbits = 0, rbits = 0;
    if (L >= 0 && L < n_in) lt = dev_in[p.in_dev[L]].dtype;
    else if (L == SCALAR_NULL_INDEX) lt = p.lsc[k].dtype;
    else if (L == SCALAR_INDEX) {
      lt = is_int_kernel(p.lsc[k].dtype) ? GDF_INT64 : GDF_FLOAT64;
      lbits = scalar_bits(p.lsc[k], scalar_mode);
    } else if (L >= n_in) lt = out_map.count(L) ? out_map[L] : GDF_invalid;
    if (R >= 0 && R < n_in) rt = dev_in[p.in_dev[R]].dtype;
    else if (R == SCALAR_NULL_INDEX) rt = p.rsc[k].dtype;
    else if (R == SCALAR_INDEX) {
      rt = is_int_kernel(p.rsc[k].dtype) ? GDF_INT64 : GDF_FLOAT64;
      rbits = scalar_bits(p.rsc[k], scalar_mode);
    } else if (R >= n_in) rt = out_map.count(R) ? out_map[R] : GDF_invalid;
    const bool unary = R == BLZ_UNARY_INDEX;
    const int t = unary ? unary_output_dtype(lt, p.unops[k]) : binary_output_dtype(lt, rt, p.binops[k]);
    int ot = GDF_invalid;
    if (is_signed_t(t) && !is_float(t)) ot = GDF_INT64;
    else if (is_float(t)) ot = GDF_FLOAT64;
    out_map[p.outs[k]] = ot;
    const bool lk = is_float(lt), rk = is_float(rt), ok = is_float(ot);

    // Operands: slot value, scalar immediate, or 0 (null scalar / undefined).
    auto operand = [&](int idx, uint64_t bits) {
      Operand o;
      if (idx == SCALAR_INDEX) { o.imm = true; o.bits = bits; }
      else if (idx >= 0) {
        int v = cur[idx];
        if (v < 0) { o.imm = true; o.bits = 0; }  // never-written slot reads 0
        else o.node = v;
      } else { o.imm = true; o.bits = 0; }
      return o;
    };
    Node nd;
    nd.kind = 1;
    nd.program = program;
    nd.a = operand(L, lbits);
    if (!unary) nd.b = operand(R, rbits); else { nd.b.imm = true; nd.b.bits = 0; }
    int ca = CV_NONE, cb = CV_NONE, natural_f = 0, code = OP_IDENT;
    int vcol = 0xFF;
    auto in_col_of = [&](int idx) { return (idx >= 0 && idx < n_in) ? p.in_dev[idx] : 0xFF; };
    bool is_bool = false;
    if (!unary) {
      const bool fp = lk || rk;
      auto arith = [&](int ci, int cf) {
        code = fp ? cf : ci;
        if (fp) { ca = lk ? CV_NONE : CV_I2F; cb = rk ? CV_NONE : CV_I2F; }
        natural_f = fp;
      };
      auto cmp = [&](int ci, int cf) {
        code = fp ? cf : ci;
        if (fp) { ca = lk ? CV_NONE : CV_I2F; cb = rk ? CV_NONE : CV_I2F; }
        natural_f = 0;
        is_bool = true;
      };
      switch (p.binops[k]) {
        case GDF_ADD: arith(OP_ADD_I, OP_ADD_F); break;
        case GDF_SUB: arith(OP_SUB_I, OP_SUB_F); break;
        case GDF_MUL: arith(OP_MUL_I, OP_MUL_F); break;
        case GDF_DIV: case GDF_FLOOR_DIV: arith(OP_DIV_I, OP_DIV_F); break;
        case GDF_MOD:
          code = OP_MOD_I; ca = lk ? CV_F2I : CV_NONE; cb = rk ? CV_F2I : CV_NONE; natural_f = 0;
          break;
        case GDF_POW:
          if (is_float(lt) || is_float(rt)) {
            code = OP_POW_F; ca = lk ? CV_NONE : CV_I2F; cb = rk ? CV_NONE : CV_I2F; natural_f = 1;
          } else {
            code = OP_POW_I; natural_f = 0;
          }
          break;
        case GDF_EQUAL: cmp(OP_EQ_I, OP_EQ_F); break;
        case GDF_NOT_EQUAL: cmp(OP_NE_I, OP_NE_F); break;
        case GDF_LESS: cmp(OP_LT_I, OP_LT_F); break;
        case GDF_GREATER: cmp(OP_GT_I, OP_GT_F); break;
        case GDF_LESS_EQUAL: cmp(OP_LE_I, OP_LE_F); break;
        case GDF_GREATER_EQUAL: cmp(OP_GE_I, OP_GE_F); break;
        case GDF_COALESCE:
          code = OP_COALESCE;
          ca = (lk && !ok) ? CV_F2I : (!lk && ok) ? CV_I2F : CV_NONE;
          cb = (rk && !ok) ? CV_F2I : (!rk && ok) ? CV_I2F : CV_NONE;
          natural_f = ok;
          vcol = in_col_of(L);
          break;
        default:
          fail("binary operator without a kernel branch (op " + std::to_string(p.binops[k]) + ")");
      }
    } else {
      const int u = p.unops[k];
      auto math = [&](int c) { code = c; ca = lk ? CV_NONE : CV_I2F; natural_f = 1; };
      auto datepart = [&](int c32, int c64) {
        code = lt == GDF_DATE32 ? c32 : c64; ca = lk ? CV_F2I : CV_NONE; natural_f = 0;
      };
      switch (u) {
        case BLZ_FLOOR: math(OP_FLOOR); break;
        case BLZ_CEIL: math(OP_CEIL); break;
        case BLZ_SIN: math(OP_SIN); break;
        case BLZ_COS: math(OP_COS); break;
        case BLZ_ASIN: math(OP_ASIN); break;
        case BLZ_ACOS: math(OP_ACOS); break;
        case BLZ_TAN: math(OP_TAN); break;
        case BLZ_COTAN: math(OP_COTAN); break;
        case BLZ_ATAN: math(OP_ATAN); break;
        case BLZ_ABS: math(OP_ABS); break;
        case BLZ_LN: math(OP_LN); break;
        case BLZ_LOG: math(OP_LOG10); break;
        case BLZ_CAST_FLOAT64: math(OP_IDENT); break;
        case BLZ_NOT: code = lk ? OP_NOT_F : OP_NOT_I; natural_f = 0; is_bool = true; break;
        case BLZ_YEAR: datepart(OP_YEAR32, OP_YEAR64); break;
        case BLZ_MONTH: datepart(OP_MONTH32, OP_MONTH64); break;
        case BLZ_DAY: datepart(OP_DAY32, OP_DAY64); break;
        case BLZ_HOUR: datepart(OP_ZERO, OP_HOUR64); break;
        case BLZ_MINUTE: datepart(OP_ZERO, OP_MINUTE64); break;
        case BLZ_SECOND: datepart(OP_ZERO, OP_SECOND64); break;
        case BLZ_IS_NULL: code = OP_IS_NULL; vcol = in_col_of(L); natural_f = 0; is_bool = true; break;
        case BLZ_IS_NOT_NULL: code = OP_IS_NOT_NULL; vcol = in_col_of(L); natural_f = 0; is_bool = true; break;
        default: code = OP_IDENT; natural_f = lk; break;  // `computed = left_value`
      }
    }
    int co = CV_NONE;
    if (natural_f && !ok) co = CV_F2I;
    else if (!natural_f && ok) co = CV_I2F;
    if (co == CV_I2F) is_bool = false;
    // Immediates are converted here once.
    if (nd.a.imm && ca) { nd.a.bits = host_conv(nd.a.bits, ca); ca = CV_NONE; }
    if (nd.b.imm && cb) { nd.b.bits = host_conv(nd.b.bits, cb); cb = CV_NONE; }
    // AND of two 0/1 values (GDF_MUL from the `AND` token) is a bitwise and.
    if (code == OP_MUL_I && ca == CV_NONE && cb == CV_NONE && co == CV_NONE && !nd.a.imm &&
        !nd.b.imm && B.nodes[nd.a.node].is_bool && B.nodes[nd.b.node].is_bool) {
      code = OP_AND_B;
      is_bool = true;
    }
    nd.op.code = (uint8_t)code;
    nd.op.conv = (uint8_t)(ca | (cb << 2) | (co << 4));
    nd.op.vcol = (uint8_t)vcol;
    nd.is_bool = is_bool;
    B.nodes.push_back(nd);
    cur[p.outs[k]] = (int)B.nodes.size() - 1;

    // Validity (process_valids, interpreter_ops.cuh:507-571).
    auto vop = [&](int idx) {
      if (idx == SCALAR_INDEX) return V_ALL;
      if (idx >= 0) return vcur[idx];
      return V_NONE;  // null scalar (-3) and undefined operands
    };
    if (!unary) {
      const int vc = p.binops[k] == GDF_COALESCE ? VOP_OR : VOP_AND;
      vcur[p.outs[k]] = B.vcombine(vc, vop(L), vop(R), program);
    } else {
      vcur[p.outs[k]] = vop(L);
    }
  }
  PlanResult r;
  for (int f : p.final_pos) {
    r.finals.push_back(cur[f] >= 0 ? cur[f] : B.zero());
    r.vfinals.push_back(vcur[f]);
  }
  return r;
}

// Sethi-Ullman "need" of a value (inputs are resident: 0).
int compute_need(Builder &B, int id) {
  Node &n = B.nodes[id];
  if (n.kind != 1) return 0;
  if (n.need > 0) return n.need;
  int na = n.a.imm ? 0 : compute_need(B, n.a.node);
  int nb = n.b.imm ? 0 : compute_need(B, n.b.node);
  n.need = std::max(1, na == nb ? na + (na > 0 ? 1 : 0) : std::max(na, nb));
  return n.need;
}

void schedule(Builder &B, int id, std::vector<int> &order) {
  Node &n = B.nodes[id];
  if (n.order >= 0 || n.kind == 0) return;
  if (n.kind == 2) {  // constant zero gets a real op so it owns a slot
    n.order = (int)order.size();
    order.push_back(id);
    return;
  }
  int ca = n.a.imm ? -1 : n.a.node, cb = n.b.imm ? -1 : n.b.node;
  int na = ca >= 0 ? compute_need(B, ca) : 0, nb = cb >= 0 ? compute_need(B, cb) : 0;
  if (nb > na) std::swap(ca, cb);
  if (ca >= 0) schedule(B, ca, order);
  if (cb >= 0) schedule(B, cb, order);
  n.order = (int)order.size();
  order.push_back(id);
}

void mark_vneeded(Builder &B, int v) {
  if (v < 0) return;
  VNode &n = B.vnodes[v];
  if (n.needed) return;
  n.needed = true;
  if (n.kind == 1) { mark_vneeded(B, n.a); mark_vneeded(B, n.b); }
}

template <typename T>
void put(std::vector<uint8_t> &blob, size_t off, const T &v) {
  std::memcpy(blob.data() + off, &v, sizeof(T));
}

int flip_cmp(int c) {  // a OP b  ==  b FLIP(OP) a
  switch (c) {
    case OP_LT_I: return OP_GT_I; case OP_GT_I: return OP_LT_I;
    case OP_LE_I: return OP_GE_I; case OP_GE_I: return OP_LE_I;
    case OP_LT_F: return OP_GT_F; case OP_GT_F: return OP_LT_F;
    case OP_LE_F: return OP_GE_F; case OP_GE_F: return OP_LE_F;
    default: return c;
  }
}
bool is_cmp_code(int c) { return c >= OP_EQ_I && c <= OP_GE_F; }

// Multiplier and shift for truncating int64 division by a constant d (|d| >= 2):
// q = (mulhi(M, n) [+ n if d > 0 and M < 0] [- n if d < 0 and M > 0]) >> s, plus
// one if negative. The signed magic-number algorithm of Granlund & Montgomery
// (PLDI 1994) in the formulation of Warren, "Hacker's Delight", 2nd ed., 10-4.
void magic_signed(int64_t d, int64_t &M, int &s) {
  const uint64_t two63 = 1ull << 63;
  const uint64_t ad = d < 0 ? (uint64_t)0 - (uint64_t)d : (uint64_t)d;
  const uint64_t t = two63 + ((uint64_t)d >> 63);
  const uint64_t anc = t - 1 - t % ad;  // |nc|
  int p = 63;
  uint64_t q1 = two63 / anc, r1 = two63 - q1 * anc;  // 2^p / |nc|
  uint64_t q2 = two63 / ad, r2 = two63 - q2 * ad;    // 2^p / |d|
  uint64_t delta;
  do {
    p++;
    q1 *= 2;
    r1 *= 2;
    if (r1 >= anc) { q1++; r1 -= anc; }
    q2 *= 2;
    r2 *= 2;
    if (r2 >= ad) { q2++; r2 -= ad; }
    delta = ad - r2;
  } while (q1 < delta || (q1 == delta && r1 == 0));
  M = (int64_t)(q2 + 1);
  if (d < 0) M = -M;
  s = p - 64;
}

// Dispatch leaf of a lowered op, once the rows per lane R (and so the byte offsets
// of temporaries, 8 * 256 R apart) are known.
void assign_leaf(VmOp &op, int R) {
  const int tstride = 8 * 256 * R;
  if (op.cls == VC_TMP) op.off = op.arg2 * tstride;
  switch (op.code) {
    case VM_STORE: op.off = op.arg * tstride; op.leaf = kLeafStore; return;
    case VM_EMIT: op.leaf = kLeafEmit; return;
    case VM_I2F: op.leaf = kLeafI2F; return;
    case VM_NOT_I: op.leaf = kLeafNotI; return;
    case VM_NOT_F: op.leaf = kLeafNotF; return;
    case VM_DIV_I: case VM_MOD_I: {
      const int64_t d = (int64_t)op.imm;
      if (op.cls == VC_IMM && op.cv == CV_NONE && d != 0 && d != 1 && d != -1) {
        int64_t M;
        int sh;
        magic_signed(d, M, sh);
        op.imm2 = (uint64_t)M;
        op.gen = (uint8_t)sh;
        op.leaf = op.code == VM_DIV_I ? kLeafDivK : kLeafModK;
      } else {
        op.leaf = kLeafGeneral;
      }
      return;
    }
    default: break;
  }
  int hc = -1;
  switch (op.cls) {
    case VC_IMM: hc = HC_IMM; break;  // converted on the host
    case VC_TMP: hc = op.cv == CV_NONE ? HC_TMP : -1; break;
    case VC_L64: hc = op.cv == CV_NONE ? HC_L64 : op.cv == CV_I2F ? HC_L64_F : -1; break;
    case VC_L32I: hc = op.cv == CV_NONE ? HC_L32I : op.cv == CV_I2F ? HC_L32I_F : -1; break;
    case VC_L32F: hc = op.cv == CV_NONE ? HC_L32F : -1; break;
    default: break;
  }
  int ho = -1;
  switch (op.code) {
    case VM_LOAD: ho = HO_LOAD; break;
    case VM_ADD_I: ho = HO_ADD_I; break;
    case VM_ADD_F: ho = HO_ADD_F; break;
    case VM_SUB_I: ho = HO_SUB_I; break;
    case VM_SUB_F: ho = HO_SUB_F; break;
    case VM_RSUB_I: ho = HO_RSUB_I; break;
    case VM_RSUB_F: ho = HO_RSUB_F; break;
    case VM_MUL_I: ho = HO_MUL_I; break;
    case VM_MUL_F: ho = HO_MUL_F; break;
    case VM_AND_B: ho = HO_AND_B; break;
    case VM_CMP: case VM_CMPAND:
      for (int i = 0; i < 12; i++)
        if (kHotCmp[i] == op.arg) ho = (op.code == VM_CMP ? HO_CMP : HO_CMPAND) + i;
      break;
    case VM_RANGE_I: ho = HO_RANGE_I; break;
    case VM_RANGE_F: ho = HO_RANGE_F; break;
    default: break;
  }
  op.leaf = hc >= 0 && ho >= 0 ? ho * HC_COUNT_ + hc : kLeafGeneral;
}

// Folds the accumulator moves of an op stream into the ops beside them (kFusePre,
// kFuseStore, kFuseEmit): a LOAD of a temporary or a tile column into the op after
// it, a STORE or an EMIT into the op before it. A LOAD of an integer column followed
// by I2F becomes one LOAD of the column widened to binary64 first. The values every
// op computes are unchanged; only the number of dispatches falls.
std::vector<VmOp> fuse_moves(const std::vector<VmOp> &in) {
  auto target = [](const VmOp &o) {
    const int l = o.leaf & kLeafMask;
    return l < kHotLeaves || (l >= kLeafI2F && l <= kLeafModK);
  };
  auto is_load = [](const VmOp &o) {
    const int l = o.leaf;  // no flags yet
    return l < kHotLeaves && l / HC_COUNT_ == HO_LOAD && l % HC_COUNT_ != HC_IMM;
  };
  // LOAD of an integer column, then I2F: one LOAD of the widened column
  std::vector<VmOp> a;
  for (size_t k = 0; k < in.size(); k++) {
    VmOp op = in[k];
    if (is_load(op) && k + 1 < in.size() && in[k + 1].leaf == kLeafI2F) {
      const int hc = op.leaf % HC_COUNT_;
      if (hc == HC_L64 || hc == HC_L32I) {
        op.leaf = HO_LOAD * HC_COUNT_ + (hc == HC_L64 ? HC_L64_F : HC_L32I_F);
        op.cv = CV_I2F;
        a.push_back(op);
        k++;
        continue;
      }
    }
    a.push_back(op);
  }
  // LOAD into the next op
  std::vector<VmOp> b;
  for (size_t k = 0; k < a.size(); k++) {
    const VmOp &op = a[k];
    if (is_load(op) && k + 1 < a.size() && target(a[k + 1]) && (uint32_t)op.off < (1u << 20)) {
      VmOp nx = a[k + 1];
      nx.leaf |= kFusePre;
      const uint32_t w = (uint32_t)op.off | ((uint32_t)(op.leaf % HC_COUNT_) << 20);
      std::memcpy(&nx.code, &w, 4);  // code, arg, cls, cv
      b.push_back(nx);
      k++;
      continue;
    }
    b.push_back(op);
  }
  // STORE / EMIT into the previous op
  std::vector<VmOp> c;
  for (const VmOp &op : b) {
    if (!c.empty() && target(c.back())) {
      VmOp &pv = c.back();
      if (op.leaf == kLeafStore && !(pv.leaf & kFuseStore) && op.arg < 256) {
        uint32_t w;
        std::memcpy(&w, &pv.code, 4);
        if (!(pv.leaf & kFusePre)) w = 0;
        w = (w & 0x00FFFFFFu) | ((uint32_t)op.arg << 24);
        std::memcpy(&pv.code, &w, 4);
        pv.leaf |= kFuseStore;
        continue;
      }
      if (op.leaf == kLeafEmit && !(pv.leaf & kFuseEmit)) {
        pv.vcol = op.arg;   // gen word bits 8-15: output
        pv.arg2 = op.gen;   // bits 16-23: StoreKind
        pv.leaf |= kFuseEmit;
        continue;
      }
    }
    c.push_back(op);
  }
  return c;
}

// Conjunctions of comparisons (filters such as a date range, a discount range and
// a quantity bound): a predicate opening with LOAD x; CMP imm opens with
// CMPAND(x CMP imm) instead (the predicate kernel starts with acc = 1), the
// CMPANDs of a run are grouped by operand (AND of 0/1 values commutes), and a
// lower and an upper bound on one operand become one range op.
bool cmp_is_float(int sub) {
  return sub == OP_LT_F || sub == OP_LE_F || sub == OP_GT_F || sub == OP_GE_F || sub == OP_EQ_F || sub == OP_NE_F;
}
bool cmp_lower(int sub) { return sub == OP_GT_I || sub == OP_GE_I || sub == OP_GT_F || sub == OP_GE_F; }
bool cmp_upper(int sub) { return sub == OP_LT_I || sub == OP_LE_I || sub == OP_LT_F || sub == OP_LE_F; }
void fuse_conjunctions(std::vector<VmOp> &v, bool predicate) {
  auto has_operand = [](const VmOp &o) { return o.cls != VC_NONE && o.cls != VC_IMM; };
  if (predicate && v.size() >= 2 && v[0].code == VM_LOAD && has_operand(v[0]) && v[1].code == VM_CMP &&
      v[1].cls == VC_IMM && is_cmp_code(v[1].arg)) {
    VmOp c = v[0];
    c.code = VM_CMPAND;
    c.arg = v[1].arg;
    c.imm2 = v[1].imm;
    c.imm = 0;
    v.erase(v.begin());
    v[0] = c;
  }
  auto key = [](const VmOp &o) { return (o.cls << 16) | (o.arg2 << 8) | o.cv; };
  for (size_t i = 0; i < v.size();) {
    size_t j = i;
    while (j < v.size() && v[j].code == VM_CMPAND) j++;
    if (j - i >= 2) {
      std::stable_sort(v.begin() + i, v.begin() + j, [&](const VmOp &a, const VmOp &b) { return key(a) < key(b); });
      for (size_t k = i; k + 1 < j; k++) {
        VmOp &a = v[k], &b = v[k + 1];
        if (a.code != VM_CMPAND || b.code != VM_CMPAND || key(a) != key(b) || !has_operand(a)) continue;
        if (cmp_is_float(a.arg) != cmp_is_float(b.arg)) continue;
        const VmOp *lo = cmp_lower(a.arg) && cmp_upper(b.arg) ? &a : cmp_lower(b.arg) && cmp_upper(a.arg) ? &b : nullptr;
        if (!lo) continue;
        const VmOp *hi = lo == &a ? &b : &a;
        VmOp r = a;
        if (cmp_is_float(a.arg)) {
          r.code = VM_RANGE_F;
          r.imm2 = lo->imm2;
          r.imm = hi->imm2;
          r.gen = (uint8_t)((lo->arg == OP_GE_F ? 1 : 0) | (hi->arg == OP_LE_F ? 2 : 0));
        } else {
          int64_t l = (int64_t)lo->imm2, h = (int64_t)hi->imm2;
          if (lo->arg == OP_GE_I) {
            if (l == INT64_MIN) continue;
            l -= 1;
          }
          if (hi->arg == OP_LE_I) {
            if (h == INT64_MAX) continue;
            h += 1;
          }
          r.code = VM_RANGE_I;
          r.imm2 = (uint64_t)l;
          r.imm = (uint64_t)h;
        }
        a = r;
        b.code = VM_COUNT_;  // removed below
        k++;
      }
      v.erase(std::remove_if(v.begin() + i, v.begin() + j, [](const VmOp &o) { return o.code == VM_COUNT_; }),
              v.begin() + j);
      j = i;
      while (j < v.size() && (v[j].code == VM_CMPAND || v[j].code == VM_RANGE_I || v[j].code == VM_RANGE_F)) j++;
    }
    i = j > i ? j : i + 1;
  }
}

// Lowers the accumulator form to VmOps and lays out the tile buffer (kTileBuffers).
// Returns false (with the reason) when the plan does not fit the tile VM.
bool compile_vm(const CompileRequest &req, Builder &B, const std::vector<AccOp> &acc_p,
                const std::vector<AccOp> &acc_q, int acc_temps, const std::vector<VOp> &vops_p,
                const std::vector<VOp> &vops_q, const std::vector<int> &out_node, const std::vector<int> &out_v,
                int stencil_v, int n_vslots, Compiled &out, std::string &why) {
  const int n_dev = (int)req.inputs.size();
  const int n_out = (int)req.outputs.size();
  if (n_vslots > kVmVSlots) { why = "plan needs more than 32 validity registers"; return false; }
  if (req.n_rows > (int64_t)INT32_MAX * 256) { why = "table too large"; return false; }

  // Which inputs the predicate reads (staged) and which only the projection of a
  // filter reads (fetched from HBM for kept rows).
  std::vector<bool> used_p(n_dev, false), used_q(n_dev, false);
  for (auto &a : acc_p) if (a.okind == OK_INPUT) used_p[a.oidx] = true;
  for (auto &a : acc_q) if (a.okind == OK_INPUT) used_q[a.oidx] = true;
  std::vector<bool> staged(n_dev, false);
  for (int c = 0; c < n_dev; c++) staged[c] = used_p[c] || (used_q[c] && !req.compact);
  for (int c = 0; c < n_dev; c++) {
    if (staged[c] && (uintptr_t)req.inputs[c].data % 16) { why = "unaligned column"; return false; }
    if (B.vnodes[c].needed && (uintptr_t)req.inputs[c].valid % 16) { why = "unaligned bitmap"; return false; }
  }

  bool gather_pass = false;  // lowering the projection of a filter: operands from HBM
  auto operand = [&](VmOp &v, const AccOp &a, int conv_b) {
    switch (a.okind) {
      case OK_IMM: v.cls = VC_IMM; v.imm = host_conv(a.imm, conv_b); return;
      case OK_TEMP: v.cls = VC_TMP; v.arg2 = a.oidx; v.cv = (uint8_t)conv_b; return;
      case OK_INPUT: {
        const int c = a.oidx;
        const int lk = load_kind(req.inputs[c].dtype);
        const bool g = !gather_pass && !staged[c];
        v.arg2 = (uint8_t)c;
        v.cv = (uint8_t)conv_b;
        switch (lk) {
          case LD_I64: case LD_F64: v.cls = g ? VC_G64 : VC_L64; break;
          case LD_I32: v.cls = g ? VC_G32I : VC_L32I; break;
          case LD_F32: v.cls = g ? VC_G32F : VC_L32F; break;
          case LD_I16: v.cls = g ? VC_G16I : VC_L16I; break;
          case LD_I8: v.cls = g ? VC_G8I : VC_L8I; break;
          case LD_U8: v.cls = g ? VC_G8U : VC_L8U; break;
          default: v.cls = VC_IMM; v.imm = host_conv(0, conv_b); v.cv = 0; break;  // never read: 0
        }
        return;
      }
      default: v.cls = VC_IMM; v.imm = 0; return;
    }
  };
  auto new_vm = [](int code) { VmOp v{}; v.code = (uint8_t)code; v.vcol = 0xFF; return v; };
  bool slow = false, vcol_needed = false;
  auto conv_op = [&](std::vector<VmOp> &dst, int cv) {
    if (cv == CV_I2F) dst.push_back(new_vm(VM_I2F));
    else if (cv == CV_F2I) dst.push_back(new_vm(VM_F2I));
  };
  auto lower = [&](const std::vector<AccOp> &src, std::vector<VmOp> &dst) {
    for (const AccOp &a : src) {
      const int ca = a.conv & 3, cb = (a.conv >> 2) & 3, co = (a.conv >> 4) & 3;
      switch (a.code) {
        case A_LOAD: { VmOp v = new_vm(VM_LOAD); operand(v, a, cb); dst.push_back(v); break; }
        case A_STORE: { VmOp v = new_vm(VM_STORE); v.arg = a.oidx; dst.push_back(v); break; }
        case A_EMIT: {
          const OutSpec &s = req.outputs[a.oidx];
          VmOp v = new_vm(VM_EMIT);
          v.arg = a.oidx;
          v.cls = VC_NONE;
          v.gen = (uint8_t)(s.kind == 0 ? store_kind_program(s.dtype) : store_kind_alias(s.dtype));
          dst.push_back(v);
          break;
        }
        case A_CMPAND: {
          VmOp v = new_vm(VM_CMPAND);
          operand(v, a, cb);
          v.arg = a.sub;
          v.imm2 = a.imm2;
          dst.push_back(v);
          break;
        }
        case A_BIN: {
          const int sub = a.sub;
          const bool rev = a.flags & F_REV;
          int hot = -1, harg = 0;
          switch (sub) {
            case OP_ADD_I: hot = VM_ADD_I; break;
            case OP_ADD_F: hot = VM_ADD_F; break;
            case OP_SUB_I: hot = rev ? VM_RSUB_I : VM_SUB_I; break;
            case OP_SUB_F: hot = rev ? VM_RSUB_F : VM_SUB_F; break;
            case OP_MUL_I: hot = VM_MUL_I; break;
            case OP_MUL_F: hot = VM_MUL_F; break;
            case OP_AND_B: hot = VM_AND_B; break;
            case OP_DIV_F: hot = rev ? VM_RDIV_F : VM_DIV_F; break;
            case OP_DIV_I: hot = rev ? VM_RDIV_I : VM_DIV_I; break;
            case OP_MOD_I: hot = rev ? VM_RMOD_I : VM_MOD_I; break;
            case OP_COALESCE: hot = rev ? VM_RCOALESCE : VM_COALESCE; break;
            default:
              if (is_cmp_code(sub)) { hot = VM_CMP; harg = rev ? flip_cmp(sub) : sub; }
          }
          if (hot == VM_COALESCE || hot == VM_RCOALESCE) vcol_needed = true;
          if (hot >= 0) {
            conv_op(dst, ca);
            VmOp v = new_vm(hot);
            operand(v, a, cb);
            v.arg = (uint8_t)harg;
            v.vcol = a.vcol;
            dst.push_back(v);
            conv_op(dst, co);
          } else {
            VmOp v = new_vm(VM_GENERIC);
            operand(v, a, 0);
            v.cv = 0;
            v.arg = (uint8_t)sub;
            v.gen = (uint8_t)(cb | (ca << 2) | (co << 4) | (rev ? VG_REV : 0));
            v.vcol = a.vcol;
            dst.push_back(v);
            slow = true;
          }
          break;
        }
        case A_UN: {
          const int sub = a.sub;
          int hot = -1;
          switch (sub) {
            case OP_FLOOR: hot = VM_FLOOR; break;
            case OP_CEIL: hot = VM_CEIL; break;
            case OP_ABS: hot = VM_ABS_F; break;
            case OP_NOT_I: hot = VM_NOT_I; break;
            case OP_NOT_F: hot = VM_NOT_F; break;
            case OP_IS_NULL: hot = VM_IS_NULL; break;
            case OP_IS_NOT_NULL: hot = VM_IS_NOT_NULL; break;
            case OP_IDENT: hot = VM_COUNT_; break;
            case OP_ZERO: hot = VM_LOAD; break;
            default: break;
          }
          if (hot >= 0) {
            conv_op(dst, ca);
            if (hot == VM_LOAD) { VmOp v = new_vm(VM_LOAD); v.cls = VC_IMM; v.imm = 0; dst.push_back(v); }
            else if (hot != VM_COUNT_) { VmOp v = new_vm(hot); v.vcol = a.vcol; dst.push_back(v); }
            conv_op(dst, co);
          } else {
            VmOp v = new_vm(VM_GENERIC);
            v.cls = VC_NONE;
            v.arg = (uint8_t)sub;
            v.gen = (uint8_t)((ca << 2) | (co << 4) | VG_UNARY);
            v.vcol = a.vcol;
            dst.push_back(v);
            slow = true;
          }
          break;
        }
        default: why = "unknown accumulator op"; return false;
      }
    }
    return true;
  };
  std::vector<VmOp> vm_p, vm_q;
  if (!lower(acc_p, vm_p)) return false;
  gather_pass = req.compact;
  if (!lower(acc_q, vm_q)) return false;
  fuse_conjunctions(vm_p, true);
  fuse_conjunctions(vm_q, false);
  // The gather pass of a filter stages the kept rows of the columns its ops read
  // into a pass buffer (256 R rows x each width) before evaluating them.
  std::vector<int> goff(n_dev, -1);

  // Tile buffers: staged column slices, then needed bitmaps, each 16-byte aligned.
  int n_valid_out = 0;
  for (auto &o : req.outputs) n_valid_out += o.valid != nullptr;
  std::vector<int> lds_off(n_dev, 0), lds_voff(n_dev, -1);
  auto layout = [&](int t, int &buf) {
    size_t off = 0;
    for (int c = 0; c < n_dev; c++) {
      lds_off[c] = 0;
      if (!staged[c]) continue;
      lds_off[c] = (int)off;
      off += (size_t)t * std::max(1, wid(req.inputs[c].dtype));
      off = (off + 15) & ~size_t(15);
    }
    for (int c = 0; c < n_dev; c++) {
      lds_voff[c] = -1;
      if (!B.vnodes[c].needed) continue;
      lds_voff[c] = (int)off;
      off += (size_t)t / 8;
      off = (off + 15) & ~size_t(15);
    }
    buf = (int)std::max<size_t>(off, 16);
    return (size_t)kTileBuffers * buf;
  };
  // Shape: rows per lane R and tile rows T = 256 R P (P passes of the four waves).
  // Temporaries live in LDS (256 R rows each), after the validity registers (four
  // waves x registers x R words). The default is R = 4 and one pass
  // per tile (T = 1024): a small tile buffer keeps two workgroups per CU, whose
  // loads and evaluations overlap; the autotuner may choose otherwise.
  static const int shapes[] = {4, 8, 2};
  const int budget = std::min(req.lds_budget > 0 ? req.lds_budget : 64 * 1024, 160 * 1024 - 4096);
  int R = 0, T = 0, buf = 0;
  size_t lds_total = 0;
  int tmp_bytes = 0;
  auto al16 = [](int x) { return (x + 15) & ~15; };
  for (int r : shapes) {
    if (req.force_r > 0 && req.force_r != r) continue;
    const int tb = acc_temps * 256 * r * 8 + 4 * n_vslots * r * 8;
    int g = 0;  // gather pass buffers of a filter (four waves' regions of 64 r rows)
    for (int c = 0; c < n_dev && req.compact; c++)
      if (used_q[c]) g = al16(g + 64 * r * std::max(1, wid(req.inputs[c].dtype)));
    g *= 4;
    // default passes per tile: the fewest (P >= 1) that stage at least 16 KiB per
    // tile, so narrow programs do not pay a tile's fixed costs for a few KiB
    int b1 = 0;
    const size_t one = layout(256 * r, b1);  // one pass
    int pmax = 1;
    while (pmax < 64 / r && one / kTileBuffers * pmax < 16384) pmax *= 2;
    for (int p = 64 / r; p >= 1; p /= 2) {
      const int t = 256 * r * p;
      if (req.force_t > 0 ? t != req.force_t : p > pmax) continue;
      int b = 0;
      const size_t tot = layout(t, b) + tb;
      if (tot > (size_t)budget) continue;
      const int gl = g + tb + 16;
      if (req.compact && gl > budget) continue;
      R = r; T = t; buf = b; lds_total = tot; tmp_bytes = tb;
      break;
    }
    if (R) break;
  }
  if (R == 0) { why = "no tile VM shape fits the plan"; return false; }
  const int NT = acc_temps;
  layout(T, buf);
  for (auto &op : vm_p)
    if (op.cls >= VC_L64 && op.cls <= VC_L8U) op.off = lds_off[op.arg2];
  std::vector<Piece> gcols;
  int gbuf = 0;
  if (req.compact) {
    for (int c = 0; c < n_dev; c++) {
      if (!used_q[c]) continue;
      const int w = std::max(1, wid(req.inputs[c].dtype));
      goff[c] = gbuf;  // within one wave's region of 64 R rows
      gcols.push_back(Piece{req.inputs[c].data, 0, 0, gbuf, 0, 0, w});
      gbuf += 64 * R * w;
      gbuf = (gbuf + 15) & ~15;
    }
  }
  for (auto &op : vm_q)
    if (op.cls >= VC_L64 && op.cls <= VC_L8U) op.off = req.compact ? goff[op.arg2] : lds_off[op.arg2];
  for (auto *v : {&vm_p, &vm_q})
    for (auto &op : *v) assign_leaf(op, R);
  // A conjunction of column-versus-constant comparisons (every predicate op a
  // compare-and / range leaf over a column) runs from registers. Its validity
  // must be an AND of raw column bitmaps (columns the input rule forces valid
  // are absent from it); each bitmap is applied by one term over that column,
  // named in the term's vcol (which compare-and / range leaves otherwise ignore).
  bool direct = req.compact && req.pred != nullptr && !vm_p.empty() && vm_p.size() <= 3;
  for (const auto &op : vm_p) {
    const int ho = op.leaf / HC_COUNT_, cl = op.leaf % HC_COUNT_;
    if (op.leaf >= kHotLeaves || ho < HO_CMPAND || ho > HO_RANGE_F || cl < HC_L64) direct = false;
  }
  if (direct) {
    std::vector<int> todo{stencil_v};
    while (direct && !todo.empty()) {
      const int v = todo.back();
      todo.pop_back();
      if (v == V_ALL) continue;
      if (v < 0) { direct = false; break; }
      const VNode &vn = B.vnodes[v];
      if (vn.kind == 1) {
        if (vn.code != VOP_AND) direct = false;
        todo.push_back(vn.a);
        todo.push_back(vn.b);
        continue;
      }
      bool placed = false;
      for (auto &op : vm_p)
        if (op.arg2 == vn.aux && (op.vcol == 0xFF || op.vcol == vn.aux)) { op.vcol = (uint8_t)vn.aux; placed = true; break; }
      if (!placed) direct = false;
    }
  }
  out.vm_direct_pred = direct;
  out.vm_direct_terms = direct ? (int)vm_p.size() : 0;
  out.vm_direct_wmask = 0;
  for (size_t k = 0; direct && k < vm_p.size(); k++) {
    const int cl = vm_p[k].leaf % HC_COUNT_;
    if (cl == HC_L64 || cl == HC_L64_F) out.vm_direct_wmask |= 1 << k;
  }
  // the register-load predicate reads its terms as they are
  if (!direct) vm_p = fuse_moves(vm_p);
  vm_q = fuse_moves(vm_q);

  std::vector<Piece> pieces;
  for (int c = 0; c < n_dev; c++) {
    if (!staged[c]) continue;
    const int w = std::max(1, wid(req.inputs[c].dtype));
    const int slice = T * w;
    for (int off = 0; off < slice; off += kPieceBytes)
      pieces.push_back(Piece{req.inputs[c].data, slice, off, lds_off[c] + off,
                             std::min(kPieceBytes, slice - off), 0, w});
  }
  for (int c = 0; c < n_dev; c++) {
    if (lds_voff[c] < 0) continue;
    const int slice = T / 8;
    for (int off = 0; off < slice; off += kPieceBytes)
      pieces.push_back(Piece{req.inputs[c].valid, slice, off, lds_voff[c] + off,
                             std::min(kPieceBytes, slice - off), 1, 0});
  }

  // Blob: header | InCol | OutCol | VmOp | Piece | VOp.
  auto al = [](size_t x) { return (x + 15) & ~size_t(15); };
  const size_t off_in = al(sizeof(ProgHeader));
  const size_t off_out = al(off_in + sizeof(InCol) * n_dev);
  const size_t off_ops = al(off_out + sizeof(OutCol) * n_out);
  // one descriptor past the last op: the kernels load op k + 1 while op k runs
  const size_t off_pieces = al(off_ops + sizeof(VmOp) * (vm_p.size() + vm_q.size() + 1));
  // bare-column outputs (no ops: the projection / gather kernel copies them)
  std::vector<CopyCol> copies;
  for (int o = 0; o < n_out; o++) {
    const OutSpec &s = req.outputs[o];
    if (s.kind != 1 || store_kind_alias(s.dtype) == ST_NONE) continue;
    copies.push_back(CopyCol{req.inputs[s.index].data, s.data, std::max(1, wid(s.dtype)), 0, 0});
  }
  const size_t off_gcols = al(off_pieces + sizeof(Piece) * pieces.size());
  const size_t off_copy = al(off_gcols + sizeof(Piece) * gcols.size());
  const size_t off_vops = al(off_copy + sizeof(CopyCol) * copies.size());
  const size_t total = al(off_vops + sizeof(VOp) * (vops_p.size() + vops_q.size()));
  if ((int)total > blob_capacity()) { why = "plan too large for the program buffer"; return false; }
  auto vsrc = [&](int v) { return v == V_ALL ? -1 : v == V_NONE ? -2 : B.vnodes[v].vslot; };

  out.blob.assign(total, 0);
  ProgHeader h{};
  h.n_rows = req.n_rows;
  h.n_in = n_dev;
  h.n_out = n_out;
  h.n_ops_p = (int)vm_p.size();
  h.n_ops_q = (int)vm_q.size();
  h.n_vops_p = (int)vops_p.size();
  h.n_vops_q = (int)vops_q.size();
  h.n_vin = n_dev;
  h.n_slots = 0;
  h.n_vslots = n_vslots;
  h.stencil_slot = 0;
  // The predicate's value is the INT8 stencil column the reference writes
  // (narrowed); a bare-column stencil is the widened column value.
  h.stencil_width = (req.compact && !req.pred) ? 8 : 1;
  h.stencil_vsrc = vsrc(stencil_v);
  int any_valid = 0;
  for (auto &o : req.outputs) any_valid |= o.valid != nullptr;
  h.any_valid_out = any_valid;
  h.tile_rows = T;
  h.n_chunks = 0;
  // gather kernel LDS: pass buffers (four waves') | validity registers |
  // temporaries | chunk prefix (one program: two int64)
  const int g_kw = 0;
  const int g_buf = 0;
  const int vr_bytes = 4 * n_vslots * R * 8;
  const int g_vr = g_buf + 4 * gbuf;
  const int g_tmp = g_vr + vr_bytes;
  const int g_cpre = (g_vr + tmp_bytes + 7) & ~7;  // tmp_bytes: validity registers + temporaries
  h.lds_vbits_off = 0;
  h.n_gcols = (int)gcols.size();
  h.off_gcols = (int)off_gcols;
  h.n_copy = (int)copies.size();
  h.off_copy = (int)off_copy;
  h.gbuf_bytes = gbuf;
  h.g_kw_off = g_kw;
  h.g_buf_off = g_buf;
  h.g_tmp_off = g_tmp;
  h.lds_tmp_off = kTileBuffers * buf + vr_bytes;
  h.n_valid_out = n_valid_out;
  h.lds_bytes = (int)lds_total;
  h.acc_mode = 2;
  h.n_temps = acc_temps;
  h.buf_bytes = buf;
  h.n_pieces = (int)pieces.size();
  h.off_pieces = (int)off_pieces;
  h.off_chunks = (int)off_pieces;
  h.off_in = (int)off_in;
  h.off_out = (int)off_out;
  h.off_ops = (int)off_ops;
  h.off_vops = (int)off_vops;
  h.blob_bytes = (int)total;
  put(out.blob, 0, h);
  for (int c = 0; c < n_dev; c++) {
    InCol ic{};
    ic.data = req.inputs[c].data;
    ic.valid = req.inputs[c].valid;
    ic.load = load_kind(req.inputs[c].dtype);
    ic.vslot = B.vnodes[c].needed ? B.vnodes[c].vslot : -1;
    ic.lds_off = lds_off[c];
    ic.lds_voff = lds_voff[c];
    put(out.blob, off_in + sizeof(InCol) * c, ic);
  }
  int vbuf_next = 0;
  for (int o = 0; o < n_out; o++) {
    const OutSpec &s = req.outputs[o];
    OutCol oc{};
    oc.data = s.data;
    oc.valid = s.valid;
    oc.store = s.kind == 0 ? store_kind_program(s.dtype) : store_kind_alias(s.dtype);
    oc.slot = 0;
    oc.src_kind = (int16_t)s.kind;
    oc.vsrc = vsrc(out_v[o]);
    oc.vbuf = s.valid ? vbuf_next++ : -1;
    put(out.blob, off_out + sizeof(OutCol) * o, oc);
  }
  size_t p = off_ops;
  for (auto &op : vm_p) { put(out.blob, p, op); p += sizeof(VmOp); }
  for (auto &op : vm_q) { put(out.blob, p, op); p += sizeof(VmOp); }
  for (size_t k = 0; k < pieces.size(); k++) put(out.blob, off_pieces + sizeof(Piece) * k, pieces[k]);
  for (size_t k = 0; k < gcols.size(); k++) put(out.blob, off_gcols + sizeof(Piece) * k, gcols[k]);
  for (size_t k = 0; k < copies.size(); k++) put(out.blob, off_copy + sizeof(CopyCol) * k, copies[k]);
  p = off_vops;
  for (auto &op : vops_p) { put(out.blob, p, op); p += sizeof(VOp); }
  for (auto &op : vops_q) { put(out.blob, p, op); p += sizeof(VOp); }
  out.S = NT;
  out.R = R;
  out.n_slots = 0;
  out.n_vslots = n_vslots;
  out.n_ops = (int)(vm_p.size() + vm_q.size());
  out.n_vops = (int)(vops_p.size() + vops_q.size());
  out.T = T;
  out.lds_bytes = (int)lds_total;
  out.acc = false;
  out.vm = true;
  // op set level: 2 with out-of-line ops, 1 with hot ops outside the common set
  int level = slow ? 2 : 0;
  for (auto *v : {&vm_p, &vm_q})
    for (auto &op : *v)
      if (level == 0 && op.code > VM_I2F && op.code != VM_GENERIC && op.leaf == kLeafGeneral) level = 1;
  out.vm_slow = level;
  out.lds_gather = g_cpre + 16;
  out.vm_g_cpre_off = g_cpre;
  out.vm_bufb = buf;
  out.vm_vr_off = kTileBuffers * buf;
  out.vm_g_vr_off = g_vr;
  out.vm_tmp_off = kTileBuffers * buf + vr_bytes;
  out.vm_g_buf_off = g_buf;
  out.vm_g_wave_bytes = gbuf;
  out.vm_nvout = n_valid_out;
  out.vm_g_tmp_off = g_tmp;
  (void)vcol_needed;
  return true;
}

}  // namespace

int column_load_kind(int dtype) { return load_kind(dtype); }

Compiled compile_program(const CompileRequest &req) {
  Compiled out;
  out.n_rows = req.n_rows;
  const int n_dev = (int)req.inputs.size();
  if (n_dev > kMaxSlots) fail("more than 64 input columns");
  Builder B;
  for (int c = 0; c < n_dev; c++) {
    Node n; n.kind = 0; n.aux = c;
    B.nodes.push_back(n);
    B.input_node.push_back(c);
    VNode v; v.kind = 0; v.aux = c;
    B.vnodes.push_back(v);
    B.input_vnode.push_back(c);
  }
  PlanResult pr, qr;
  if (req.compact && req.pred) pr = build_plan(B, *req.pred, req.inputs, 0, req.scalar_mode);
  if (req.proj) qr = build_plan(B, *req.proj, req.inputs, 1, req.scalar_mode, req.compact);

  // Roots.
  int stencil_node = -1, stencil_v = V_ALL;
  if (req.compact) {
    if (req.pred) {
      if (pr.finals.size() != 1) fail("predicate plan must have exactly one output");
      stencil_node = pr.finals[0];
      stencil_v = pr.vfinals[0];
    } else {
      if (req.bare_stencil_input < 0 || req.bare_stencil_input >= n_dev) fail("bad stencil column");
      stencil_node = B.in_node(req.bare_stencil_input);
      stencil_v = req.inputs[req.bare_stencil_input].valid ? B.vraw(req.bare_stencil_input) : V_ALL;
    }
  }
  std::vector<int> out_node(req.outputs.size(), -1), out_v(req.outputs.size(), V_ALL);
  for (size_t o = 0; o < req.outputs.size(); o++) {
    const OutSpec &s = req.outputs[o];
    if (s.kind == 0) {
      if (!req.proj || s.index < 0 || s.index >= (int)qr.finals.size()) fail("bad program output");
      out_node[o] = qr.finals[s.index];
      out_v[o] = qr.vfinals[s.index];
    } else {
      if (s.index < 0 || s.index >= n_dev) fail("bad gathered column");
      out_node[o] = B.in_node(s.index);
      out_v[o] = req.inputs[s.index].valid ? B.vraw(s.index) : V_ALL;
    }
  }

  // Schedule: predicate first, then the projection outputs in order.
  std::vector<int> order_p, order_q;
  if (stencil_node >= 0) schedule(B, stencil_node, order_p);
  for (size_t o = 0; o < req.outputs.size(); o++)
    if (req.outputs[o].kind == 0) schedule(B, out_node[o], order_q);

  // Liveness: position in the global op sequence; roots live to their program's end.
  std::vector<int> seq = order_p;
  seq.insert(seq.end(), order_q.begin(), order_q.end());
  const int end_p = (int)order_p.size() - 1, end_q = (int)seq.size() - 1;
  for (int i = 0; i < (int)seq.size(); i++) {
    Node &n = B.nodes[seq[i]];
    n.order = i;
    if (!n.a.imm && n.a.node >= 0) B.nodes[n.a.node].last_use = std::max(B.nodes[n.a.node].last_use, i);
    if (!n.b.imm && n.b.node >= 0) B.nodes[n.b.node].last_use = std::max(B.nodes[n.b.node].last_use, i);
  }
  if (stencil_node >= 0 && B.nodes[stencil_node].kind != 0)
    B.nodes[stencil_node].last_use = std::max(B.nodes[stencil_node].last_use, end_p + 1);
  for (size_t o = 0; o < req.outputs.size(); o++)
    if (req.outputs[o].kind == 0 && B.nodes[out_node[o]].kind != 0)
      B.nodes[out_node[o]].last_use = std::max(B.nodes[out_node[o]].last_use, end_q + 1);

  // Linear-scan slot allocation; inputs own slots [0, n_dev).
  for (int c = 0; c < n_dev; c++) B.nodes[c].slot = c;
  std::vector<int> slot_free_at(kMaxSlots * 4, -1);  // slot -> index after which it is free
  int n_slots = n_dev;
  std::vector<std::vector<int>> dying(seq.size() + 2);
  for (int i = 0; i < (int)seq.size(); i++) {
    Node &n = B.nodes[seq[i]];
    // free operands whose last use is this op
    for (const Operand *o : {&n.a, &n.b}) {
      if (o->imm || o->node < 0) continue;
      Node &c = B.nodes[o->node];
      if (c.kind != 0 && c.last_use == i && c.slot >= 0) slot_free_at[c.slot] = -1;
    }
    int s = -1;
    for (int k = n_dev; k < (int)slot_free_at.size(); k++)
      if (slot_free_at[k] < 0) { s = k; break; }
    if (s < 0) fail("out of device registers");
    slot_free_at[s] = 1;
    n.slot = s;
    n_slots = std::max(n_slots, s + 1);
    if (n.last_use < 0) slot_free_at[s] = -1;  // dead value (should not happen)
    // a value with its last use beyond this op keeps the slot; values whose life
    // ends at a program boundary are released there
    if (i == end_p) {
      for (int j = 0; j <= end_p; j++) {
        Node &m = B.nodes[seq[j]];
        if (m.last_use <= end_p + 1 && m.slot >= 0) slot_free_at[m.slot] = -1;
      }
    }
  }
  if (n_slots > kMaxSlots) fail("plan needs more than 64 device registers");

  // Validity program: needed vnodes in creation order (topological), the
  // predicate's ops before the projection's (the projection may read words the
  // predicate computed, never the reverse). Registers are reused after a word's
  // last use (linear scan); input words are loaded before the first op and the
  // roots (stencil and output validity) live to the end.
  mark_vneeded(B, stencil_v);
  for (int v : out_v) mark_vneeded(B, v);
  std::vector<int> vseq;
  for (int prog = 0; prog < 2; prog++)
    for (size_t v = n_dev; v < B.vnodes.size(); v++)
      if (B.vnodes[v].needed && B.vnodes[v].program == prog) vseq.push_back((int)v);
  const int vend = (int)vseq.size();
  std::vector<int> vlast(B.vnodes.size(), -1);
  for (int i = 0; i < vend; i++) {
    const VNode &n = B.vnodes[vseq[i]];
    vlast[n.a] = std::max(vlast[n.a], i);
    vlast[n.b] = std::max(vlast[n.b], i);
  }
  if (stencil_v >= 0) vlast[stencil_v] = vend;
  for (int v : out_v)
    if (v >= 0) vlast[v] = vend;
  std::vector<bool> vbusy(kMaxVSlots, false);
  int n_vslots = 0;
  auto valloc = [&]() {
    for (int k = 0; k < kMaxVSlots; k++)
      if (!vbusy[k]) { vbusy[k] = true; n_vslots = std::max(n_vslots, k + 1); return k; }
    fail("plan needs more than 64 validity registers");
    return -1;
  };
  for (int c = 0; c < n_dev; c++)
    if (B.vnodes[c].needed) B.vnodes[c].vslot = valloc();
  std::vector<VOp> vops_p, vops_q;
  for (int i = 0; i < vend; i++) {
    VNode &n = B.vnodes[vseq[i]];
    for (int o : {n.a, n.b})
      if (vlast[o] == i) vbusy[B.vnodes[o].vslot] = false;
    n.vslot = valloc();
    VOp op{(uint8_t)n.code, (uint8_t)n.vslot, (uint8_t)B.vnodes[n.a].vslot, (uint8_t)B.vnodes[n.b].vslot};
    (n.program == 0 ? vops_p : vops_q).push_back(op);
  }
  auto vsrc = [&](int v) { return v == V_ALL ? -1 : v == V_NONE ? -2 : B.vnodes[v].vslot; };

  // Accumulator form: expression trees lowered to acc = acc OP operand code.
  std::vector<AccOp> acc_p, acc_q;
  int acc_temps = 0;
  bool acc_ok = n_vslots <= kVmVSlots && req.kernel != 2 && req.kernel != 3;
  if (acc_ok) {
    std::vector<int> uses(B.nodes.size(), 0);
    for (auto &n : B.nodes)
      if (n.kind == 1) {
        if (!n.a.imm && n.a.node >= 0) uses[n.a.node]++;
        if (!n.b.imm && n.b.node >= 0) uses[n.b.node]++;
      }
    std::vector<int> mat(B.nodes.size(), -1);  // materialised node -> temp
    std::vector<bool> tbusy(8, false);
    std::vector<AccOp> *code = nullptr;
    auto alloc_t = [&]() {
      for (int t = 0; t < 8; t++)
        if (!tbusy[t]) { tbusy[t] = true; acc_temps = std::max(acc_temps, t + 1); return t; }
      throw std::runtime_error("accumulator form needs more than 8 temporaries");
    };
    auto is_leaf = [&](int id) { return B.nodes[id].kind != 1 || mat[id] >= 0; };
    auto set_operand = [&](AccOp &a, const Operand &o) {
      if (o.imm) { a.okind = OK_IMM; a.imm = o.bits; return; }
      const Node &n = B.nodes[o.node];
      if (mat[o.node] >= 0) { a.okind = OK_TEMP; a.oidx = (uint8_t)mat[o.node]; }
      else if (n.kind == 0) { a.okind = OK_INPUT; a.oidx = (uint8_t)n.aux; }
      else { a.okind = OK_IMM; a.imm = 0; }  // constant zero
    };
    auto release = [&](const Operand &o) {  // a use of a materialised node is done
      if (o.imm || o.node < 0 || mat[o.node] < 0) return;
      if (--uses[o.node] == 0) tbusy[mat[o.node]] = false;
    };
    auto new_op = [](int code) { AccOp a{}; a.code = (uint8_t)code; a.vcol = 0xFF; return a; };
    auto is_cmp = [](int c) { return c >= OP_EQ_I && c <= OP_GE_F; };
    auto flip = [](int c) {  // imm OP x  ==  x FLIP(OP) imm
      switch (c) {
        case OP_LT_I: return (int)OP_GT_I; case OP_GT_I: return (int)OP_LT_I;
        case OP_LE_I: return (int)OP_GE_I; case OP_GE_I: return (int)OP_LE_I;
        case OP_LT_F: return (int)OP_GT_F; case OP_GT_F: return (int)OP_LT_F;
        case OP_LE_F: return (int)OP_GE_F; case OP_GE_F: return (int)OP_LE_F;
        default: return c;
      }
    };
    std::function<void(int)> gen = [&](int id) {
      Node &n = B.nodes[id];
      if (is_leaf(id)) {
        AccOp a = new_op(A_LOAD);
        Operand o; o.node = id;
        set_operand(a, o);
        code->push_back(a);
        return;
      }
      const int op = n.op.code;
      const int ca = n.op.conv & 3, cb = (n.op.conv >> 2) & 3, co = (n.op.conv >> 4) & 3;
      const bool unary = op >= OP_FLOOR;
      if (unary) {
        if (op != OP_IS_NULL && op != OP_IS_NOT_NULL) {
          if (n.a.imm) { AccOp l = new_op(A_LOAD); set_operand(l, n.a); code->push_back(l); }
          else gen(n.a.node);
        }
        release(n.a);
        AccOp a = new_op(A_UN);
        a.sub = (uint8_t)op;
        a.conv = (uint8_t)(ca | (co << 4));
        a.vcol = n.op.vcol;
        code->push_back(a);
      } else {
        // AND of 0/1 values with a comparison against an immediate: fused.
        if (op == OP_AND_B) {
          for (int side = 0; side < 2; side++) {
            const Operand &c = side ? n.a : n.b, &x = side ? n.b : n.a;
            if (c.imm || mat[c.node] >= 0 || uses[c.node] > 1) continue;
            const Node &cn = B.nodes[c.node];
            if (cn.kind != 1 || !is_cmp(cn.op.code) || ((cn.op.conv >> 4) & 3)) continue;
            const int cca = cn.op.conv & 3, ccb = (cn.op.conv >> 2) & 3;
            const Operand *leaf = nullptr, *imm = nullptr;
            int sub = cn.op.code, lconv = 0;
            if (cn.b.imm && !cn.a.imm && is_leaf(cn.a.node)) { leaf = &cn.a; imm = &cn.b; lconv = cca; }
            else if (cn.a.imm && !cn.b.imm && is_leaf(cn.b.node)) { leaf = &cn.b; imm = &cn.a; lconv = ccb; sub = flip(sub); }
            if (!leaf) continue;
            if (x.imm) { AccOp l = new_op(A_LOAD); set_operand(l, x); code->push_back(l); }
            else gen(x.node);
            release(x);
            AccOp a = new_op(A_CMPAND);
            a.sub = (uint8_t)sub;
            set_operand(a, *leaf);
            a.conv = (uint8_t)(lconv << 2);
            a.imm2 = imm->bits;
            code->push_back(a);
            release(*leaf);
            goto done;
          }
        }
        {
          const bool a_simple = n.a.imm || is_leaf(n.a.node);
          const bool b_simple = n.b.imm || is_leaf(n.b.node);
          AccOp a = new_op(A_BIN);
          a.sub = (uint8_t)op;
          a.vcol = n.op.vcol;
          if (b_simple) {
            if (n.a.imm) { AccOp l = new_op(A_LOAD); set_operand(l, n.a); code->push_back(l); }
            else gen(n.a.node);
            set_operand(a, n.b);
            a.conv = (uint8_t)(ca | (cb << 2) | (co << 4));
          } else if (a_simple) {
            gen(n.b.node);
            set_operand(a, n.a);
            a.flags = F_REV;
            a.conv = (uint8_t)(cb | (ca << 2) | (co << 4));
          } else {
            const bool a_first = B.nodes[n.a.node].need >= B.nodes[n.b.node].need;
            const int first = a_first ? n.a.node : n.b.node, second = a_first ? n.b.node : n.a.node;
            gen(first);
            const int t = alloc_t();
            AccOp st = new_op(A_STORE);
            st.oidx = (uint8_t)t;
            code->push_back(st);
            gen(second);
            a.okind = OK_TEMP;
            a.oidx = (uint8_t)t;
            if (a_first) { a.flags = F_REV; a.conv = (uint8_t)(cb | (ca << 2) | (co << 4)); }
            else a.conv = (uint8_t)(ca | (cb << 2) | (co << 4));
            tbusy[t] = false;
          }
          code->push_back(a);
          release(n.a);
          release(n.b);
        }
      done:;
      }
      if (uses[id] > 1) {  // shared value: keep it in a temporary for its other uses
        const int t = alloc_t();
        AccOp st = new_op(A_STORE);
        st.oidx = (uint8_t)t;
        code->push_back(st);
        mat[id] = t;
      }
    };
    try {
      for (auto &n : B.nodes) n.need = 0;
      for (size_t i = 0; i < B.nodes.size(); i++) compute_need(B, (int)i);
      code = &acc_p;
      if (stencil_node >= 0) gen(stencil_node);
      std::fill(tbusy.begin(), tbusy.end(), false);
      std::fill(mat.begin(), mat.end(), -1);
      code = &acc_q;
      for (size_t o = 0; o < req.outputs.size(); o++) {
        // bare-column outputs are copied by the kernels, raw
        if (req.outputs[o].kind == 1) continue;
        gen(out_node[o]);
        AccOp e = new_op(A_EMIT);
        e.oidx = (uint8_t)o;
        code->push_back(e);
      }
      if ((int)(acc_p.size() + acc_q.size()) > kMaxOps) throw std::runtime_error("too many ops");
    } catch (const std::runtime_error &) {
      acc_ok = false;
    }
  }

  if (req.kernel == 1 || req.kernel == 2) fail("kernel kind not available (1 and 2 are superseded by 4)");
  if (req.kernel < 0 || req.kernel > 5) fail("unknown kernel kind");
  // ---- Tile VM: the accumulator form, pre-decoded, over LDS-staged tiles ----
  if (req.kernel == 4 && !acc_ok) fail("plan does not fit the tile VM");
  if (acc_ok && (req.kernel == 0 || req.kernel == 4)) {
    std::string why;
    if (compile_vm(req, B, acc_p, acc_q, acc_temps, vops_p, vops_q, out_node, out_v, stencil_v, n_vslots,
                   out, why))
      return out;
    if (req.kernel == 4) fail(why);
  }

  // Register-file shape.
  const int need = std::max(n_slots, n_vslots);
  int S = need <= 8 ? 8 : need <= 16 ? 16 : need <= 32 ? 32 : 64;
  int R = S == 8 ? 2 : 1;
  if (req.force_s > 0) { S = req.force_s; R = req.force_r > 0 ? req.force_r : 1; }

  // Device ops.
  auto emit = [&](const std::vector<int> &ord, std::vector<DevOp> &dst) {
    for (int id : ord) {
      Node &n = B.nodes[id];
      DevOp op = n.op;
      if (n.kind == 2) { op = DevOp{}; op.code = OP_ZERO; op.flags = F_A_IMM | F_B_IMM; op.vcol = 0xFF; }
      op.dst = (uint8_t)n.slot;
      op.flags = 0;
      if (n.kind == 1) {
        if (n.a.imm) { op.flags |= F_A_IMM; op.imm_a = n.a.bits; } else op.a = (uint8_t)B.nodes[n.a.node].slot;
        if (n.b.imm) { op.flags |= F_B_IMM; op.imm_b = n.b.bits; } else op.b = (uint8_t)B.nodes[n.b.node].slot;
      } else {
        op.flags = F_A_IMM | F_B_IMM;
      }
      dst.push_back(op);
    }
  };
  std::vector<DevOp> ops_p, ops_q;
  emit(order_p, ops_p);
  emit(order_q, ops_q);

  // The direct kernel reads inputs straight from HBM: no staging layout.
  const int n_out = (int)req.outputs.size();
  int n_valid_out = 0;
  for (auto &o : req.outputs) n_valid_out += o.valid != nullptr;
  if (S < need) fail("register file too small for the plan");

  // Assemble the blob.
  size_t off_in = (sizeof(ProgHeader) + 15) & ~size_t(15);
  size_t off_out = off_in + sizeof(InCol) * n_dev;
  off_out = (off_out + 15) & ~size_t(15);
  size_t off_ops = off_out + sizeof(OutCol) * n_out;
  off_ops = (off_ops + 15) & ~size_t(15);
  const size_t n_ops_total = ops_p.size() + ops_q.size();
  size_t off_vops = off_ops + sizeof(DevOp) * n_ops_total;
  off_vops = (off_vops + 15) & ~size_t(15);
  size_t total = off_vops + sizeof(VOp) * (vops_p.size() + vops_q.size());
  total = (total + 15) & ~size_t(15);
  if ((int)total > blob_capacity()) fail("plan too large for the on-chip program buffer");
  out.blob.assign(total, 0);
  ProgHeader h{};
  h.n_rows = req.n_rows;
  h.n_in = n_dev;
  h.n_out = n_out;
  h.n_ops_p = (int)ops_p.size();
  h.n_ops_q = (int)ops_q.size();
  h.acc_mode = 0;
  h.n_temps = 0;
  h.n_vops_p = (int)vops_p.size();
  h.n_vops_q = (int)vops_q.size();
  h.n_vin = n_dev;
  h.n_slots = n_slots;
  h.n_vslots = n_vslots;
  h.stencil_slot = stencil_node >= 0 ? B.nodes[stencil_node].slot : -1;
  h.stencil_width = (req.compact && !req.pred) ? 8 : 1;
  h.stencil_vsrc = vsrc(stencil_v);
  int any_valid = 0;
  for (auto &o : req.outputs) any_valid |= o.valid != nullptr;
  h.any_valid_out = any_valid;
  h.tile_rows = 0;
  h.n_chunks = 0;
  h.lds_vbits_off = 0;
  h.n_valid_out = n_valid_out;
  h.lds_bytes = 0;
  h.off_chunks = (int)off_vops;
  h.off_in = (int)off_in;
  h.off_out = (int)off_out;
  h.off_ops = (int)off_ops;
  h.off_vops = (int)off_vops;
  h.blob_bytes = (int)total;
  put(out.blob, 0, h);
  for (int c = 0; c < n_dev; c++) {
    InCol ic{};
    ic.data = req.inputs[c].data;
    ic.valid = req.inputs[c].valid;
    ic.load = load_kind(req.inputs[c].dtype);
    ic.vslot = B.vnodes[c].needed ? B.vnodes[c].vslot : -1;
    ic.lds_off = 0;
    ic.lds_voff = -1;
    put(out.blob, off_in + sizeof(InCol) * c, ic);
  }
  int vbuf_next = 0;
  for (int o = 0; o < n_out; o++) {
    const OutSpec &s = req.outputs[o];
    OutCol oc{};
    oc.data = s.data;
    oc.valid = s.valid;
    oc.store = s.kind == 0 ? store_kind_program(s.dtype) : store_kind_alias(s.dtype);
    oc.slot = (int16_t)B.nodes[out_node[o]].slot;
    oc.src_kind = (int16_t)s.kind;
    oc.vsrc = vsrc(out_v[o]);
    oc.vbuf = s.valid ? vbuf_next++ : -1;
    put(out.blob, off_out + sizeof(OutCol) * o, oc);
  }
  size_t p = off_ops;
  for (auto &op : ops_p) { put(out.blob, p, op); p += sizeof(DevOp); }
  for (auto &op : ops_q) { put(out.blob, p, op); p += sizeof(DevOp); }
  p = off_vops;
  for (auto &op : vops_p) { put(out.blob, p, op); p += sizeof(VOp); }
  for (auto &op : vops_q) { put(out.blob, p, op); p += sizeof(VOp); }
  out.S = S;
  out.R = R;
  out.n_slots = n_slots;
  out.n_vslots = n_vslots;
  out.n_ops = (int)(ops_p.size() + ops_q.size());
  out.n_vops = (int)(vops_p.size() + vops_q.size());
  out.T = 0;
  out.lds_bytes = 0;
  out.acc = false;
  if (req.kernel == 5) {
    // ---- compiled tier: the SSA graph written out as a kernel ----
    CodegenInput ci;
    ci.B = &B;
    ci.order_p = &order_p;
    ci.order_q = &order_q;
    ci.compact = req.compact;
    ci.stencil_node = stencil_node;
    ci.stencil_v = stencil_v;
    ci.stencil_width = h.stencil_width;
    for (int c = 0; c < n_dev; c++) {
      ci.in_load.push_back(load_kind(req.inputs[c].dtype));
      ci.in_width.push_back(wid(req.inputs[c].dtype));
      ci.in_valid.push_back(req.inputs[c].valid != nullptr);
    }
    for (int o = 0; o < n_out; o++) {
      const OutSpec &s = req.outputs[o];
      ci.out_node.push_back(out_node[o]);
      ci.out_v.push_back(out_v[o]);
      ci.out_kind.push_back(s.kind);
      ci.out_store.push_back(s.kind == 0 ? store_kind_program(s.dtype) : store_kind_alias(s.dtype));
      ci.out_input.push_back(s.kind == 1 ? s.index : -1);
      ci.out_valid.push_back(s.valid != nullptr);
    }
    // Rows per lane: about 64 bytes per lane of the columns a step holds (a filter's
    // predicate columns, a projection's columns), so every load of a step is in
    // flight without spilling (one 4-byte column: 16 rows, 16 bytes per row: 8,
    // 48: 2), unless forced (r).
    int bpr = 0;
    {
      std::vector<bool> used(n_dev, false);
      const std::vector<int> &ord = req.compact ? order_p : order_q;
      for (int id : ord) {
        const Node &nd = B.nodes[id];
        if (nd.kind != 1) continue;
        if (!nd.a.imm && nd.a.node >= 0 && nd.a.node < n_dev) used[nd.a.node] = true;
        if (!nd.b.imm && nd.b.node >= 0 && nd.b.node < n_dev) used[nd.b.node] = true;
      }
      if (req.compact && stencil_node >= 0 && stencil_node < n_dev) used[stencil_node] = true;
      if (!req.compact)
        for (int o = 0; o < n_out; o++) {
          if (req.outputs[o].kind == 1) used[req.outputs[o].index] = true;
          else if (out_node[o] < n_dev) used[out_node[o]] = true;
        }
      for (int c = 0; c < n_dev; c++)
        if (used[c]) bpr += wid(req.inputs[c].dtype);
    }
    int rs = bpr <= 4 ? 16 : bpr <= 16 ? 8 : bpr <= 32 ? 4 : 2;
    if (req.force_r > 0) {
      if (req.force_r != 1 && req.force_r != 2 && req.force_r != 4 && req.force_r != 8 && req.force_r != 16)
        fail("compiled kernel: rows per lane must be 1, 2, 4, 8 or 16");
      rs = req.force_r;
    }
    ci.rs = rs;
    const CodegenOutput g = generate_kernel_source(ci);
    out.jit = true;
    out.jit_src = g.src;
    out.jit_name = g.name;
    out.jit_side_bytes = g.side_bytes;
    out.jit_null_outputs = g.null_outputs;
    out.R = g.rs;
    out.T = req.compact ? g.tile_rows : 0;
    if (req.compact) {  // the slot scan reads the slot shape from the header
      h.tile_rows = out.T;
      put(out.blob, 0, h);
    }
  }
  return out;
}

}  // namespace blz

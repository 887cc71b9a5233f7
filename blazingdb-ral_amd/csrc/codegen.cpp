/*
 * codegen.cpp — writes one plan's SSA graph out as a HIP kernel (see codegen.h).
 *
 * Semantics follow the interpreter kernels op for op: an input column is widened
 * into a 64-bit slot as interpreter_ops.cuh:386-452 widens it (and as the tile VM's
 * operand classes do), each op node is op_t<code>(conv(a), conv(b), raw validity
 * bit) with the result conversion after it (the DevOp of program.cpp's build_plan),
 * outputs are narrowed as interpreter_ops.cuh:317-381 narrows them, bare-column
 * outputs are copied by width (materialize_column), and validity is the AND / OR
 * program over 64-row words (process_valids, interpreter_ops.cuh:507-571). The
 * filter kernel keeps a row when its predicate value is non-zero in its low
 * stencil_width bytes and the predicate's validity bit is set (the stencil of
 * CalciteInterpreter.cpp:509-573), and writes kept rows in row order.
 */
#include "codegen.h"

#include <cstdio>
#include <set>
#include <sstream>
#include <stdexcept>

#include "program.h"

namespace blz {
namespace {

const char *kPrelude = R"(#include "device_common.h"
using namespace blz;
template <bool B> struct CgTag { static constexpr bool value = B; };
// Set bits of m below this lane.
DEV uint32_t cg_rank(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
)";

// DevOpCode names (program.h order), so the generated source reads as the plan.
const char *const kOpNames[] = {
    "OP_ADD_I", "OP_ADD_F", "OP_SUB_I", "OP_SUB_F", "OP_MUL_I", "OP_MUL_F", "OP_DIV_I", "OP_DIV_F",
    "OP_MOD_I", "OP_POW_I", "OP_POW_F", "OP_EQ_I", "OP_EQ_F", "OP_NE_I", "OP_NE_F", "OP_LT_I",
    "OP_LT_F", "OP_GT_I", "OP_GT_F", "OP_LE_I", "OP_LE_F", "OP_GE_I", "OP_GE_F", "OP_COALESCE",
    "OP_AND_B", "OP_FLOOR", "OP_CEIL", "OP_SIN", "OP_COS", "OP_ASIN", "OP_ACOS", "OP_TAN",
    "OP_COTAN", "OP_ATAN", "OP_ABS", "OP_LN", "OP_LOG10", "OP_NOT_I", "OP_NOT_F", "OP_YEAR32",
    "OP_MONTH32", "OP_DAY32", "OP_YEAR64", "OP_MONTH64", "OP_DAY64", "OP_HOUR64", "OP_MINUTE64",
    "OP_SECOND64", "OP_ZERO", "OP_IS_NULL", "OP_IS_NOT_NULL", "OP_IDENT"};
static_assert(sizeof(kOpNames) / sizeof(kOpNames[0]) == OP_COUNT_, "one name per DevOpCode");
const char *op_name(int code) {
  if (code < 0 || code >= OP_COUNT_) throw std::runtime_error("compiled tier: op code " + std::to_string(code));
  return kOpNames[code];
}

const char *uint_of_width(int w) {
  switch (w) {
    case 1: return "uint8_t";
    case 2: return "uint16_t";
    case 4: return "uint32_t";
    case 8: return "uint64_t";
    default: throw std::runtime_error("compiled tier: column width " + std::to_string(w));
  }
}

// Widening of raw column bits `r` into a 64-bit slot.
std::string widen(int load, const std::string &r) {
  switch (load) {
    case LD_I8: return "(uint64_t)(int64_t)(int8_t)" + r;
    case LD_U8: return "(uint64_t)(uint8_t)" + r;
    case LD_I16: return "(uint64_t)(int64_t)(int16_t)" + r;
    case LD_I32: return "(uint64_t)(int64_t)(int32_t)" + r;
    case LD_F32: return "d2b((double)__builtin_bit_cast(float, (uint32_t)" + r + "))";
    case LD_I64: case LD_F64: return "(uint64_t)" + r;
    default: return "0ull";
  }
}

std::string conv_expr(int cv, const std::string &x) {
  if (cv == CV_I2F) return "d2b((double)(int64_t)(" + x + "))";
  if (cv == CV_F2I) return "(uint64_t)f2i(b2d(" + x + "))";
  return x;
}

std::string hex64(uint64_t v) {
  char b[32];
  std::snprintf(b, sizeof b, "0x%llxull", (unsigned long long)v);
  return b;
}

// Narrowing store of slot value v (StoreKind).
const char *store_type(int st) {
  switch (st) {
    case ST_I8: return "int8_t";
    case ST_I16: return "int16_t";
    case ST_I32: return "int32_t";
    case ST_I64: case ST_F64: return "uint64_t";
    case ST_F32: return "float";
    default: return nullptr;
  }
}
std::string store_value(int st, const std::string &v) {
  switch (st) {
    case ST_I8: return "(int8_t)" + v;
    case ST_I16: return "(int16_t)" + v;
    case ST_I32: return "(int32_t)" + v;
    case ST_F32: return "(float)b2d(" + v + ")";
    default: return v;
  }
}

struct Gen {
  const CodegenInput &in;
  const Builder &B;
  int n_dev;
  std::ostringstream o;
  explicit Gen(const CodegenInput &i) : in(i), B(*i.B), n_dev((int)i.in_load.size()) {}

  bool per_lane_bits = false;  // COALESCE / IS_NULL read vb<c>[u] (gather) instead of a word
  bool out_base = false;       // stores through ob<k> (the unit's base) instead of o<k>
  std::string val(int node) const {
    if (node < n_dev) return "x" + std::to_string(node);
    return "v" + std::to_string(node);
  }
  std::string operand(const Operand &a) const { return a.imm ? hex64(a.bits) : val(a.node); }

  // Inputs read by the value nodes of `order` (and by the extra roots).
  void inputs_of(const std::vector<int> &order, std::set<int> &cols, std::set<int> &vcols) const {
    for (int id : order) {
      const Node &n = B.nodes[id];
      if (n.kind != 1) continue;
      if (!n.a.imm && n.a.node >= 0 && n.a.node < n_dev) cols.insert(n.a.node);
      if (!n.b.imm && n.b.node >= 0 && n.b.node < n_dev) cols.insert(n.b.node);
      if (n.op.vcol != 0xFF) vcols.insert(n.op.vcol);
    }
  }
  // Validity nodes under `root` (creation order is topological).
  void vneeded(int root, std::set<int> &vs) const {
    if (root < 0 || vs.count(root)) return;
    vs.insert(root);
    const VNode &v = B.vnodes[root];
    if (v.kind == 1) { vneeded(v.a, vs); vneeded(v.b, vs); }
  }
  // Validity value as an expression: lane u (< RS) holds the word of group u.
  std::string vword(int v) const {
    if (v == V_ALL) return "~0ull";
    if (v == V_NONE) return "0ull";
    const VNode &n = B.vnodes[v];
    if (n.kind == 0) return "mv" + std::to_string(n.aux);
    return "y" + std::to_string(v);
  }
  // Validity value as this lane's bit (gather: per-lane bits of the source rows).
  std::string vbit(int v) const {
    if (v == V_ALL) return "true";
    if (v == V_NONE) return "false";
    const VNode &n = B.vnodes[v];
    if (n.kind == 0) return "vb" + std::to_string(n.aux) + "[u]";
    return "z" + std::to_string(v);
  }
  void emit_vnode_bits(const std::set<int> &vs, const std::string &ind) {
    for (int v : vs) {
      const VNode &n = B.vnodes[v];
      if (n.kind != 1) continue;
      o << ind << "const bool z" << v << " = " << vbit(n.a) << (n.code == VOP_OR ? " | " : " & ") << vbit(n.b)
        << ";\n";
    }
  }
  // y<v> = ... for the op vnodes of `vs`, in creation order (word of the lane's group).
  void emit_vnodes(const std::set<int> &vs, const std::string &ind) {
    for (int v : vs) {
      const VNode &n = B.vnodes[v];
      if (n.kind != 1) continue;
      o << ind << "const uint64_t y" << v << " = " << vword(n.a) << (n.code == VOP_OR ? " | " : " & ")
        << vword(n.b) << ";\n";
    }
  }
  // x<c> = widened value of input c in row u (raw bits in <pre>r<c>[u]).
  void emit_inputs(const std::set<int> &cols, const std::string &ind, const std::string &pre) {
    for (int c : cols)
      o << ind << "const uint64_t x" << c << " = "
        << widen(in.in_load[c], pre + "r" + std::to_string(c) + "[u]") << ";\n";
  }
  void emit_nodes(const std::vector<int> &order, const std::string &ind) {
    for (int id : order) {
      const Node &n = B.nodes[id];
      if (n.kind == 2) {
        o << ind << "const uint64_t v" << id << " = 0ull;\n";
        continue;
      }
      if (n.kind != 1) continue;
      const int ca = n.op.conv & 3, cb = (n.op.conv >> 2) & 3, co = (n.op.conv >> 4) & 3;
      std::string vb = "true";
      if (n.op.vcol != 0xFF && in.in_valid[n.op.vcol])
        vb = per_lane_bits ? "vb" + std::to_string(n.op.vcol) + "[u]"
                           : "((readlane64(mv" + std::to_string(n.op.vcol) + ", u) >> lane) & 1) != 0";
      const std::string e = std::string("op_t<") + op_name(n.op.code) + ">(" + conv_expr(ca, operand(n.a)) +
                            ", " + conv_expr(cb, operand(n.b)) + ", " + vb + ")";
      o << ind << "const uint64_t v" << id << " = " << conv_expr(co, e) << ";\n";
    }
  }
  // Loads of rows (g0 + u) * 64 + lane of `cols` into <pre>r<c>[u]; clamp: rows
  // past the last are read as the last row (in bounds; never kept or stored).
  void emit_load_body(const std::set<int> &cols, const std::string &ind, const std::string &pre, bool clamp) {
    o << ind << "#pragma unroll\n" << ind << "for (int u = 0; u < RS; u++) {\n";
    o << ind << "  const int64_t r = (g0 + u) * 64 + lane;\n";
    o << ind << "  const int64_t rc = " << (clamp ? "r < n ? r : n - 1" : "r") << ";\n";
    for (int c : cols)
      o << ind << "  " << pre << "r" << c << "[u] = __builtin_nontemporal_load(c" << c << " + rc);\n";
    o << ind << "}\n";
  }
  // Rows (g0 + u) * 64 + lane of `cols` into r<c>[u], from a uniform base plus the
  // lane offset and a constant per group; in a step reaching past the last row
  // (CLAMP) the lanes past it load the last row (in range: a filter's last tile can
  // hold slots that start past the last row) and then hold 0, so that no load sits
  // behind a branch of its own and waits alone.
  void emit_loads(const std::set<int> &cols, const std::string &ind) {
    for (int c : cols) o << ind << uint_of_width(in.in_width[c]) << " r" << c << "[RS];\n";
    if (cols.empty()) return;
    for (int c : cols)
      o << ind << "const GLOBAL " << uint_of_width(in.in_width[c]) << " *b" << c << " = c" << c << " + g0 * 64;\n";
    o << ind << "#pragma unroll\n" << ind << "for (int u = 0; u < RS; u++) {\n";
    o << ind << "  const uint32_t lo = (uint32_t)(u * 64 + lane);\n";
    o << ind << "  const bool in_range = !CLAMP || (g0 + u) * 64 + lane < n;\n";
    for (int c : cols)
      o << ind << "  r" << c << "[u] = __builtin_nontemporal_load(in_range ? b" << c << " + lo : c" << c
        << " + (n - 1));\n";
    o << ind << "}\n";
    o << ind << "#pragma unroll\n" << ind << "for (int u = 0; u < RS; u++) {\n";
    o << ind << "  const bool in_range = !CLAMP || (g0 + u) * 64 + lane < n;\n";
    for (int c : cols) o << ind << "  if (!in_range) r" << c << "[u] = 0;\n";
    o << ind << "}\n";
  }
  // Bitmap words of the groups g0 .. g0 + RS - 1: lane u < RS holds group u's word
  // (mv<c>; 0 in the other lanes and past the last row).
  void emit_bitmap_words(const std::set<int> &bm, const std::string &ind, const std::string &clamp) {
    for (int c : bm) {
      o << ind << "uint64_t mv" << c << " = 0;\n";
      // whole steps: every lane loads (lanes past RS reload word g0, then hold 0), so
      // the load is not behind a branch of its own
      o << ind << "if (" << clamp << ") { if (lane < RS) mv" << c << " = load_vword(m" << c << ", g0 + lane, n); }\n"
        << ind << "else { const uint64_t t = ((const GLOBAL uint64_t *)m" << c << ")[g0 + (lane < RS ? lane : 0)];"
        << " mv" << c << " = lane < RS ? t : 0ull; }\n";
    }
  }
  // Stores of the outputs at row index `pos` (row u); a bare-column output copies
  // the input's raw bits (<pre>r<c>[u] when the input is in `pre_cols`).
  // Output data is written once and not read again by the kernel: non-temporal
  // stores (the 500M-row gather at s = 0.9: 10.1 -> 9.0-9.8 ms).
  void emit_stores(const std::string &pos, const std::string &ind, const std::string &pre,
                   const std::set<int> *pre_cols = nullptr) {
    for (size_t k = 0; k < in.out_node.size(); k++) {
      const std::string dst = std::string(out_base ? "ob" : "o") + std::to_string(k) + " + " + pos;
      if (in.out_kind[k] == 1) {
        const int c = in.out_input[k];
        const std::string p = pre_cols && pre_cols->count(c) ? pre : "";
        o << ind << "__builtin_nontemporal_store(" << p << "r" << c << "[u], " << dst << ");\n";
      } else {
        if (!store_type(in.out_store[k])) continue;
        o << ind << "__builtin_nontemporal_store(" << store_value(in.out_store[k], val(in.out_node[k])) << ", "
          << dst << ");\n";
      }
    }
  }
};

}  // namespace

const char *codegen_prelude() { return kPrelude; }

CodegenOutput generate_kernel_source(const CodegenInput &in) {
  Gen g(in);
  const Builder &B = *in.B;
  const int n_dev = g.n_dev;
  const int n_out = (int)in.out_node.size();
  auto &o = g.o;
  CodegenOutput res;
  res.rs = in.rs;

  // Columns and bitmaps each phase reads.
  std::set<int> cols_p, cols_q, vcols;
  if (in.compact) {
    g.inputs_of(*in.order_p, cols_p, vcols);
    if (in.stencil_node >= 0 && in.stencil_node < n_dev) cols_p.insert(in.stencil_node);
  }
  std::set<int> q_reads;  // inputs the projection's values read
  g.inputs_of(*in.order_q, q_reads, vcols);
  for (int k = 0; k < n_out; k++)
    if (in.out_kind[k] == 0 && in.out_node[k] >= 0 && in.out_node[k] < n_dev) q_reads.insert(in.out_node[k]);
  cols_q = q_reads;
  for (int k = 0; k < n_out; k++)
    if (in.out_kind[k] == 1) cols_q.insert(in.out_input[k]);
  for (int c : cols_p) cols_q.erase(c);  // read once, in the predicate phase
  std::set<int> vs_p, vs_q;  // validity nodes: the predicate's, the outputs'
  if (in.compact) g.vneeded(in.stencil_v, vs_p);
  for (int k = 0; k < n_out; k++)
    if (in.out_valid[k]) g.vneeded(in.out_v[k], vs_q);
  std::set<int> bitmaps;  // raw bitmap words loaded per group
  for (int v : vs_p) if (B.vnodes[v].kind == 0) bitmaps.insert(B.vnodes[v].aux);
  for (int v : vs_q) if (B.vnodes[v].kind == 0) bitmaps.insert(B.vnodes[v].aux);
  for (int c : vcols) if (in.in_valid[c]) bitmaps.insert(c);
  std::set<int> all_cols = cols_p;
  all_cols.insert(cols_q.begin(), cols_q.end());
  bool any_valid = false;
  for (int k = 0; k < n_out; k++) any_valid |= (bool)in.out_valid[k];

  o << kPrelude;
  o << "#define RS " << in.rs << "\n";
  const std::string kname = in.compact ? "blz_cg_pred" : "blz_cg_project";
  o << "extern \"C\" __global__ void __launch_bounds__(256) " << kname
    << "(const uint8_t *__restrict__ blob_arg, Workspace ws) {\n";
  o << "  const CONSTANT uint8_t *blob = (const CONSTANT uint8_t *)blob_arg;\n"
       "  const CONSTANT ProgHeader &h = *(const CONSTANT ProgHeader *)blob;\n"
       "  const CONSTANT InCol *in = (const CONSTANT InCol *)(blob + h.off_in);\n"
       "  const CONSTANT OutCol *out = (const CONSTANT OutCol *)(blob + h.off_out);\n"
       "  (void)in; (void)out;\n"
       "  const int64_t n = h.n_rows;\n"
       "  const int lane = threadIdx.x & 63;\n";
  for (int c : all_cols)
    o << "  const GLOBAL " << uint_of_width(in.in_width[c]) << " *c" << c << " = (const GLOBAL "
      << uint_of_width(in.in_width[c]) << " *)in[" << c << "].data;\n";
  for (int c : bitmaps) o << "  const uint8_t *m" << c << " = (const uint8_t *)in[" << c << "].valid;\n";
  if (!in.compact)
    for (int k = 0; k < n_out; k++) {
      const char *t = in.out_kind[k] == 1 ? uint_of_width(in.in_width[in.out_input[k]]) : store_type(in.out_store[k]);
      if (t) o << "  GLOBAL " << t << " *o" << k << " = (GLOBAL " << t << " *)out[" << k << "].data;\n";
      if (in.out_valid[k]) o << "  uint64_t *q" << k << " = (uint64_t *)out[" << k << "].valid;\n";
      if (in.out_valid[k]) o << "  unsigned long long nn" << k << " = 0;\n";
    }

  if (!in.compact) {
    // ---------------- projection ----------------
    res.tile_rows = 64 * in.rs;
    o << "  const int64_t wave = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * 256 + threadIdx.x) >> 6));\n"
         "  const int64_t n_waves = ((int64_t)gridDim.x * 256) >> 6;\n"
         "  const int64_t n_groups = (n + 63) >> 6;\n"
         "  const int64_t n_units = (n_groups + RS - 1) / RS;\n"
         "  const int64_t n_blocks = (n_units + 3) / 4;\n";
    o << "  auto unit = [&](auto tag, int64_t g0) {\n"
         "    constexpr bool CLAMP = decltype(tag)::value;\n";
    g.emit_loads(cols_q, "    ");
    g.emit_bitmap_words(bitmaps, "    ", "CLAMP");
    for (int k = 0; k < n_out; k++)  // this unit's outputs: uniform base, lane offset, constant per group
      if (in.out_kind[k] == 1 || store_type(in.out_store[k])) o << "    const auto ob" << k << " = o" << k << " + g0 * 64;\n";
    o << "    #pragma unroll\n    for (int u = 0; u < RS; u++) {\n"
         "      const int64_t r = (g0 + u) * 64 + lane;\n";
    g.emit_inputs(cols_q, "      ", "");
    g.emit_nodes(*in.order_q, "      ");
    o << "      if (!CLAMP || r < n) {\n";
    g.out_base = true;
    g.emit_stores("(uint32_t)(u * 64 + lane)", "        ", "");
    g.out_base = false;
    o << "      }\n    }\n";
    if (any_valid) {
      // lane u < RS: the validity words of group u, stored whole; nulls counted
      g.emit_vnodes(vs_q, "    ");
      o << "    if (lane < RS && (g0 + lane) * 64 < n) {\n"
           "      const uint64_t live = low_mask(n - (g0 + lane) * 64);\n";
      for (int k = 0; k < n_out; k++) {
        if (!in.out_valid[k]) continue;
        o << "      { const uint64_t wd = " << g.vword(in.out_v[k]) << " & live;  // bits past the last row: 0\n"
          << "        store_vword(q" << k << ", g0 + lane, n, wd);\n"
          << "        nn" << k << " += __builtin_popcountll(~wd & live); }\n";
      }
      o << "    }\n";
    }
    o << "  };\n";
    o << "  for (int64_t blk = wave; blk < n_blocks; blk += n_waves)\n"
         "    for (int64_t un = blk * 4; un < blk * 4 + 4 && un < n_units; un++) {\n"
         "      if ((un + 1) * RS * 64 <= n) unit(CgTag<false>{}, un * RS);\n"
         "      else unit(CgTag<true>{}, un * RS);\n"
         "    }\n";
  } else {
    // ---------------- filter (+ projection) ----------------
    // Two kernels around the slot scan (vm_kernel.hip), with the tile VM's keep
    // words, slot counts and kept-row indices:
    //  blz_cg_pred    the predicate over every row from registers: one keep word per
    //                 64 rows, the kept count of each slot (one wave step of 64 RS
    //                 rows), and the kept rows' values of the predicate columns the
    //                 projection reads, compacted per slot (side staging);
    //  blz_cg_gather  output-driven: chunks of 64 GR outputs per wave, source rows
    //                 from the scan's indices; staged values read densely at their
    //                 slot's offset (the scan leaves each slot's exclusive prefix in
    //                 its count), other columns at the source rows; outputs and
    //                 validity words stored whole, in row order.
    res.tile_rows = 256 * in.rs;
    std::vector<int> side;  // predicate columns the projection (or a bare output) reads
    for (int c : cols_p) {
      bool need = q_reads.count(c) > 0;
      for (int k = 0; k < n_out; k++) need |= in.out_kind[k] == 1 && in.out_input[k] == c;
      if (need) side.push_back(c);
    }
    std::set<int> bm_p, bm_q;  // bitmaps: the predicate's words, the projection's bits
    for (int v : vs_p) if (B.vnodes[v].kind == 0) bm_p.insert(B.vnodes[v].aux);
    {
      std::set<int> vc_p, vc_q, dummy;
      g.inputs_of(*in.order_p, dummy, vc_p);
      g.inputs_of(*in.order_q, dummy, vc_q);
      for (int c : vc_p) if (in.in_valid[c]) bm_p.insert(c);
      for (int c : vc_q) if (in.in_valid[c]) bm_q.insert(c);
    }
    for (int v : vs_q) if (B.vnodes[v].kind == 0) bm_q.insert(B.vnodes[v].aux);
    const uint64_t st_mask = in.stencil_width >= 8 ? ~0ull : ((1ull << (8 * in.stencil_width)) - 1);
    auto side_ptr = [&](int c) {  // staged values of column c: ws.side + (bytes of earlier ones) * capacity
      int off = 0;
      for (int d : side) {
        if (d == c) break;
        off += in.in_width[d];
      }
      o << "  GLOBAL " << uint_of_width(in.in_width[c]) << " *s" << c << " = (GLOBAL "
        << uint_of_width(in.in_width[c]) << " *)(ws.side + (size_t)" << off << " * cap);\n";
    };
    for (int c : side) res.side_bytes += in.in_width[c];

    // ---- blz_cg_pred ----
    o << "  const int64_t T = 256 * RS;\n"
         "  const int64_t n_slots = ((n + T - 1) / T) * 4;\n"
         "  const int64_t cap = n_slots * 64 * RS;\n"
         "  (void)cap;\n";
    for (int c : side) side_ptr(c);
    o << "  const int64_t wave = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * 256 + threadIdx.x) >> 6));\n"
         "  const int64_t n_waves = ((int64_t)gridDim.x * 256) >> 6;\n"
         "  const int64_t n_blocks = (n_slots + 3) / 4;\n";
    o << "  auto step = [&](auto tag, int64_t sl) {\n"
         "    constexpr bool CLAMP = decltype(tag)::value;\n"
         "    const int64_t g0 = sl * RS;\n";
    g.emit_loads(cols_p, "    ");
    g.emit_bitmap_words(bm_p, "    ", "CLAMP");
    g.emit_vnodes(vs_p, "    ");
    o << "    const uint64_t svw = " << g.vword(in.stencil_v) << ";\n"
         "    (void)svw;\n";
    o << "    uint64_t myword = 0;\n    uint32_t cnt = 0;\n"
         "    #pragma unroll\n    for (int u = 0; u < RS; u++) {\n"
         "      const int64_t r = (g0 + u) * 64 + lane;\n";
    g.emit_inputs(cols_p, "      ", "");
    g.emit_nodes(*in.order_p, "      ");
    o << "      const uint64_t sv = " << (in.stencil_v == V_ALL ? std::string("~0ull") : in.stencil_v == V_NONE ? std::string("0ull") : std::string("readlane64(svw, u)")) << ";\n";
    o << "      const bool keep = (!CLAMP || r < n) && ((" << g.val(in.stencil_node) << " & " << hex64(st_mask)
      << ") != 0) && ((sv >> lane) & 1);\n";
    o << "      const uint64_t km = __builtin_amdgcn_ballot_w64(keep);\n"
         "      if (lane == u) myword = km;\n";
    if (!side.empty()) {
      o << "      if (keep) {\n"
           "        const int64_t p = sl * 64 * RS + cnt + cg_rank(km);\n";
      for (int c : side) o << "        s" << c << "[p] = r" << c << "[u];\n";
      o << "      }\n";
    }
    o << "      cnt += (uint32_t)__builtin_popcountll(km);\n    }\n"
         "    if (lane < RS) ws.keep[g0 + lane] = myword;\n"
         "    if (lane == 0) ws.wave_count[sl] = cnt;\n"
         "  };\n";
    o << "  for (int64_t blk = wave; blk < n_blocks; blk += n_waves)\n"
         "    for (int64_t sl = blk * 4; sl < blk * 4 + 4 && sl < n_slots; sl++) {\n"
         "      if ((sl + 1) * RS * 64 <= n) step(CgTag<false>{}, sl);\n"
         "      else step(CgTag<true>{}, sl);\n"
         "    }\n"
         "}\n\n";

    // ---- blz_cg_gather ----
    o << "#define GR 4\n";
    o << "extern \"C\" __global__ void __launch_bounds__(256) blz_cg_gather(const uint8_t *__restrict__ blob_arg, "
         "Workspace ws) {\n";
    o << "  const CONSTANT uint8_t *blob = (const CONSTANT uint8_t *)blob_arg;\n"
         "  const CONSTANT ProgHeader &h = *(const CONSTANT ProgHeader *)blob;\n"
         "  const CONSTANT InCol *in = (const CONSTANT InCol *)(blob + h.off_in);\n"
         "  const CONSTANT OutCol *out = (const CONSTANT OutCol *)(blob + h.off_out);\n"
         "  (void)in; (void)out;\n"
         "  const int64_t n = h.n_rows;\n"
         "  const int lane = threadIdx.x & 63;\n"
         "  const int64_t T = 256 * RS;\n"
         "  const int64_t n_slots = ((n + T - 1) / T) * 4;\n"
         "  const int64_t cap = n_slots * 64 * RS;\n"
         "  (void)cap;\n";
    std::set<int> own;  // columns read at the source rows
    for (int c : q_reads) if (!cols_p.count(c)) own.insert(c);
    for (int k = 0; k < n_out; k++)
      if (in.out_kind[k] == 1 && !cols_p.count(in.out_input[k])) own.insert(in.out_input[k]);
    for (int c : own)
      o << "  const GLOBAL " << uint_of_width(in.in_width[c]) << " *c" << c << " = (const GLOBAL "
        << uint_of_width(in.in_width[c]) << " *)in[" << c << "].data;\n";
    for (int c : side) side_ptr(c);
    for (int c : bm_q) o << "  const uint8_t *m" << c << " = (const uint8_t *)in[" << c << "].valid;\n";
    for (int k = 0; k < n_out; k++) {
      const char *t = in.out_kind[k] == 1 ? uint_of_width(in.in_width[in.out_input[k]]) : store_type(in.out_store[k]);
      if (t) o << "  GLOBAL " << t << " *o" << k << " = (GLOBAL " << t << " *)out[" << k << "].data;\n";
      if (in.out_valid[k]) o << "  uint64_t *q" << k << " = (uint64_t *)out[" << k << "].valid;\n";
      if (in.out_valid[k]) o << "  unsigned long long nn" << k << " = 0;\n";
    }
    o << "  const int64_t n_sel = (int64_t)*(volatile unsigned long long *)ws.n_sel;\n"
         "  const int64_t wave = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * 256 + threadIdx.x) >> 6));\n"
         "  const int64_t n_waves = ((int64_t)gridDim.x * 256) >> 6;\n"
         "  const int64_t n_chunks = (n_sel + 64 * GR - 1) / (64 * GR);\n"
         "  for (int64_t ch = wave; ch < n_chunks; ch += n_waves) {\n"
         "    const int64_t j0 = ch * 64 * GR;\n"
         "    bool live[GR];\n    int64_t row[GR];\n"
         "    #pragma unroll\n    for (int u = 0; u < GR; u++) {\n"
         "      const int64_t j = j0 + 64 * u + lane;\n"
         "      live[u] = j < n_sel;\n"
         "      // every lane loads (a lane past the last output reads the chunk's first\n"
         "      // index): a load behind a condition is waited for on its own\n"
         "      const uint32_t ri = __builtin_nontemporal_load(ws.idx + (live[u] ? j : j0));\n"
         "      row[u] = live[u] ? (int64_t)ri : 0;\n"
         "    }\n";
    if (!side.empty()) {
      o << "    int64_t sp[GR];  // place of the row in its slot's staged values\n"
           "    #pragma unroll\n    for (int u = 0; u < GR; u++) {\n"
           "      // lanes past the last output: row 0, slot 0, whose prefix is 0, so sp = 0\n"
           "      const int64_t sl = row[u] / (64 * RS);\n"
           "      const int64_t pre = (int64_t)__builtin_nontemporal_load(ws.wave_count + sl);\n"
           "      const int64_t jl = live[u] ? j0 + 64 * u + lane : 0;\n"
           "      sp[u] = sl * 64 * RS + (jl - pre);\n"
           "    }\n";
    }
    // raw values: staged predicate columns, the rest at the source rows
    for (int c : side) o << "    " << uint_of_width(in.in_width[c]) << " r" << c << "[GR];\n";
    for (int c : own) o << "    " << uint_of_width(in.in_width[c]) << " r" << c << "[GR];\n";
    o << "    #pragma unroll\n    for (int u = 0; u < GR; u++) {\n";
    // unconditional loads (sp and row are 0, in range, in the lanes past the last
    // output, whose values are never stored)
    for (int c : side) o << "      r" << c << "[u] = __builtin_nontemporal_load(s" << c << " + sp[u]);\n";
    for (int c : own) o << "      r" << c << "[u] = __builtin_nontemporal_load(c" << c << " + row[u]);\n";
    o << "    }\n";
    for (int c : bm_q) o << "    bool vb" << c << "[GR];\n";
    if (!bm_q.empty()) {
      o << "    #pragma unroll\n    for (int u = 0; u < GR; u++) {\n";
      // the bitmap exists (the source is written for a frame whose column has one)
      for (int c : bm_q)
        o << "      vb" << c << "[u] = ((((const GLOBAL uint8_t *)m" << c << ")[row[u] >> 3] >> (row[u] & 7)) & 1) != 0;\n";
      o << "    }\n";
    }
    o << "    #pragma unroll\n    for (int u = 0; u < GR; u++) {\n"
         "      const int64_t j = j0 + 64 * u + lane;\n";
    g.per_lane_bits = true;
    g.emit_inputs(q_reads, "      ", "");
    g.emit_nodes(*in.order_q, "      ");
    o << "      if (live[u]) {\n";
    g.emit_stores("j", "        ", "");
    o << "      }\n    }\n";
    if (any_valid) {
      for (int k = 0; k < n_out; k++)
        if (in.out_valid[k]) o << "    uint64_t ow" << k << " = 0;\n";
      o << "    #pragma unroll\n    for (int u = 0; u < GR; u++) {\n"
           "      if (j0 + 64 * u >= n_sel) break;\n";
      g.emit_vnode_bits(vs_q, "      ");
      for (int k = 0; k < n_out; k++) {
        if (!in.out_valid[k]) continue;
        o << "      { const uint64_t wd = __builtin_amdgcn_ballot_w64(live[u] && " << g.vbit(in.out_v[k]) << ");\n"
          << "        if (lane == u) ow" << k << " = wd;\n"
          << "        if (lane == 0) nn" << k << " += __builtin_popcountll(~wd & low_mask(n_sel - (j0 + 64 * u))); }\n";
      }
      o << "    }\n";
      o << "    if (lane < GR && j0 + 64 * lane < n_sel) {\n";
      for (int k = 0; k < n_out; k++)
        if (in.out_valid[k]) o << "      store_vword(q" << k << ", (j0 >> 6) + lane, n_sel, ow" << k << ");\n";
      o << "    }\n";
    }
    o << "  }\n";
  }
  // null counts: summed over the workgroup's four waves in LDS, then one atomic per
  // workgroup and output (atomics on one address serialize, and one per wave over
  // a large grid cost more than the rest of the kernel)
  std::vector<int> nk;
  for (int k = 0; k < n_out; k++)
    if (in.out_valid[k]) nk.push_back(k);
  res.null_outputs = (int)nk.size();
  if (!nk.empty()) {
    o << "  __shared__ unsigned long long s_nn[4][" << nk.size() << "];\n";
    for (size_t i = 0; i < nk.size(); i++)
      o << "  { unsigned long long t = nn" << nk[i] << ";\n"
        << "    for (int s = 32; s > 0; s >>= 1) t += __shfl_xor(t, s);\n"
        << "    if (lane == 0) s_nn[threadIdx.x >> 6][" << i << "] = t; }\n";
    o << "  __syncthreads();\n";
    for (size_t i = 0; i < nk.size(); i++)
      o << "  if (threadIdx.x == " << i << ") {\n"
        << "    const unsigned long long t = s_nn[0][" << i << "] + s_nn[1][" << i << "] + s_nn[2][" << i
        << "] + s_nn[3][" << i << "];\n"
        << "    if (t) atomicAdd(ws.null_counts + " << nk[i] << ", t);\n  }\n";
  }
  o << "}\n";
  res.src = o.str();
  res.name = kname;
  return res;
}

}  // namespace blz

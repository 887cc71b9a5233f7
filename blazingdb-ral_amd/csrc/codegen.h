/*
 * codegen.h — HIP source for one compiled plan (the compiled tier).
 *
 * The interpreter kernels (tile VM, register-file kernel) read an op stream and
 * dispatch on it per op. For a plan that runs more than once, the compiled tier
 * instead writes the plan out as straight-line HIP over registers — each input
 * column read once from HBM into registers, every SSA value a local, every
 * validity word a scalar — and compiles it at run time for gfx950 (jit.cpp). The
 * generated kernel keeps the interpreter's semantics exactly: it evaluates the
 * same SSA graph, op by op, through op_t<code> (device_common.h), and the same
 * validity program over 64-row words.
 *
 * Two kernel shapes:
 *  - projection (blz_cg_project): every row evaluated and stored, validity words
 *    written whole;
 *  - filter (+ projection): blz_cg_pred evaluates the predicate over every row from
 *    registers and writes the tile VM's keep words and slot counts, plus the kept
 *    rows' values of the predicate columns the projection reads, compacted per slot;
 *    the slot scan turns counts into offsets and kept-row indices; blz_cg_gather
 *    evaluates the projection of the kept rows, output-driven (staged values read
 *    densely, other columns at the source rows), outputs and validity words stored
 *    whole in row order.
 */
#ifndef BLZ_CODEGEN_H_
#define BLZ_CODEGEN_H_

#include <string>
#include <vector>

#include "ssa.h"

namespace blz {

struct CodegenInput {
  const Builder *B = nullptr;
  const std::vector<int> *order_p = nullptr;  // predicate nodes in evaluation order
  const std::vector<int> *order_q = nullptr;  // projection nodes in evaluation order
  bool compact = false;
  int stencil_node = -1;
  int stencil_v = V_ALL;
  int stencil_width = 1;  // bytes of the predicate value that decide "kept"
  // per output: value node, validity value, kind (0 evaluated, 1 bare input column),
  // StoreKind (evaluated) or the input column (bare), validity written
  std::vector<int> out_node, out_v, out_kind, out_store, out_input;
  std::vector<bool> out_valid;
  // per device input: LoadKind, bytes per row, bitmap present
  std::vector<int> in_load, in_width;
  std::vector<bool> in_valid;
  int rs = 8;  // 64-row groups per wave per step (rows per lane)
};

struct CodegenOutput {
  std::string src;   // kernel source (includes device_common.h)
  std::string name;  // extern "C" kernel name
  int rs = 8;
  int tile_rows = 0;  // filter kernel: rows per tile (256 * rs); projection: 64 * rs per wave step
  int side_bytes = 0;  // filter: bytes per row of the staged predicate columns
  int null_outputs = 0;  // outputs with validity (a null-count atomic per workgroup each)
};

// Throws std::runtime_error when the plan has a shape the generator does not take.
CodegenOutput generate_kernel_source(const CodegenInput &in);

// The text every generated source starts with (includes and shared helpers); the
// rest of a source is its kernels, named blz_cg_project, or blz_cg_pred and
// blz_cg_gather, and the RS / GR macros they use.
const char *codegen_prelude();

}  // namespace blz
#endif

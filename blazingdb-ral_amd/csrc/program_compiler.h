/*
 * program_compiler.h — host compiler from reference-format plans to the device
 * program blob of program.h.
 */
#ifndef BLZ_PROGRAM_COMPILER_H_
#define BLZ_PROGRAM_COMPILER_H_

#include <string>
#include <vector>

#include "../../include/blz/gdf_types.h"
#include "../../include/blz/blz_ral.h"
#include "program.h"

namespace blz {

// One plan in the reference's perform_operation format. Plan register c < n_in
// holds device input `in_dev[c]`.
struct PlanSpec {
  std::vector<int> in_dev;
  std::vector<int16_t> left, right, outs, final_pos;
  std::vector<int32_t> binops, unops;
  std::vector<gdf_scalar> lsc, rsc;
};

// What the compiler needs to know of a device input column.
struct DevInputInfo {
  void *data;
  const uint8_t *valid;
  int dtype;
  int null_count;
};

struct OutSpec {
  void *data;
  uint8_t *valid;
  int dtype;
  int kind;   // 0: final output `index` of the projection plan; 1: gathered device input `index`
  int index;
};

struct CompileRequest {
  int64_t n_rows = 0;
  std::vector<DevInputInfo> inputs;
  bool compact = false;
  const PlanSpec *pred = nullptr;   // compaction predicate; null -> bare-column stencil
  int bare_stencil_input = -1;
  const PlanSpec *proj = nullptr;
  std::vector<OutSpec> outputs;
  int scalar_mode = BLZ_SCALAR_INTENDED;
  int force_s = 0, force_r = 0;     // register-file shape override (autotuner)
  int force_t = 0;                  // staged tile rows override (autotuner); -1 = direct kernel
  int kernel = 0;                   // 0 automatic, 1 accumulator, 2 staged register file, 3 direct, 4 tile VM,
                                    // 5 compiled (generated kernel, run-time compiled)
  int lds_budget = 64 * 1024;       // dynamic LDS per workgroup the staged kernel may use
};

struct Compiled {
  std::vector<uint8_t> blob;
  int64_t n_rows = 0;
  int S = 8, R = 1;
  int n_slots = 0, n_vslots = 0, n_ops = 0, n_vops = 0;
  int T = 0;          // staged kernel tile rows (0: direct kernel)
  int lds_bytes = 0;  // staged kernel dynamic LDS per workgroup
  bool acc = false;   // accumulator-form program (S = temporaries)
  bool vm = false;    // tile VM program (S = temporaries, T = 256 R)
  int vm_slow = 0;  // op set the VM program needs: 0 common, 1 hot, 2 with out-of-line ops
  int lds_gather = 0;    // tile VM filter: LDS of the gather kernel
  int vm_bufb = 0, vm_tmp_off = 0;                      // tile VM LDS layout (projection / predicate)
  int vm_g_buf_off = 0, vm_g_tmp_off = 0, vm_g_cpre_off = 0;  // tile VM LDS layout (gather)
  int vm_g_wave_bytes = 0, vm_nvout = 0;  // gather pass buffer per wave, outputs with validity
  int vm_vr_off = 0, vm_g_vr_off = 0;     // validity registers (four waves' regions)
  bool vm_direct_pred = false;            // tile VM filter: predicate by launch_pred_direct
  int vm_direct_terms = 0;                //   with this many terms (the predicate's ops),
  int vm_direct_wmask = 0;                //   bit k: term k reads an 8-byte column
  // compiled tier (kernel 5): generated kernel source and name (codegen.h); the
  // blob is the register-file kernel's (header, input and output tables), T a
  // filter's rows per tile (4 slots of 64 R rows, as the tile VM's)
  bool jit = false;
  std::string jit_src, jit_name;  // projection kernel, or a filter's predicate kernel
  int jit_side_bytes = 0;         // filter: bytes per row of staged predicate columns
  int jit_null_outputs = 0;       // outputs with validity (null-count atomics per workgroup)
};

// Throws std::runtime_error with a message on an unsupported or malformed plan.
Compiled compile_program(const CompileRequest &req);

int unary_output_dtype(int l, int u);
int column_load_kind(int dtype);  // LoadKind of a column dtype (LD_NONE: not read)
int binary_output_dtype(int l, int r, int op);
uint64_t scalar_bits(const gdf_scalar &s, int mode);

}  // namespace blz
#endif

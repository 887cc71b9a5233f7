/*
 * program.h — the device program a reference plan is compiled into.
 *
 * The reference ships its plan to the GPU as ~16 parallel arrays and resolves the
 * operand/result kernel types (int64 vs binary64) inside the kernel with three
 * nested branches per op per row (interpreter_ops.cuh:473-505). Here the host
 * resolves everything once (program.cpp) and the kernel receives one packed blob:
 *
 *   ProgHeader | InCol[n_in] | OutCol[n_out] | DevOp[n_ops_p + n_ops_q] | VOp[...]
 *
 * Each DevOp names a fully typed operation, its operand conversions and its result
 * conversion, so the kernel's per-op dispatch is one wave-uniform switch. Validity
 * is a separate, constant-folded program over 64-row words (VOp).
 *
 * Shared by host C++ and device HIP code: plain structs, fixed widths.
 */
#ifndef BLZ_PROGRAM_H_
#define BLZ_PROGRAM_H_

#include <stdint.h>

namespace blz {

// Device register file limits. Slots are 8-byte (int64 / binary64 bits).
constexpr int kMaxSlots = 64;   // reference limit (interpreter_cpp.cu:86-217 handles <= 64)
constexpr int kMaxVSlots = 64;
constexpr int kVmVSlots = 32;  // tile VM: validity registers (LDS words per 64-row group)
// Tile VM staging buffers per workgroup: 2 overlaps a tile's loads with the
// evaluation of the one before; 1 halves the LDS per workgroup (more workgroups
// per CU, whose loads and evaluations overlap instead). Measured on MI355X, one
// buffer is faster (C1 filter 0.625 -> 0.593 ms at 100M rows; wide projections
// fit R = 4 tiles instead of R = 2: 1.4x).
constexpr int kTileBuffers = 1;
constexpr int kMaxOps = 256;    // per program

// Input load kinds: how a column widens into an 8-byte slot (interpreter_ops.cuh:386-452).
enum LoadKind : int32_t {
  LD_NONE = 0,  // dtype the reference never reads (slot holds 0)
  LD_I8, LD_I16, LD_I32, LD_I64, LD_F32, LD_F64,
  LD_U8,        // BOOL8 raw byte, used only for gathered (aliased) columns
};

// Output store kinds: narrowing writes (interpreter_ops.cuh:317-381).
enum StoreKind : int32_t {
  ST_NONE = 0,  // dtype the reference does not write
  ST_I8, ST_I16, ST_I32, ST_I64, ST_F32, ST_F64,
};

// Operand / result conversions.
enum Conv : uint8_t { CV_NONE = 0, CV_I2F = 1, CV_F2I = 2 };

// Fully typed device opcodes.
enum DevOpCode : uint8_t {
  // binary, computed in int64 (_I) or binary64 (_F)
  OP_ADD_I, OP_ADD_F, OP_SUB_I, OP_SUB_F, OP_MUL_I, OP_MUL_F, OP_DIV_I, OP_DIV_F,
  OP_MOD_I, OP_POW_I, OP_POW_F,
  OP_EQ_I, OP_EQ_F, OP_NE_I, OP_NE_F, OP_LT_I, OP_LT_F, OP_GT_I, OP_GT_F,
  OP_LE_I, OP_LE_F, OP_GE_I, OP_GE_F,
  OP_COALESCE,
  // boolean AND of two operands known to be 0/1 (an `AND` = GDF_MUL whose operands
  // are both comparison results: the product equals the bitwise and)
  OP_AND_B,
  // unary
  OP_FLOOR, OP_CEIL, OP_SIN, OP_COS, OP_ASIN, OP_ACOS, OP_TAN, OP_COTAN, OP_ATAN,
  OP_ABS, OP_LN, OP_LOG10, OP_NOT_I, OP_NOT_F,
  OP_YEAR32, OP_MONTH32, OP_DAY32, OP_YEAR64, OP_MONTH64, OP_DAY64,
  OP_HOUR64, OP_MINUTE64, OP_SECOND64, OP_ZERO,
  OP_IS_NULL, OP_IS_NOT_NULL,
  OP_IDENT,
  OP_COUNT_
};

// One device op: 8 bytes of control + two 8-byte immediates.
struct DevOp {
  uint8_t code;   // DevOpCode
  uint8_t dst;    // destination slot
  uint8_t a, b;   // operand slots (ignored when the matching *_IMM flag is set)
  uint8_t flags;  // bit0 A is immediate, bit1 B is immediate
  uint8_t conv;   // bits 0-1 A conversion, 2-3 B conversion, 4-5 result conversion
  uint8_t vcol;   // input column whose raw validity COALESCE / IS_[NOT_]NULL read; 0xFF = none
  uint8_t pad;
  uint64_t imm_a, imm_b;
};
static_assert(sizeof(DevOp) == 24, "DevOp layout");

constexpr uint8_t F_A_IMM = 1, F_B_IMM = 2;

// Accumulator form (the compiler's intermediate, lowered to VmOps): each op
// combines the accumulator with at most one operand, so results need no
// write-back and inputs are read straight from the LDS tile.
enum AccCode : uint8_t {
  A_LOAD = 0,   // acc = operand
  A_STORE,      // temp[oidx] = acc
  A_EMIT,       // output oidx <- acc (narrowing store at the row's output position)
  A_BIN,        // acc = sub(acc, operand), or sub(operand, acc) when F_REV
  A_UN,         // acc = sub(acc)
  A_CMPAND,     // acc = acc & (operand sub imm2), acc and the comparison both 0/1
};
enum OperandKind : uint8_t { OK_NONE = 0, OK_IMM = 1, OK_INPUT = 2, OK_TEMP = 3 };
constexpr uint8_t F_REV = 1;
struct AccOp {
  uint8_t code;   // AccCode
  uint8_t sub;    // DevOpCode of the operation (A_BIN / A_UN / comparison of A_CMPAND)
  uint8_t okind;  // OperandKind
  uint8_t oidx;   // input column / temp index (operand), temp (A_STORE), output (A_EMIT)
  uint8_t conv;   // bits 0-1 accumulator conversion, 2-3 operand conversion, 4-5 result conversion
  uint8_t flags;  // F_REV: the operand is the left-hand side
  uint8_t vcol;   // input column whose raw validity COALESCE / IS_[NOT_]NULL read; 0xFF = none
  uint8_t pad;
  uint64_t imm;   // immediate operand
  uint64_t imm2;  // A_CMPAND: the comparison's immediate
};
static_assert(sizeof(AccOp) == 24, "AccOp layout");

// Tile VM (vm_kernel): the accumulator form pre-decoded on the host so the
// kernel's dispatch is one switch per op with no further field decoding. The
// operand of an op is fetched by class: an immediate, a temporary, a column slice
// of the LDS tile (byte offset `off` within the tile buffer), or a column read
// straight from HBM for the rows a compaction keeps (input `arg2`).
// VM_EMIT's `arg` is the output, VM_STORE's the temporary.
enum VmCode : uint8_t {
  VM_LOAD = 0,  // acc = B
  VM_STORE,     // tmp[arg] = acc
  VM_EMIT,      // output arg <- acc; gen = StoreKind
  VM_ADD_I, VM_ADD_F, VM_SUB_I, VM_SUB_F, VM_RSUB_I, VM_RSUB_F, VM_MUL_I, VM_MUL_F,
  VM_AND_B,     // acc &= B (both 0/1)
  VM_CMP,       // acc = acc CMP B; arg = comparison DevOpCode
  VM_CMPAND,    // acc &= (B CMP imm2); arg = comparison DevOpCode
  VM_RANGE_I,   // acc &= (imm2 < B < imm) as int64 (bounds made exclusive on the host)
  VM_RANGE_F,   // acc &= (imm2 <(=) B <(=) imm) as binary64; gen bit 0: lower bound
                // inclusive, bit 1: upper bound inclusive
  VM_I2F, VM_F2I,                   // acc conversion
  VM_FLOOR, VM_CEIL, VM_ABS_F, VM_NOT_I, VM_NOT_F,
  VM_DIV_F, VM_RDIV_F, VM_DIV_I, VM_RDIV_I, VM_MOD_I, VM_RMOD_I,
  VM_COALESCE, VM_RCOALESCE,        // acc = valid(vcol) ? acc : B  /  valid(vcol) ? B : acc
  VM_IS_NULL, VM_IS_NOT_NULL,       // acc = !valid(vcol)  /  valid(vcol)
  VM_GENERIC,   // transcendentals, POWER, date parts: row by row, out of line;
                // arg = DevOpCode, gen = conversions / flags
  VM_COUNT_
};
enum VmCls : uint8_t {
  VC_NONE = 0, VC_IMM, VC_TMP,
  VC_L64, VC_L32I, VC_L32F, VC_L16I, VC_L8I, VC_L8U,   // LDS tile slice
  VC_G64, VC_G32I, VC_G32F, VC_G16I, VC_G8I, VC_G8U,   // HBM column (kept rows only)
};
// VM_GENERIC flag bits (`gen`): 0-1 operand conversion, 2-3 accumulator conversion,
// 4-5 result conversion, 6 operand is the left-hand side, 7 unary.
constexpr uint8_t VG_REV = 0x40, VG_UNARY = 0x80;
// Hot leaves: the common ops with an operand of a common class, each a
// straight-line case of the kernel's single dispatch switch (leaf = op * kHotCls
// + class). Everything else runs through the general path (kLeafGeneral), which
// decodes code / cls / cv.
enum VmHotOp : uint8_t {
  HO_LOAD = 0, HO_ADD_I, HO_ADD_F, HO_SUB_I, HO_SUB_F, HO_RSUB_I, HO_RSUB_F, HO_MUL_I, HO_MUL_F, HO_AND_B,
  HO_CMP,         // 12 comparisons, acc CMP B, in the order of kHotCmp
  HO_CMPAND = HO_CMP + 12,  // 12 comparisons, acc &= (B CMP imm2)
  HO_RANGE_I = HO_CMPAND + 12, HO_RANGE_F,
  HO_COUNT_
};
enum VmHotCls : uint8_t {
  HC_IMM = 0, HC_TMP, HC_L64, HC_L32I, HC_L32F,
  HC_L32I_F,  // int32 column widened to binary64
  HC_L64_F,   // int64 column converted to binary64
  HC_COUNT_
};
// Comparison order of the HO_CMP / HO_CMPAND leaves (DevOpCode values).
constexpr int kHotCmp[12] = {OP_LT_I, OP_LE_I, OP_GT_I, OP_GE_I, OP_EQ_I, OP_NE_I,
                             OP_LT_F, OP_LE_F, OP_GT_F, OP_GE_F, OP_EQ_F, OP_NE_F};
constexpr int kHotLeaves = HO_COUNT_ * HC_COUNT_;
constexpr int kLeafStore = kHotLeaves;      // tmp = acc (off: byte offset of the temporary)
constexpr int kLeafEmit = kHotLeaves + 1;   // output arg <- acc
constexpr int kLeafI2F = kHotLeaves + 2;    // acc int64 -> binary64
constexpr int kLeafNotI = kHotLeaves + 3;   // acc = (acc == 0)
constexpr int kLeafNotF = kHotLeaves + 4;   // acc = !(binary64 acc != 0)
constexpr int kLeafDivK = kHotLeaves + 5;   // acc = acc / imm (int64, imm not in {-1, 0, 1}): by
constexpr int kLeafModK = kHotLeaves + 6;   // acc = acc % imm   the magic multiplier imm2, shift gen
constexpr int kLeafGeneral = kHotLeaves + 7;
// Fused moves. A hot leaf or one of kLeafI2F .. kLeafModK may carry, in bits 16-18
// of its leaf word, the accumulator moves around it that would otherwise be ops of
// their own (each op costs one dispatch, whatever it does):
//   kFusePre    before the op, acc = operand of hot class (code word bits 20-23) at
//               byte offset (bits 0-19), as HO_LOAD does;
//   kFuseStore  after it, temporary (code word bits 24-31) = acc, as kLeafStore;
//   kFuseEmit   after it, output (gen word bits 8-15) <- acc with StoreKind (bits
//               16-23), as kLeafEmit.
// The code word (code, arg, cls, cv) is not read by these leaves; it then holds the
// fused fields.
constexpr int kLeafMask = 0xFFFF;
constexpr int kFusePre = 1 << 16;
constexpr int kFuseStore = 1 << 17;
constexpr int kFuseEmit = 1 << 18;
struct VmOp {
  int32_t leaf;   // dispatch id (set once the tile shape is known)
  int32_t off;    // VC_L*: byte offset of the column slice in the tile buffer;
                  // VC_TMP operand / VM_STORE: byte offset of the temporary (t * 8 * 256 R)
  uint8_t code;   // VmCode
  uint8_t arg;    // temp / output / comparison / DevOpCode
  uint8_t cls;    // VmCls of the operand (VC_NONE for ops without one)
  uint8_t cv;     // operand conversion (Conv) applied after the fetch
  uint8_t gen;    // VM_GENERIC flags; VM_EMIT: StoreKind
  uint8_t vcol;   // input whose raw validity COALESCE / IS_[NOT_]NULL read; 0xFF none
  uint8_t arg2;   // VC_G* / VC_L*: input column; VC_TMP: temporary
  uint8_t pad;
  uint64_t imm;   // VC_IMM operand
  uint64_t imm2;  // VM_CMPAND immediate
};
static_assert(sizeof(VmOp) == 32, "VmOp layout");

// Tile VM staging: one piece of up to 1 KiB of a tile's column (or bitmap) slice,
// moved HBM -> LDS by one wave (64 lanes x 16 bytes, direct to LDS).
struct Piece {
  const void *base;      // column data / bitmap base
  int32_t slice;         // bytes of the slice per full tile (T * width, or T / 8)
  int32_t off;           // offset of the piece within the slice
  int32_t lds;           // LDS offset of the piece within a tile buffer
  int32_t bytes;         // bytes of the piece (<= 1024)
  int32_t bitmap;        // 1: a bitmap slice (partial tile: ceil(rows / 8) bytes)
  int32_t width;         // data slice: bytes per row
};
static_assert(sizeof(Piece) == 32, "Piece layout");
constexpr int kPieceBytes = 1024;

// An output that is a bare input column: the projection / gather kernel copies the
// rows' bytes straight from the column (as materialize_column does).
struct CopyCol {
  const void *src;  // input column data
  void *dst;        // output data
  int32_t width;    // bytes per row
  int32_t pad0;
  int64_t pad1;
};
static_assert(sizeof(CopyCol) == 32, "CopyCol layout");

// Validity op over 64-row words.
enum VOpCode : uint8_t { VOP_AND = 0, VOP_OR = 1 };
struct VOp {
  uint8_t code, dst, a, b;
};

// Input column descriptor.
struct InCol {
  const void *data;
  const uint8_t *valid;  // raw bitmap (may be null = all valid)
  int32_t load;          // LoadKind
  int32_t vslot;         // validity slot this column's words load into, or -1 (forced / unused)
  int32_t lds_off;       // staged kernel: LDS offset of the tile's values
  int32_t lds_voff;      // staged kernel: LDS offset of the tile's bitmap words (-1: none)
};


// Output descriptor. `src_kind`: 0 = program slot (of the program that owns the
// output), 1 = raw input column (a gathered alias).
struct OutCol {
  void *data;
  uint8_t *valid;        // null = no validity written
  int32_t store;         // StoreKind
  int16_t slot;          // data slot (src_kind 0) or input column (src_kind 1)
  int16_t src_kind;
  int32_t vsrc;          // >= 0 validity slot; -1 = constant all-valid; -2 = constant all-null
  int32_t vbuf;          // staged compaction: index of this output's LDS validity buffer
};

// Per-launch device scratch. Every counter must be zero at launch.
struct Workspace {
  unsigned long long *null_counts;  // [n_out] nulls per output (outputs with validity)
  unsigned int *tile_counter;       // compaction: dynamic tile ids
  unsigned long long *tile_state;   // compaction: [n_tiles] look-back state words
  unsigned long long *n_sel;        // compaction: rows kept
  // tile VM filter (written before they are read; no zeroing needed)
  uint64_t *keep;                   // keep word per 64-row group
  uint32_t *wave_count;             // kept rows per (tile, wave) slot
  uint32_t *idx;                    // row of each kept row, in output order (n_sel entries)
  // compiled filter: the kept rows' values of the predicate columns the projection
  // reads, compacted per slot (slot s's kept rows at s * slot_rows ..), one array of
  // n_tiles * tile_rows values per column, back to back
  uint8_t *side;
};

struct ProgHeader {
  int64_t n_rows;
  int32_t n_in, n_out;
  int32_t n_ops_p, n_ops_q;      // predicate ops (compaction mode), projection ops
  int32_t n_vops_p, n_vops_q;
  int32_t n_vin;                 // validity slots loaded from input bitmaps [0, n_vin)
  int32_t n_slots, n_vslots;     // register-file sizes used
  int32_t stencil_slot;          // compaction: slot holding the predicate value
  int32_t stencil_width;         // compaction: bytes of the predicate value that decide "non-zero"
  int32_t stencil_vsrc;          // compaction: validity of the predicate (same codes as OutCol::vsrc)
  int32_t any_valid_out;         // some output has validity
  int32_t tile_rows;             // staged kernel: rows per tile (T)
  int32_t n_chunks;              // staged kernel: HBM -> LDS pieces per tile
  int32_t lds_vbits_off;         // staged kernel: LDS offset of compacted-validity words
  int32_t n_valid_out;           // outputs with validity
  int32_t lds_bytes;             // staged kernel: dynamic LDS per workgroup
  int32_t acc_mode;              // ops are AccOp (1: accumulator form) / VmOp (2: tile VM) instead of DevOp
  int32_t n_temps;               // accumulator form: temporaries used
  int32_t buf_bytes;             // tile VM: bytes of one of the two tile buffers
  int32_t n_pieces;              // tile VM: staging pieces per tile
  int32_t off_pieces;            // tile VM: offset of the Piece table
  int32_t n_gcols;               // tile VM filter: columns the gather pass stages (Piece: base, width, lds)
  int32_t n_copy, off_copy;      // tile VM: bare-column outputs (CopyCol table)
  int32_t off_gcols;             // tile VM filter: offset of that table
  int32_t gbuf_bytes;            // tile VM filter: bytes of the gather pass buffer (256 R rows)
  int32_t lds_tmp_off;           // tile VM: LDS temporaries (projection / predicate kernels)
  int32_t g_tmp_off;             // tile VM filter: LDS temporaries of the gather kernel
  int32_t g_kw_off;              // tile VM filter: gather kernel's keep words + counts
  int32_t g_buf_off;             // tile VM filter: gather kernel's pass buffer
  // offsets (bytes from the blob start)
  int32_t off_in, off_out, off_ops, off_vops, off_chunks;
  int32_t blob_bytes;
};

}  // namespace blz
#endif

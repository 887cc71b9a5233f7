/*
 * kernels.h — host-callable launchers of the gfx950 kernels (kernels.hip).
 */
#ifndef BLZ_KERNELS_H_
#define BLZ_KERNELS_H_

#include <hip/hip_runtime.h>

#include <mutex>
#include <stdint.h>

#include "program.h"

namespace blz {


// Register-file interpreter (project or fused filter+project/gather), the general
// fallback for plans outside the tile VM's limits.
// (s, r) in {(8,1), (8,2), (16,1), (32,1), (64,1)}: s slots, r rows per lane;
// lds_bytes must be 0.
hipError_t launch_interp(int s, int r, bool compact, int grid, int lds_bytes, hipStream_t st,
                         const uint8_t *blob, Workspace ws);
int interp_blocks_per_cu(int s, int r, bool compact, int lds_bytes);
// Tile VM launch descriptor (device memory): one compiled program, or a batch of
// independent programs processed by one persistent launch. Global tile and
// scan-chunk ids run over the programs in order (gather chunks likewise, from the
// programs' kept-row counts); each program keeps its own workspace. The LDS
// offsets are the maxima over the programs.
struct VmPlanRef {
  const uint8_t *blob;   // ProgHeader | ... of the program (acc_mode 2)
  Workspace ws;          // the program's workspace
  int64_t tile0;         // first global tile of the program
  int64_t chunk0;        // first global slot-scan chunk
};
struct VmBatch {
  int32_t n_plans;
  int32_t bufb;                      // stride of the two tile buffers
  int32_t tmp_off;                   // LDS temporaries of the projection / predicate kernels
  int32_t g_buf_off, g_tmp_off;     // gather kernel: pass buffers, temporaries
  int32_t g_cpre_off;                // gather kernel: per-program chunk prefix (n_plans + 1 int64)
  int32_t g_wave_bytes;              // gather kernel: one wave's pass buffer (64 R rows)
  int32_t vr_off, g_vr_off;          // validity registers (projection / predicate, gather kernel)
  int32_t vr_wave_words;             // one wave's validity registers (registers x R words)
  int64_t n_tiles, n_chunks;        // totals over the programs
  unsigned int *chunk_counter;       // slot scan: dynamic chunk ids (zero before the scan)
  const VmPlanRef *plans;            // [n_plans]
};

// Tile VM (vm_device.h): r rows per lane in {8, 4, 2}; tile = 256 r P rows;
// lds_bytes: two tile buffers, validity registers and temporaries (modes 0, 1), or
// the gather kernel's pass buffers, validity registers, temporaries and chunk
// prefix (mode 2). ops: the op set compiled in (0 common, 1 hot, 2 hot +
// out-of-line: transcendentals, POWER, date parts).
// mode: 0 projection, 1 predicate (keep words + slot counts), 2 gather of the kept
// rows (after launch_slot_scan).
hipError_t launch_vm(int r, int mode, int ops, int grid, int lds_bytes, hipStream_t st, const VmBatch *bd);
int vm_blocks_per_cu(int r, int mode, int ops, int lds_bytes);
// The predicate of a single program whose ops are compare-and / range leaves over
// columns (Compiled::vm_direct_pred), from registers: same keep words and slot
// counts as launch_vm mode 1.
// n_terms <= 3; wmask bit k: term k reads an 8-byte column.
hipError_t launch_pred_direct(const VmBatch *bd, int n_terms, int wmask, int grid, hipStream_t st);
// Exclusive prefix of every program's slot counts (ws.wave_count, written over the
// counts; total to
// ws.n_sel) and, from it and the keep words, the row of every kept row in output
// order (ws.idx): chunks of scan_chunk_slots(tile rows) slots, combined within a
// program by a single-pass decoupled look-back over its ws.tile_state (zero before
// the scan).
hipError_t launch_slot_scan(const VmBatch *bd, int64_t n_chunks, hipStream_t st);
constexpr int64_t kScanChunk = 512;  // most slots per scan chunk
// Slots per scan chunk of a program with `tile_rows`-row tiles (four slots per
// tile, tile_rows / 256 keep words per slot): 2048 keep words (131072 rows) per
// chunk, at most kScanChunk slots, so a program's scan has as many workgroups
// whatever its slot size.
constexpr int64_t scan_chunk_slots(int64_t tile_rows) {
  return tile_rows / 256 <= 4 ? kScanChunk : 2048 / (tile_rows / 256);
}
// Seeded synthetic column in place (dist: 0 int uniform [a,b], 1 real uniform [a,b),
// 2 normal(a,b), 3 lognormal(a,b), 4 k/100 for k in [a,b], 5 a + row).
hipError_t launch_synth(void *data, uint8_t *valid, int64_t n, int dtype, uint64_t seed, int dist, double a,
                        double b, double null_rate, hipStream_t st);
// Zeroes `n` byte ranges (dst[i], bytes[i]; 8-byte multiples, 8-byte aligned) in
// one launch.
hipError_t launch_fill_ranges(void *const *dst, const int64_t *bytes, int n, int grid, hipStream_t st);
// materialize_column: value width and index width (a signed integer) in {1,2,4,8}.
hipError_t launch_gather(int width, const void *in, const uint8_t *in_valid, int64_t n_in, const void *idx,
                         int idx_width, int64_t n_idx, void *out, uint8_t *out_valid,
                         unsigned long long *null_count, int grid, hipStream_t st);
hipError_t launch_sequence(int32_t *out, int64_t n, int32_t init, int32_t step, int grid, hipStream_t st);
hipError_t launch_fill_zero(void *p, int64_t bytes, int grid, hipStream_t st);
hipError_t launch_null_count(const uint8_t *valid, int64_t n, unsigned long long *out, int grid,
                             hipStream_t st);

// Reduction of a column's valid rows (op 0 SUM, 1 MIN, 2 MAX; load: its LoadKind):
// parts[0..grid) the workgroups' partials, parts[grid] the result (value as the
// interpreter's 64-bit pattern: int64, or binary64 for float columns; cnt = valid
// rows). Deterministic for a given grid; reduce_grid gives the grid for a row
// count (every grid from 1 to kRedMaxGrid is valid). The data of a column of
// width w is w-byte aligned.
struct RedPart {
  uint64_t v;
  unsigned long long cnt;
};
constexpr int kRedLoads = 4;        // 16-byte loads in flight per lane
constexpr int kRedMaxGrid = 2048;   // workgroups of the first pass
int reduce_grid(int load, int64_t n);
hipError_t launch_reduce(int op, const void *data, int load, const uint8_t *valid, int64_t n, RedPart *parts,
                         int grid, hipStream_t st);

// dst[0..bytes) = src[0..bytes); bytes a multiple of 16, both 16-byte aligned.
hipError_t launch_copy(void *dst, const void *src, int64_t bytes, int grid, hipStream_t st);

// The allocation-free entry point's scratch on the current device: a file-scope
// device array of *bytes bytes (each device holds its own copy until the process
// ends), resolved to its device address once per device. *lock serialises the
// calls that use it on one device.
hipError_t noalloc_scratch(uint8_t **ptr, size_t *bytes, std::mutex **lock);

int blob_capacity();   // bytes of LDS the program blob may occupy
int block_threads();   // threads per workgroup of every kernel

}  // namespace blz
#endif

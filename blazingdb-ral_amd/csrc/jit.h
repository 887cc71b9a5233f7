/*
 * jit.h — run-time compilation of generated kernels (codegen.h) for gfx950 with
 * hiprtc, and their launch.
 *
 * A source text is compiled once per process (the code object is kept, keyed by
 * the text) and loaded once per device; programs whose plans generate the same
 * text share the kernel. The kernels take the program blob and the workspace, as
 * the interpreter kernels do, so a compiled program runs on any frame of the same
 * shape.
 */
#ifndef BLZ_JIT_H_
#define BLZ_JIT_H_

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "program.h"

namespace blz {

// The kernel `name` of `src` on the current device: compiled and loaded on first
// use. Throws std::runtime_error carrying the compiler's log on failure.
hipFunction_t jit_function(const std::string &src, const std::string &name);

// Compiles `src` (cached as jit_function caches it) without loading it: needs no
// device. Throws as jit_function does.
void jit_compile_only(const std::string &src);

// Compiles the generated sources of `srcs` that are not compiled yet as one
// program: the prelude once, then each source's kernels under names of their own,
// so the compiler's fixed cost (most of a small plan's compile time) is paid once.
// jit_function then finds each source compiled (its kernels load from the shared
// code object). A source that is not a generated one, or a combined compile that
// fails, is compiled on its own; a source that does not compile is left for
// jit_function to report. Needs no device.
void jit_compile_many(const std::vector<std::string> &srcs);

// Workgroups of 256 threads of `f` that stay resident on one CU.
int jit_blocks_per_cu(hipFunction_t f);

hipError_t launch_jit(hipFunction_t f, int grid, hipStream_t st, const uint8_t *blob, const Workspace &ws);

// Compilations done and reused since the process began; milliseconds spent compiling.
void jit_counters(long long *compiles, long long *hits, double *compile_ms);

}  // namespace blz
#endif

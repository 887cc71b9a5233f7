/*
 * blz_ral.h — C ABI of the MI355X expression-evaluation / filter / gather path.
 *
 * Plain C, plain pointers and sizes. Every device pointer is HBM memory of the
 * current HIP device. Every entry point returns BLZ_OK (0) or a negative status and
 * writes a message into `err` (when err != NULL, err_len > 0); nothing throws.
 * `stream` is a hipStream_t (NULL = the default stream). Calls enqueue their work
 * on `stream` and return when it is complete (the reference's calls are blocking:
 * interpreter_cpp.cu:282-285), except blz_program_launch, which only enqueues.
 *
 * The C++ operator API of the reference (perform_operation, evaluate_expression,
 * process_filter, execute_project_plan, materialize_column, gdf_sequence) is
 * provided over this ABI by include/blz/ral.hpp.
 */
#ifndef BLZ_RAL_H_
#define BLZ_RAL_H_

#include <stddef.h>
#include <stdint.h>

#include "gdf_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define BLZ_OK 0
#define BLZ_ERR_INVALID (-1)     /* malformed plan / arguments (std::runtime_error in the reference) */
#define BLZ_ERR_HIP (-2)         /* HIP runtime error (CheckCudaErrors in Utils.cuh:25-55)          */
#define BLZ_ERR_OOM (-3)         /* allocation failed or over budget (cuDF::Allocator OutOfMemory)   */
#define BLZ_ERR_UNSUPPORTED (-4) /* operator / dtype outside this path (e.g. string literals)        */

/* Float-literal widening (interpreter_ops.cuh:156-175). INTENDED keeps the literal's
 * binary64 value; SHIPPED reproduces the shipped `static_cast<int64_t>(value)` whose
 * bits the kernel then reads as a double. */
#define BLZ_SCALAR_INTENDED 0
#define BLZ_SCALAR_SHIPPED 1

typedef void *blz_stream_t;

/* ---- plan level ------------------------------------------------------------ */

/* perform_operation (src/Interpreter/interpreter_cpp.h:25-37, implementation
 * interpreter_cpp.cu:219-286): evaluate a register-file plan of n_operations ops
 * over input_columns[0..n_inputs) (all input_columns[0]->size rows) into the
 * caller-allocated output_columns. Register c < n_inputs holds input c; register
 * -2 = scalar, -3 = null scalar, right register -1 = unary op. Writes data and, for
 * outputs with a validity buffer, 64-row validity words. out_null_counts[k] (if
 * non-NULL) receives output k's null count; output_columns[k]->null_count is left
 * unchanged, as in the reference (interpreter_cpp.cu:280). Plans use at most 64
 * registers (interpreter_cpp.cu:86-217). */
int blz_perform_operation(gdf_column *const *output_columns, int n_outputs,
                          gdf_column *const *input_columns, int n_inputs,
                          const column_index_type *left_inputs,
                          const column_index_type *right_inputs,
                          const column_index_type *outputs, int n_operations,
                          const column_index_type *final_output_positions,
                          const gdf_binary_operator *operators,
                          const gdf_unary_operator *unary_operators,
                          const gdf_scalar *left_scalars, const gdf_scalar *right_scalars,
                          int scalar_mode, int32_t *out_null_counts, blz_stream_t stream,
                          char *err, size_t err_len);

/* As blz_perform_operation, allocating no device memory (no counterpart in the
 * reference: for callers that cannot call hipMalloc, such as code under HIP graph
 * capture, tight serving loops and harnesses that forbid it). The plan's launch
 * buffers (program blob, launch descriptor, null counts, counters: at most 16 KiB,
 * whatever the row count) live in a file-scope device array that each GPU holds
 * until the process ends; a plan whose buffers would not fit gives BLZ_ERR_INVALID.
 * Calls on one GPU share that array and keep the last call's contents: the library
 * clears what must start at zero and serialises the calls on one GPU (each one
 * blocks until its work is done), so they never use the array at the same time. */
int blz_perform_operation_noalloc(gdf_column *const *output_columns, int n_outputs,
                                  gdf_column *const *input_columns, int n_inputs,
                                  const column_index_type *left_inputs,
                                  const column_index_type *right_inputs,
                                  const column_index_type *outputs, int n_operations,
                                  const column_index_type *final_output_positions,
                                  const gdf_binary_operator *operators,
                                  const gdf_unary_operator *unary_operators,
                                  const gdf_scalar *left_scalars, const gdf_scalar *right_scalars,
                                  int scalar_mode, int32_t *out_null_counts, blz_stream_t stream,
                                  char *err, size_t err_len);

/* ---- expression level (host planner + kernels) ------------------------------- */

/* evaluate_expression (src/LogicalFilter.h:20-23, LogicalFilter.cpp:271-371):
 * plan `expression` (Calcite text, e.g. ">($1, 5)") over the frame's columns and
 * evaluate it into the caller-allocated `output` (its dtype decides the narrowing).
 * A bare "$k" expression copies column k into `output` (LogicalFilter.cpp:281-287;
 * output must then have column k's dtype). *out_null_count (if non-NULL) receives
 * the output's null count; output->null_count is not written (LogicalFilter.cpp:368). */
int blz_evaluate_expression(gdf_column *const *frame, int n_frame, const char *expression,
                            gdf_column *output, int scalar_mode, int32_t *out_null_count,
                            blz_stream_t stream, char *err, size_t err_len);

/* process_filter (src/CalciteInterpreter.cpp:509-573): `query_part` is a
 * "LogicalFilter(condition=[...])" line (or the bare condition). Keeps the rows
 * whose predicate is non-zero and valid, in row order, for EVERY frame column.
 * out_columns[0..n_frame) are filled with callee-allocated columns (free each with
 * blz_free_column): dtype of the input, size = rows kept, validity iff the input
 * has validity, null_count set (ColumnManipulation.cu:105-113). *n_kept = rows kept. */
int blz_process_filter(gdf_column *const *frame, int n_frame, const char *query_part,
                       gdf_column *out_columns, int64_t *n_kept, int scalar_mode,
                       blz_stream_t stream, char *err, size_t err_len);

/* execute_project_plan (src/CalciteInterpreter.h:61-63, CalciteInterpreter.cpp:123-336):
 * `query_part` is "LogicalProject(name=[expr], ...)". Evaluated expressions share
 * one plan and one launch; "$k" outputs alias frame column k (no copy, is_alias[i]
 * = 1; do not free those). Writes *n_out columns (<= max_out) with null counts set
 * (CalciteInterpreter.cpp:332-335). */
int blz_execute_project(gdf_column *const *frame, int n_frame, const char *query_part,
                        gdf_column *out_columns, int32_t *is_alias, int max_out, int *n_out,
                        int scalar_mode, blz_stream_t stream, char *err, size_t err_len);

/* process_aggregate without GROUP BY (src/operators/GroupBy.h:18, GroupBy.cpp:531-588
 * with compute_aggregations :292-336 and aggregations_without_groupby :230-290):
 * `query_part` is a "LogicalAggregate(group=[{}], name=[OP(input)], ...)" line, OP
 * one of SUM, MIN, MAX, AVG, COUNT; input "$k", an expression (evaluated first) or
 * empty (COUNT(): column 0). out_columns[0..*n_out) receive one-row callee-allocated
 * columns (free with blz_free_column): COUNT -> INT64 rows - null_count (always
 * valid); SUM -> INT64 (integer input, wrapping) or FLOAT64 over the valid rows;
 * MIN / MAX -> the input dtype; AVG -> FLOAT64 sum / count. SUM, MIN, MAX, AVG of no
 * valid row are null. A column whose null_count is 0 counts every row as valid, and
 * an evaluated input has null_count 0 whatever its bitmap holds, as in the reference
 * (evaluate_expression does not update it, LogicalFilter.cpp:368): COUNT(expr) is the
 * row count and AVG(expr) divides the sum of every row by it. A GROUP BY
 * (group=[{...}] not empty) or COUNT_DISTINCT gives BLZ_ERR_UNSUPPORTED. */
int blz_process_aggregate(gdf_column *const *frame, int n_frame, const char *query_part,
                          gdf_column *out_columns, int max_out, int *n_out, int scalar_mode,
                          blz_stream_t stream, char *err, size_t err_len);

/* LogicalProject over LogicalFilter in ONE pass over the frame: the result equals
 * blz_process_filter followed by blz_execute_project on its output (projection of
 * row-local expressions commutes with the filter). Every output is callee-allocated
 * (aliases are gathered copies, is_alias[i] = 0). */
int blz_filter_project(gdf_column *const *frame, int n_frame, const char *filter_query_part,
                       const char *project_query_part, gdf_column *out_columns, int max_out,
                       int *n_out, int64_t *n_kept, int scalar_mode, blz_stream_t stream,
                       char *err, size_t err_len);

/* ---- gather / mask / sequence ------------------------------------------------ */

/* materialize_column (src/ColumnManipulation.cuh:16-18, ColumnManipulation.cu:83-150):
 * output[i] = input[row_indices[i]] for row_indices[i] >= 0; output validity bit i =
 * input validity bit row_indices[i] (when output->valid != NULL); sets
 * output->null_count. row_indices holds signed integers of its dtype's width (1, 2, 4
 * or 8 bytes, as ColumnManipulation.cu:120-134 dispatches them); output->size must
 * equal row_indices->size. */
/* materialize_column over n_columns columns of one table sharing row_indices, as one
 * blocking call (no reference counterpart as one call: its callers loop over
 * materialize_column, JoinOperator.cpp:130-135, OrderBy.cpp:66,
 * distribution/primitives.cpp:239,426,667). Each outputs[c] as
 * blz_materialize_column(inputs[c], outputs[c], row_indices); every input has the
 * same size. */
int blz_materialize_columns(const gdf_column *const *inputs, gdf_column *const *outputs, int n_columns,
                            const gdf_column *row_indices, blz_stream_t stream, char *err, size_t err_len);
int blz_materialize_column(const gdf_column *input, gdf_column *output,
                           const gdf_column *row_indices, blz_stream_t stream, char *err,
                           size_t err_len);

/* gdf_sequence (src/Utils.cuh:187-190, Utils.cu:4-12): data[i] = init + i * step. */
int blz_gdf_sequence(int32_t *data, size_t size, int32_t init_val, int32_t step,
                     blz_stream_t stream, char *err, size_t err_len);

/* cudf::apply_boolean_mask as process_filter uses it (CalciteInterpreter.cpp:543):
 * keeps input rows whose stencil value is non-zero and valid, in order, into a
 * callee-allocated `out` (free with blz_free_column). */
int blz_apply_boolean_mask(const gdf_column *input, const gdf_column *stencil, gdf_column *out,
                           blz_stream_t stream, char *err, size_t err_len);

/* gdf_column_cpp::update_null_count (src/GDFColumn.cu:519-528). */
int blz_update_null_count(gdf_column *column, blz_stream_t stream, char *err, size_t err_len);

/* ---- prepared programs (no reference counterpart: plan once, launch many) ---- */

typedef struct blz_program_s *blz_program;

/* Compiles a fused filter+project (filter_query_part may be NULL for a pure
 * projection; project_query_part NULL keeps every frame column) against fixed
 * device buffers the program allocates and owns: out_columns[0..*n_out) receive
 * them (capacity: every frame row; sizes valid after blz_program_result).
 * kernel: 0 automatic (the compiled tier, 5, when s and t are 0 and the plan compiles;
 * else the tile VM when the plan fits it, else the direct kernel),
 * 3 direct register-file kernel, 4 tile VM (direct-to-LDS tile staging;
 * a filter runs as predicate, slot-scan and gather launches), 5 compiled: the
 * plan written out as a kernel and compiled for the device at creation (one
 * launch; a filter keeps rows in one pass with a decoupled look-back). r / t (0 =
 * automatic) force the tile VM's rows per lane (2, 4, 8) and tile rows (256 r P);
 * s / r force the direct kernel's slots x rows (8,1) (8,2) (16,1) (32,1) (64,1). */
int blz_program_create(blz_program *prog, gdf_column *const *frame, int n_frame,
                       const char *filter_query_part, const char *project_query_part,
                       gdf_column *out_columns, int max_out, int *n_out, int scalar_mode,
                       int kernel, int s, int r, int t, char *err, size_t err_len);
/* Enqueues one execution on `stream` (no allocation, no synchronisation). */
int blz_program_launch(blz_program prog, blz_stream_t stream, char *err, size_t err_len);
/* Waits for the last launch and reads back rows kept and per-output null counts. */
int blz_program_result(blz_program prog, blz_stream_t stream, int64_t *n_kept,
                       int32_t *null_counts, char *err, size_t err_len);
/* Launches `iters` times with a HIP event pair around each kernel on `stream`;
 * returns the mean kernel time in milliseconds. */
int blz_program_time(blz_program prog, blz_stream_t stream, int iters, double *mean_ms,
                     char *err, size_t err_len);
/* As blz_program_time, per kernel launch of the program: stage_ms[0..*n_stages)
 * (3 for a tile-VM filter: predicate, slot scan, gather of the kept rows; else 1). */
int blz_program_time_stages(blz_program prog, blz_stream_t stream, int iters, double *stage_ms,
                            int *n_stages, char *err, size_t err_len);
/* Register-file shape and sizes chosen for the program. */
int blz_program_info(blz_program prog, int *s, int *r, int *n_ops, int *n_vops, int *n_slots,
                     int *grid);
/* Tile rows (0 = unstaged kernel), dynamic LDS bytes per workgroup and the kernel
 * kind chosen (1 accumulator, 2 staged register file, 3 unstaged). */
int blz_program_tile(blz_program prog, int *tile_rows, int *lds_bytes, int *kernel);
/* How a tile-VM filter evaluates its predicate: *direct = 1 when it reads the
 * columns into registers (a conjunction of at most three column-versus-constant
 * comparisons, validity an AND of column bitmaps), with *terms terms (*wmask bit k:
 * term k reads an 8-byte column). */
int blz_program_predicate(blz_program prog, int *direct, int *terms, int *wmask);
int blz_program_destroy(blz_program prog);
/* The kernel source the compiled tier (kernel 5) writes for a filter / projection
 * over a frame of this shape (dtypes, bitmaps present, null counts, rows); nothing
 * is allocated or launched and column data is not read. *src_len receives the full
 * length; src (src_cap bytes) the text, truncated. No reference counterpart. */
int blz_program_source(gdf_column *const *frame, int n_frame, const char *filter_query_part,
                       const char *project_query_part, int scalar_mode, char *src, size_t src_cap,
                       size_t *src_len, char *err, size_t err_len);
/* Compiles a generated kernel source for gfx950 without a device (the compiled
 * tier's compiler, for inspection and tests). No reference counterpart. */
int blz_jit_compile(const char *src, char *err, size_t err_len);
/* Compiles n generated sources that are not compiled yet as one program (the
 * compiler's fixed cost paid once; programs created later over these plans find
 * their kernels compiled). Sources that do not compile are left to be reported
 * when used. No device needed. No reference counterpart. */
int blz_jit_compile_many(const char *const *srcs, int n, char *err, size_t err_len);
/* Compilations done and cache hits since the process began, and milliseconds spent
 * compiling. No reference counterpart. */
int blz_jit_counters(long long *compiles, long long *hits, double *compile_ms);

/* ---- batches (no reference counterpart: many independent queries, one launch) ----
 *
 * A batch compiles n independent queries (each its own frame, optional filter and
 * projection, the same semantics as blz_program_create) and executes all of them
 * with one persistent launch per kernel kind: a projection kernel over every
 * projection-only query's tiles, then the predicate, slot-scan and gather kernels
 * over every filtered query's tiles, regardless of n. Results stay in
 * batch-owned device columns until the batch is destroyed.
 *
 * r / t: rows per lane (2, 4 or 8) and tile rows of every query; r = 0 runs the
 * autotuner, which times each candidate shape on the batch itself (a few
 * executions each) and keeps the fastest, then compiles the largest queries (at
 * least 2^20 rows, at most 40, in one run-time compilation) for the compiled tier
 * and runs each of them as a launch of its own where that times faster than the
 * tile VM (the whole batch is timed again, and kept only if faster); r = 0 with
 * t < 0 runs every such query that compiles on the compiled tier, untimed.
 * use_graph != 0
 * captures the launch sequence into a HIP graph once and replays it on every
 * blz_batch_launch. */
typedef struct blz_batch_s *blz_batch;
typedef struct {
  gdf_column *const *frame;
  int n_frame;
  const char *filter_query_part;   /* NULL: no filter */
  const char *project_query_part;  /* NULL: every frame column */
  int scalar_mode;
} blz_query;

int blz_batch_create(blz_batch *batch, const blz_query *queries, int n_queries, int r, int t, int use_graph,
                     char *err, size_t err_len);
/* Enqueues one execution of every query on `stream` (no allocation). */
int blz_batch_launch(blz_batch batch, blz_stream_t stream, char *err, size_t err_len);
/* Waits for `stream`, then describes query q's results: rows kept (rows for a
 * projection), per-output null counts and the output columns (views owned by the
 * batch, sized to the rows kept). */
int blz_batch_result(blz_batch batch, blz_stream_t stream, int q, int64_t *n_kept, int32_t *null_counts,
                     gdf_column *out_columns, int max_out, int *n_out, char *err, size_t err_len);
/* Mean milliseconds of one batch execution over `iters` (HIP events on `stream`). */
int blz_batch_time(blz_batch batch, blz_stream_t stream, int iters, double *mean_ms, char *err, size_t err_len);
/* r, tile rows, number of queries, whether a graph is replayed, autotuner's
 * candidates timed. */
int blz_batch_info(blz_batch batch, int *r, int *t, int *n_queries, int *graph, int *tuned_candidates);
/* The kernel kind running query q (as blz_program_tile reports it: 4 tile VM,
 * 5 compiled tier, ...) and whether it runs inside a shared launch (*shared = 1)
 * or as a launch of its own. */
int blz_batch_query_kernel(blz_batch batch, int q, int *kernel, int *shared);
int blz_batch_destroy(blz_batch batch);

/* ---- serving (no reference counterpart; the reference runs one detached thread
 * per query, CalciteInterpreter.cpp:929) ----
 *
 * A request loop over device-resident tables: submitted queries (the frame's
 * columns must stay valid until the query is released) are admitted in arrival
 * order while their estimated outputs + workspace fit hbm_budget_bytes and one of
 * n_streams stream slots is free, batched (up to max_batch per launch) and
 * launched, so batches admitted while others run overlap on the device. Results
 * are identical to running each query alone. r: rows per lane of every batch.
 * Outputs and workspaces come from the server's own device arena (4 GiB chunks,
 * reserved as admitted queries need them and held until blz_server_destroy);
 * the device's stream-ordered memory pool is left as it is. */
typedef struct blz_server_s *blz_server;
int blz_server_create(blz_server *server, size_t hbm_budget_bytes, int max_batch, int n_streams, int r,
                      char *err, size_t err_len);
int blz_server_submit(blz_server server, const blz_query *query, int64_t *ticket, char *err, size_t err_len);
/* Blocks until the query has run; describes its results like blz_batch_result and
 * its latency (submission to completion observed by the loop). */
int blz_server_wait(blz_server server, int64_t ticket, int64_t *n_kept, int32_t *null_counts,
                    gdf_column *out_columns, int max_out, int *n_out, double *latency_ms, char *err,
                    size_t err_len);
/* Frees the query's outputs and returns their budget. */
int blz_server_release(blz_server server, int64_t ticket);
/* Submits queries[i] at arrival_s[i] seconds after the call (a client thread),
 * waits for each, and reports per-query latency and rows kept. tickets_out NULL:
 * each query is released once it has completed; otherwise tickets_out[i] receives
 * query i's ticket (0 if its submission failed) and its results stay resident
 * until the caller reads them (blz_server_wait) and releases it. */
int blz_server_replay(blz_server server, const blz_query *queries, const double *arrival_s, int n,
                      double *latency_ms, int64_t *n_kept, double *wall_s, int64_t *tickets_out, char *err,
                      size_t err_len);
int blz_server_stats(blz_server server, int64_t *n_batches, int64_t *peak_inflight, size_t *used_bytes);
/* ---- ingest: delimited text to device columns ----
 * csv_parser::parse (io/data_parser/CSVParser.cpp:178-238) with the reader options
 * of init_default_csv_args (:26-47): `delimiter` and `line_terminator` bytes, quote
 * '"' (minimal quoting, no doubled quotes), blank lines skipped, no header, no NA
 * detection. The file's columns have dtypes[0 .. n_file_columns); the columns
 * column_indices[0 .. n_indices) (NULL: all, in order) are parsed into callee-
 * allocated out_columns (free with blz_free_column; no validity, null_count 0).
 * Parsed dtypes: INT8/16/32/64, FLOAT32/64, DATE32, DATE64, TIMESTAMP (ms);
 * others (strings) may be skipped but not requested (BLZ_ERR_UNSUPPORTED).
 * skip_rows drops that many rows first (csv_parser's constructor takes it and
 * leaves it unused, CSVParser.cpp:153-169). Integers: sign and digits up to the
 * first other byte, wrapping; reals: correctly rounded when the decimal has at most
 * 15-16 significant digits and a small exponent (binary64: m <= 2^53, |e| <= 22;
 * binary32: m <= 2^24, |e| <= 10), otherwise within a few ulp; dates
 * YYYY-MM-DD[(T| )hh:mm:ss[.fff]][Z]. A missing field reads as 0. *n_rows = rows. */
int blz_csv_read(const void *host_bytes, size_t n_bytes, char delimiter, char line_terminator, int skip_rows,
                 const int32_t *dtypes, int n_file_columns, const int32_t *column_indices, int n_indices,
                 gdf_column *out_columns, int64_t *n_rows, blz_stream_t stream, char *err, size_t err_len);
/* The same over bytes already in device memory (nothing is read past n_bytes). */
int blz_csv_parse_device(const void *device_bytes, size_t n_bytes, char delimiter, char line_terminator,
                         int skip_rows, const int32_t *dtypes, int n_file_columns, const int32_t *column_indices,
                         int n_indices, gdf_column *out_columns, int64_t *n_rows, blz_stream_t stream, char *err,
                         size_t err_len);
/* Measurement (no reference counterpart): as blz_csv_parse_device, and stage_ms[0..4)
 * receives the device time of its four kernels (row-start counts, their scan, row
 * starts, fields), from HIP events on `stream`. */
int blz_csv_parse_device_timed(const void *device_bytes, size_t n_bytes, char delimiter, char line_terminator,
                               int skip_rows, const int32_t *dtypes, int n_file_columns,
                               const int32_t *column_indices, int n_indices, gdf_column *out_columns,
                               int64_t *n_rows, float *stage_ms, blz_stream_t stream, char *err, size_t err_len);
/* data_loader::load_data (io/DataLoader.cpp:26-118): parses each file and
 * concatenates their columns in file order. */
int blz_csv_load_files(const char *const *paths, int n_paths, char delimiter, char line_terminator, int skip_rows,
                       const int32_t *dtypes, int n_file_columns, const int32_t *column_indices, int n_indices,
                       gdf_column *out_columns, int64_t *n_rows, blz_stream_t stream, char *err, size_t err_len);

/* Destroying a server frees what its queries still hold; results it parked in the
 * result repository must be freed first. */
int blz_server_destroy(blz_server server);

/* ---- asynchronous result hand-off: the result repository ----
 * result_set_repository (ResultSetRepository.h:29-63, .cpp), one per process.
 * Connections group query tokens; a token's frame is parked once (update_token)
 * and read by any number of readers (get_result, which blocks until it is parked
 * and counts a reference). Errors are the reference's std::runtime_error texts
 * ("Connection does not exist", "Token does not exist", "Result set does not
 * exist", "Column does not exist") with BLZ_ERR_INVALID. CUDA-IPC export of parked
 * columns (ral-message.cuh:23-46) is not provided: parked columns are device
 * pointers of this process. */

/* init_session (ResultSetRepository.cpp:115-133): a new random connection id. */
int blz_rs_init_session(uint64_t *connection, char *err, size_t err_len);
/* register_query (:36-44, add_token :22-34): token not ready, listed under the
 * connection (which is created when unknown, as add_token does). */
int blz_rs_register_query(uint64_t connection, uint64_t token, char *err, size_t err_len);
/* update_token (:45-112): parks n_columns columns (names may be NULL), the query's
 * duration and error text (NULL or "" for none), gives every column a column token
 * and wakes the token's readers. The repository takes the columns' device buffers
 * (allocated with hipMalloc) and frees them with the result; buffers shared by two
 * columns are freed once (the reference deduplicates the frame first, :59-60). An
 * unknown token leaves the buffers with the caller. */
int blz_rs_update_token(uint64_t token, const gdf_column *columns, const char *const *names, int n_columns,
                        double duration_ms, const char *error_message, char *err, size_t err_len);
/* get_result (:186-205): blocks until the token's frame is parked, counts one
 * reference and describes it: columns[i] (no copy: the repository's buffers),
 * column_tokens[i], the duration and the error text (empty on success). Either
 * array may be NULL; columns must hold the frame's columns. */
int blz_rs_get_result(uint64_t connection, uint64_t token, gdf_column *columns, uint64_t *column_tokens,
                      int max_columns, int *n_columns, double *duration_ms, char *error_message,
                      size_t error_len, char *err, size_t err_len);
/* Name of parked column `column` of a token (the projection's expression name). */
int blz_rs_column_name(uint64_t token, int column, char *name, size_t name_len);
/* try_free_result (:166-184): *found = whether the token exists; the last
 * reference (or none: a result never read) frees the frame, any other drops one.
 * Unknown connection: "Connection does not exist". */
int blz_rs_try_free_result(uint64_t connection, uint64_t token, int *found, char *err, size_t err_len);
/* remove_all_connection_tokens (:152-164): frees the connection's results that no
 * reader holds and forgets the connection. */
int blz_rs_remove_all_connection_tokens(uint64_t connection);
/* get_column (:207-226): the parked column with that column token. */
int blz_rs_get_column(uint64_t connection, uint64_t column_token, gdf_column *out, char *err, size_t err_len);
/* evaluate_query's asynchronous form (CalciteInterpreter.cpp:920-960) over the
 * serving loop: submits the query and, once it has run, parks its outputs (named
 * after the projection's expressions; "$i" for a bare filter) and its latency in
 * ms, or its error, under `token`, which must be registered. A query that cannot
 * be planned parks its error at once (the call returns BLZ_OK, as evaluate_query
 * does). The parked frame keeps the query's outputs and HBM budget until the
 * result is freed. */
int blz_server_submit_token(blz_server server, const blz_query *query, uint64_t token, char *err,
                            size_t err_len);

/* ---- benchmark inputs (no reference counterpart) ----
 * Fills an allocated column in place with seeded synthetic values: row i depends
 * only on (seed, i). dist: 0 integer uniform [a, b]; 1 real uniform [a, b);
 * 2 normal(mean a, sd b); 3 lognormal(a, b); 4 k / 100 for integer k uniform in
 * [a, b]; 5 a + i. null_rate > 0 also writes the validity bitmap. */
int blz_synth_column(gdf_column *column, uint64_t seed, int dist, double a, double b, double null_rate,
                     blz_stream_t stream, char *err, size_t err_len);

/* ---- memory / streams (HIP plumbing for hosts without a HIP toolchain) ------- */

int blz_device_malloc(void **ptr, size_t bytes);
int blz_device_free(void *ptr);
int blz_memcpy_htod(void *dst, const void *src, size_t bytes, blz_stream_t stream);
int blz_memcpy_dtoh(void *dst, const void *src, size_t bytes, blz_stream_t stream);
int blz_memset_device(void *dst, int value, size_t bytes, blz_stream_t stream);
/* Device-to-device copy by a plain streaming kernel (bytes and both pointers 16-byte
 * multiples); its hardware byte counters calibrate traffic measurements. */
int blz_device_copy(void *dst, const void *src, size_t bytes, blz_stream_t stream);
int blz_stream_create(blz_stream_t *stream);
int blz_stream_synchronize(blz_stream_t stream);
/* Frees the launch buffers (device memory and pinned readback words) the one-shot
 * calls keep per (device, stream) between calls (a stream's next call allocates
 * them again); waits for calls in flight. No reference counterpart. */
int blz_release_launch_buffers(void);
int blz_stream_destroy(blz_stream_t stream);
/* Frees data and validity of a callee-allocated column and zeroes the struct. */
int blz_free_column(gdf_column *column);
int blz_device_count(void);
const char *blz_version(void);

/* ---- planner only (host; no device work) -------------------------------------- */

/* Plans `expression` as evaluate_expression does (LogicalFilter.cpp:271-371) for a
 * frame of the given dtypes: fills up to `cap` ops and returns the op count in
 * *n_ops, the used frame columns (slot order) in in_cols / *n_in, the final
 * register in *final_pos. */
int blz_plan_expression(const int32_t *frame_dtypes, int n_frame, const char *expression,
                        int cap, column_index_type *left, column_index_type *right,
                        column_index_type *outputs, int32_t *binops, int32_t *unops,
                        gdf_scalar *lscalars, gdf_scalar *rscalars, int *n_ops,
                        int32_t *in_cols, int *n_in, column_index_type *final_pos,
                        char *err, size_t err_len);

/* add_expression_to_plan (src/LogicalFilter.h:26-42, LogicalFilter.cpp:44-267): the
 * ops of one expression as the reference appends them to a plan: registers
 * [0, num_inputs) are the inputs, num_inputs + expression_position the
 * expression's output, temporaries from num_inputs + num_outputs (lowest free
 * first); new_input_indices[k] (k < n_frame) is frame column k's input register or
 * -1. Writes the ops (at most cap) to the arrays and their count to *n_ops; the
 * caller appends them to its plan. */
int blz_add_expression_to_plan(const char *expression, column_index_type expression_position,
                               column_index_type num_outputs, column_index_type num_inputs,
                               const column_index_type *new_input_indices, int n_frame, int cap,
                               column_index_type *left, column_index_type *right, column_index_type *outputs,
                               int32_t *binops, int32_t *unops, gdf_scalar *lscalars, gdf_scalar *rscalars,
                               int *n_ops, char *err, size_t err_len);

/* parse_project_plan (src/CalciteInterpreter.h:63, CalciteInterpreter.cpp:123-309)
 * for a frame of the given dtypes: every evaluated expression of the LogicalProject
 * in one register plan (ops: at most cap), the frame columns it reads in register
 * order (in_cols[0..*n_in)), the evaluated outputs' registers (final_pos[0..*n_eval))
 * and, for each output i of the projection, out_kind[i] (0 evaluated, 1 a "$k"
 * column), out_index[i] (the evaluated output's number, or k) and out_dtype[i]. */
int blz_parse_project_plan(const int32_t *frame_dtypes, int n_frame, const char *query_part, int cap,
                           column_index_type *left, column_index_type *right, column_index_type *outputs,
                           int32_t *binops, int32_t *unops, gdf_scalar *lscalars, gdf_scalar *rscalars, int *n_ops,
                           int32_t *in_cols, int *n_in, column_index_type *final_pos, int *n_eval,
                           int32_t *out_kind, int32_t *out_index, int32_t *out_dtype, int max_out, int *n_out,
                           char *err, size_t err_len);

#ifdef __cplusplus
}
#endif
#endif /* BLZ_RAL_H_ */

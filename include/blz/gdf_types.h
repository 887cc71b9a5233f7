/*
 * gdf_types.h — column, scalar, dtype and operator types of the expression/filter path.
 *
 * These mirror the legacy cuDF `gdf_column` / `gdf_scalar` / `gdf_dtype` /
 * `gdf_binary_operator` types that the BlazingSQL RAL compiles against (they come
 * from the external, un-vendored cuDF headers; field names as used in
 * src/GDFColumn.cu:302-338 and src/CalciteExpressionParsing.cpp:337-417) and the
 * RAL's own `gdf_unary_operator` (src/gdf_wrapper/gdf_types.cuh:7-31).
 * Field names and order are kept so that reference call sites compile unchanged.
 * Enumerator ORDINALS of gdf_dtype / gdf_binary_operator are this library's own
 * (cuDF's are not in the reference tree); gdf_unary_operator follows
 * gdf_types.cuh:7-31 exactly.
 *
 * Plain C: usable from C, C++, HIP and through ctypes.
 */
#ifndef BLZ_GDF_TYPES_H_
#define BLZ_GDF_TYPES_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t gdf_size_type;  /* row counts (< 2^31), as in cuDF's legacy API   */
typedef int32_t gdf_index_type;
typedef uint8_t gdf_valid_type; /* validity bitmap byte, LSB-first                 */
typedef short column_index_type;/* src/Interpreter/interpreter_cpp.h:20           */

/* Register-file sentinels of a plan (src/Interpreter/interpreter_cpp.h:21-22 and
 * interpreter_ops.cuh:231-232): -1 = "no right operand" (unary op),
 * -2 = scalar operand, -3 = null scalar operand.                                 */
#define BLZ_UNARY_INDEX ((column_index_type)-1)
#define SCALAR_INDEX ((column_index_type)-2)
#define SCALAR_NULL_INDEX ((column_index_type)-3)

typedef enum {
  GDF_invalid = 0,
  GDF_INT8,
  GDF_INT16,
  GDF_INT32,
  GDF_INT64,
  GDF_FLOAT32,
  GDF_FLOAT64,
  GDF_BOOL8,
  GDF_DATE32,   /* days since epoch, int32          */
  GDF_DATE64,   /* milliseconds since epoch, int64  */
  GDF_TIMESTAMP,/* int64                            */
  GDF_CATEGORY,
  GDF_STRING,
  GDF_STRING_CATEGORY,
  N_GDF_TYPES
} gdf_dtype;

typedef enum {
  GDF_ADD = 0,
  GDF_SUB,
  GDF_MUL,
  GDF_DIV,
  GDF_TRUE_DIV,
  GDF_FLOOR_DIV,
  GDF_MOD,
  GDF_POW,
  GDF_EQUAL,
  GDF_NOT_EQUAL,
  GDF_LESS,
  GDF_GREATER,
  GDF_LESS_EQUAL,
  GDF_GREATER_EQUAL,
  GDF_BITWISE_AND,
  GDF_BITWISE_OR,
  GDF_BITWISE_XOR,
  GDF_LOGICAL_AND,
  GDF_LOGICAL_OR,
  GDF_COALESCE,
  GDF_INVALID_BINARY
} gdf_binary_operator;

/* src/gdf_wrapper/gdf_types.cuh:7-31 (same order). BLZ_CAST_FLOAT64 is an
 * extension placed after the reference's sentinel: (double)int64, round to
 * nearest even; the reference's CAST is a parser no-op.                          */
typedef enum {
  BLZ_FLOOR = 0,
  BLZ_CEIL,
  BLZ_SIN,
  BLZ_COS,
  BLZ_ASIN,
  BLZ_ACOS,
  BLZ_TAN,
  BLZ_COTAN,
  BLZ_ATAN,
  BLZ_ABS,
  BLZ_NOT,
  BLZ_LN,
  BLZ_LOG,
  BLZ_YEAR,
  BLZ_MONTH,
  BLZ_DAY,
  BLZ_HOUR,
  BLZ_MINUTE,
  BLZ_SECOND,
  BLZ_IS_NULL,
  BLZ_IS_NOT_NULL,
  BLZ_INVALID_UNARY,
  BLZ_CAST_FLOAT64
} gdf_unary_operator;

/* cuDF's legacy status code, as project_plan_params carries it
 * (src/CalciteInterpreter.h:31); this path sets GDF_SUCCESS or throws. */
typedef enum {
  GDF_SUCCESS = 0,
  GDF_CUDA_ERROR,
  GDF_UNSUPPORTED_DTYPE,
  GDF_COLUMN_SIZE_MISMATCH
} gdf_error;

typedef union {
  int8_t si08;
  int16_t si16;
  int32_t si32;
  int64_t si64;
  uint8_t ui08;
  uint16_t ui16;
  uint32_t ui32;
  uint64_t ui64;
  float fp32;
  double fp64;
  int32_t dt32;
  int64_t dt64;
  int64_t tmst;
} gdf_data;

typedef struct {
  gdf_data data;
  gdf_dtype dtype;
  int32_t is_valid; /* bool in cuDF; fixed 4-byte width here for a stable ABI */
} gdf_scalar;

typedef struct {
  void *category; /* NVCategory* in the reference; unused on this path */
} gdf_dtype_extra_info;

typedef struct gdf_column_ {
  void *data;              /* device pointer, size * width bytes                    */
  gdf_valid_type *valid;   /* device pointer or NULL; LSB-first, >= ceil(size/8) B  */
  gdf_size_type size;
  gdf_dtype dtype;
  gdf_size_type null_count;
  gdf_dtype_extra_info dtype_info;
  char *col_name;
} gdf_column;

/* Byte width per dtype (src/Traits/RuntimeTraits.cpp:8-59). 0 = not fixed width. */
static inline int blz_dtype_width(gdf_dtype t) {
  switch (t) {
    case GDF_INT8: case GDF_BOOL8: return 1;
    case GDF_INT16: return 2;
    case GDF_INT32: case GDF_FLOAT32: case GDF_DATE32: case GDF_CATEGORY:
    case GDF_STRING_CATEGORY: return 4;
    case GDF_INT64: case GDF_FLOAT64: case GDF_DATE64: case GDF_TIMESTAMP: return 8;
    default: return 0;
  }
}

/* Validity allocation used by gdf_column_cpp (src/GDFColumn.cu:163-173): the byte
 * count rounded up to a multiple of 64 bytes. Buffers of this size can be read and
 * written as whole 64-bit words.                                                  */
static inline size_t gdf_valid_allocation_size(gdf_size_type n) {
  size_t bytes = ((size_t)n + 7) / 8;
  return ((bytes + 63) / 64) * 64;
}

#ifdef __cplusplus
}
#endif
#endif /* BLZ_GDF_TYPES_H_ */

"""blz_ral — Python host binding of libblzral.so (the MI355X expression/filter path).

A thin ctypes layer over include/blz/blz_ral.h, mirroring the reference's C++
operator API (perform_operation, evaluate_expression, process_filter,
execute_project_plan, materialize_column, gdf_sequence) for hosts and tests
written in Python. Host data is numpy; device memory is owned by DeviceColumn.

The library is required: importing works without it, but every call raises
BlzError if libblzral.so is missing or no HIP device is present. There is no
CPU fallback.
"""
import ctypes as C
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libblzral.so")

# dtype ordinals of include/blz/gdf_types.h
GDF_invalid, GDF_INT8, GDF_INT16, GDF_INT32, GDF_INT64, GDF_FLOAT32, GDF_FLOAT64, \
    GDF_BOOL8, GDF_DATE32, GDF_DATE64, GDF_TIMESTAMP, GDF_CATEGORY, GDF_STRING, \
    GDF_STRING_CATEGORY = range(14)
(GDF_ADD, GDF_SUB, GDF_MUL, GDF_DIV, GDF_TRUE_DIV, GDF_FLOOR_DIV, GDF_MOD, GDF_POW,
 GDF_EQUAL, GDF_NOT_EQUAL, GDF_LESS, GDF_GREATER, GDF_LESS_EQUAL, GDF_GREATER_EQUAL,
 GDF_BITWISE_AND, GDF_BITWISE_OR, GDF_BITWISE_XOR, GDF_LOGICAL_AND, GDF_LOGICAL_OR,
 GDF_COALESCE, GDF_INVALID_BINARY) = range(21)
(BLZ_FLOOR, BLZ_CEIL, BLZ_SIN, BLZ_COS, BLZ_ASIN, BLZ_ACOS, BLZ_TAN, BLZ_COTAN, BLZ_ATAN,
 BLZ_ABS, BLZ_NOT, BLZ_LN, BLZ_LOG, BLZ_YEAR, BLZ_MONTH, BLZ_DAY, BLZ_HOUR, BLZ_MINUTE,
 BLZ_SECOND, BLZ_IS_NULL, BLZ_IS_NOT_NULL, BLZ_INVALID_UNARY, BLZ_CAST_FLOAT64) = range(23)
SCALAR_INDEX, SCALAR_NULL_INDEX = -2, -3
SCALAR_INTENDED, SCALAR_SHIPPED = 0, 1
BLZ_OK, BLZ_ERR_INVALID, BLZ_ERR_HIP, BLZ_ERR_OOM, BLZ_ERR_UNSUPPORTED = 0, -1, -2, -3, -4

NP_OF = {GDF_INT8: np.int8, GDF_INT16: np.int16, GDF_INT32: np.int32, GDF_INT64: np.int64,
         GDF_FLOAT32: np.float32, GDF_FLOAT64: np.float64, GDF_BOOL8: np.uint8,
         GDF_DATE32: np.int32, GDF_DATE64: np.int64, GDF_TIMESTAMP: np.int64,
         GDF_STRING_CATEGORY: np.int32}
GDF_OF = {np.dtype(np.int8): GDF_INT8, np.dtype(np.int16): GDF_INT16,
          np.dtype(np.int32): GDF_INT32, np.dtype(np.int64): GDF_INT64,
          np.dtype(np.float32): GDF_FLOAT32, np.dtype(np.float64): GDF_FLOAT64,
          np.dtype(np.uint8): GDF_BOOL8, np.dtype(np.bool_): GDF_BOOL8}


class BlzError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class GdfData(C.Union):
    _fields_ = [("si08", C.c_int8), ("si16", C.c_int16), ("si32", C.c_int32),
                ("si64", C.c_int64), ("ui08", C.c_uint8), ("ui16", C.c_uint16),
                ("ui32", C.c_uint32), ("ui64", C.c_uint64), ("fp32", C.c_float),
                ("fp64", C.c_double), ("dt32", C.c_int32), ("dt64", C.c_int64),
                ("tmst", C.c_int64)]


class GdfScalar(C.Structure):
    _fields_ = [("data", GdfData), ("dtype", C.c_int32), ("is_valid", C.c_int32)]


class GdfColumn(C.Structure):
    _fields_ = [("data", C.c_void_p), ("valid", C.c_void_p), ("size", C.c_int32),
                ("dtype", C.c_int32), ("null_count", C.c_int32), ("category", C.c_void_p),
                ("col_name", C.c_char_p)]


class BlzQuery(C.Structure):
    _fields_ = [("frame", C.POINTER(C.POINTER(GdfColumn))), ("n_frame", C.c_int),
                ("filter_query_part", C.c_char_p), ("project_query_part", C.c_char_p),
                ("scalar_mode", C.c_int)]


_lib = None


def lib():
    """Loads libblzral.so (raises BlzError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BlzError(BLZ_ERR_HIP, f"{LIB_PATH} is missing: build it first "
                                        "(python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        P, I, I16P, I32P, ERR, SZ = C.c_void_p, C.c_int, C.POINTER(C.c_int16), \
            C.POINTER(C.c_int32), C.c_char_p, C.c_size_t
        colpp = C.POINTER(C.POINTER(GdfColumn))
        colp = C.POINTER(GdfColumn)
        scp = C.POINTER(GdfScalar)
        sig = {
            "blz_perform_operation": [colpp, I, colpp, I, I16P, I16P, I16P, I, I16P, I32P, I32P,
                                      scp, scp, I, I32P, P, ERR, SZ],
            "blz_perform_operation_noalloc": [colpp, I, colpp, I, I16P, I16P, I16P, I, I16P, I32P, I32P,
                                              scp, scp, I, I32P, P, ERR, SZ],
            "blz_evaluate_expression": [colpp, I, C.c_char_p, colp, I, I32P, P, ERR, SZ],
            "blz_process_filter": [colpp, I, C.c_char_p, colp, C.POINTER(C.c_int64), I, P, ERR, SZ],
            "blz_execute_project": [colpp, I, C.c_char_p, colp, I32P, I, C.POINTER(C.c_int), I, P,
                                    ERR, SZ],
            "blz_filter_project": [colpp, I, C.c_char_p, C.c_char_p, colp, I, C.POINTER(C.c_int),
                                   C.POINTER(C.c_int64), I, P, ERR, SZ],
            "blz_process_aggregate": [colpp, I, C.c_char_p, colp, I, C.POINTER(C.c_int), I, P, ERR, SZ],
            "blz_materialize_column": [colp, colp, colp, P, ERR, SZ],
            "blz_materialize_columns": [colpp, colpp, I, colp, P, ERR, SZ],
            "blz_gdf_sequence": [P, SZ, C.c_int32, C.c_int32, P, ERR, SZ],
            "blz_apply_boolean_mask": [colp, colp, colp, P, ERR, SZ],
            "blz_update_null_count": [colp, P, ERR, SZ],
            "blz_program_create": [C.POINTER(P), colpp, I, C.c_char_p, C.c_char_p, colp, I,
                                   C.POINTER(C.c_int), I, I, I, I, I, ERR, SZ],
            "blz_program_tile": [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
            "blz_program_predicate": [P] + [C.POINTER(C.c_int)] * 3,
            "blz_program_launch": [P, P, ERR, SZ],
            "blz_program_result": [P, P, C.POINTER(C.c_int64), I32P, ERR, SZ],
            "blz_program_time": [P, P, I, C.POINTER(C.c_double), ERR, SZ],
            "blz_program_time_stages": [P, P, I, C.POINTER(C.c_double), C.POINTER(C.c_int), ERR, SZ],
            "blz_program_info": [P] + [C.POINTER(C.c_int)] * 6,
            "blz_program_destroy": [P],
            "blz_program_source": [colpp, I, C.c_char_p, C.c_char_p, I, C.c_char_p, SZ, C.POINTER(SZ), ERR, SZ],
            "blz_jit_compile": [C.c_char_p, ERR, SZ],
            "blz_jit_compile_many": [C.POINTER(C.c_char_p), C.c_int, ERR, SZ],
            "blz_jit_counters": [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(C.c_double)],
            "blz_batch_create": [C.POINTER(P), C.POINTER(BlzQuery), I, I, I, I, ERR, SZ],
            "blz_batch_launch": [P, P, ERR, SZ],
            "blz_batch_result": [P, P, I, C.POINTER(C.c_int64), I32P, colp, I, C.POINTER(C.c_int), ERR, SZ],
            "blz_batch_time": [P, P, I, C.POINTER(C.c_double), ERR, SZ],
            "blz_batch_info": [P] + [C.POINTER(C.c_int)] * 5,
            "blz_batch_query_kernel": [P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)],
            "blz_batch_destroy": [P],
            "blz_server_create": [C.POINTER(P), SZ, I, I, I, ERR, SZ],
            "blz_server_submit": [P, C.POINTER(BlzQuery), C.POINTER(C.c_int64), ERR, SZ],
            "blz_server_wait": [P, C.c_int64, C.POINTER(C.c_int64), I32P, colp, I, C.POINTER(C.c_int),
                                C.POINTER(C.c_double), ERR, SZ],
            "blz_server_release": [P, C.c_int64],
            "blz_server_replay": [P, C.POINTER(BlzQuery), C.POINTER(C.c_double), I, C.POINTER(C.c_double),
                                  C.POINTER(C.c_int64), C.POINTER(C.c_double), C.POINTER(C.c_int64), ERR, SZ],
            "blz_server_stats": [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(SZ)],
            "blz_server_destroy": [P],
            "blz_server_submit_token": [P, C.POINTER(BlzQuery), C.c_uint64, ERR, SZ],
            "blz_csv_read": [P, SZ, C.c_char, C.c_char, I, I32P, I, I32P, I, colp, C.POINTER(C.c_int64), P, ERR,
                             SZ],
            "blz_csv_parse_device": [P, SZ, C.c_char, C.c_char, I, I32P, I, I32P, I, colp,
                                     C.POINTER(C.c_int64), P, ERR, SZ],
            "blz_csv_parse_device_timed": [P, SZ, C.c_char, C.c_char, I, I32P, I, I32P, I, colp,
                                           C.POINTER(C.c_int64), C.POINTER(C.c_float), P, ERR, SZ],
            "blz_csv_load_files": [C.POINTER(C.c_char_p), I, C.c_char, C.c_char, I, I32P, I, I32P, I, colp,
                                   C.POINTER(C.c_int64), P, ERR, SZ],
            "blz_rs_init_session": [C.POINTER(C.c_uint64), ERR, SZ],
            "blz_rs_register_query": [C.c_uint64, C.c_uint64, ERR, SZ],
            "blz_rs_update_token": [C.c_uint64, colp, C.POINTER(C.c_char_p), I, C.c_double, C.c_char_p, ERR, SZ],
            "blz_rs_get_result": [C.c_uint64, C.c_uint64, colp, C.POINTER(C.c_uint64), I, C.POINTER(C.c_int),
                                  C.POINTER(C.c_double), C.c_char_p, SZ, ERR, SZ],
            "blz_rs_column_name": [C.c_uint64, I, C.c_char_p, SZ],
            "blz_rs_try_free_result": [C.c_uint64, C.c_uint64, C.POINTER(C.c_int), ERR, SZ],
            "blz_rs_remove_all_connection_tokens": [C.c_uint64],
            "blz_rs_get_column": [C.c_uint64, C.c_uint64, colp, ERR, SZ],
            "blz_synth_column": [colp, C.c_uint64, I, C.c_double, C.c_double, C.c_double, P, ERR, SZ],
            "blz_device_malloc": [C.POINTER(P), SZ],
            "blz_device_free": [P],
            "blz_memcpy_htod": [P, P, SZ, P],
            "blz_memcpy_dtoh": [P, P, SZ, P],
            "blz_memset_device": [P, I, SZ, P],
            "blz_device_copy": [P, P, SZ, P],
            "blz_stream_create": [C.POINTER(P)],
            "blz_stream_synchronize": [P],
            "blz_stream_destroy": [P],
            "blz_release_launch_buffers": [],
            "blz_free_column": [colp],
            "blz_device_count": [],
            "blz_plan_expression": [I32P, I, C.c_char_p, I, I16P, I16P, I16P, I32P, I32P, scp, scp,
                                    C.POINTER(C.c_int), I32P, C.POINTER(C.c_int), I16P, ERR, SZ],
            "blz_add_expression_to_plan": [C.c_char_p, C.c_int16, C.c_int16, C.c_int16, I16P, I, I, I16P, I16P,
                                           I16P, I32P, I32P, scp, scp, C.POINTER(C.c_int), ERR, SZ],
            "blz_parse_project_plan": [I32P, I, C.c_char_p, I, I16P, I16P, I16P, I32P, I32P, scp, scp,
                                       C.POINTER(C.c_int), I32P, C.POINTER(C.c_int), I16P, C.POINTER(C.c_int),
                                       I32P, I32P, I32P, I, C.POINTER(C.c_int), ERR, SZ],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.blz_version.restype = C.c_char_p
        L.blz_version.argtypes = []
        _lib = L
    return _lib


def _call(fn, *args):
    err = C.create_string_buffer(1024)
    rc = fn(*args, err, len(err))
    if rc != BLZ_OK:
        raise BlzError(rc, err.value.decode(errors="replace"))
    return rc


def _check(rc, what):
    if rc != BLZ_OK:
        raise BlzError(rc, what)


def device_count():
    return lib().blz_device_count()


# ---- validity bitmaps ---------------------------------------------------------------

def valid_alloc_size(n):
    """gdf_valid_allocation_size (GDFColumn.cu:163-173): bytes rounded to 64."""
    return ((n + 7) // 8 + 63) // 64 * 64


def pack_valid(mask: np.ndarray) -> np.ndarray:
    """bool[n] -> LSB-first bitmap padded to valid_alloc_size(n) bytes."""
    n = len(mask)
    out = np.zeros(max(valid_alloc_size(n), 8), dtype=np.uint8)
    if n:
        packed = np.packbits(np.asarray(mask, dtype=bool), bitorder="little")
        out[:len(packed)] = packed
    return out


def unpack_valid(bits: np.ndarray, n: int) -> np.ndarray:
    return np.unpackbits(np.asarray(bits, dtype=np.uint8), bitorder="little")[:n].astype(bool)


# ---- device columns -------------------------------------------------------------------

class DeviceBuffer:
    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p(0)
        if self.nbytes:
            _check(lib().blz_device_malloc(C.byref(self.ptr), self.nbytes), "device malloc failed")

    @property
    def value(self):
        return self.ptr.value

    def upload(self, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        if arr.nbytes:
            _check(lib().blz_memcpy_htod(self.ptr, arr.ctypes.data, arr.nbytes, None), "htod failed")

    def download(self, nbytes, dtype=np.uint8):
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        if nbytes:
            _check(lib().blz_memcpy_dtoh(out.ctypes.data, self.ptr, nbytes, None), "dtoh failed")
        return out

    def free(self):
        if self.ptr.value:
            lib().blz_device_free(self.ptr)
            self.ptr = C.c_void_p(0)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceColumn:
    """A gdf_column whose data (and optional validity) live in device memory."""

    def __init__(self, dtype: int, size: int, with_valid: bool = False, null_count: int = 0,
                 _adopt: Optional[GdfColumn] = None):
        self.dtype = dtype
        self.size = size
        self._owned = _adopt is None
        self.col = GdfColumn()
        if _adopt is not None:
            self.col = _adopt
            self._data = self._valid = None
            return
        w = np.dtype(NP_OF[dtype]).itemsize
        self._data = DeviceBuffer(max(size * w, 0) + 64)
        self._valid = DeviceBuffer(max(valid_alloc_size(size), 64)) if with_valid else None
        self.col.data = self._data.value
        self.col.valid = self._valid.value if self._valid else None
        self.col.size = size
        self.col.dtype = dtype
        self.col.null_count = null_count

    @classmethod
    def from_numpy(cls, arr: np.ndarray, dtype: Optional[int] = None, valid=None,
                   null_count: Optional[int] = None):
        arr = np.ascontiguousarray(arr)
        if dtype is None:
            dtype = GDF_OF[arr.dtype]
        arr = arr.astype(NP_OF[dtype], copy=False)
        c = cls(dtype, len(arr), with_valid=valid is not None)
        c._data.upload(arr)
        if valid is not None:
            if isinstance(valid, np.ndarray) and valid.dtype == np.uint8 and len(valid) != len(arr):
                bits = valid  # already a bitmap
                vmask = unpack_valid(bits, len(arr))
            else:
                vmask = np.asarray(valid, dtype=bool)
                bits = pack_valid(vmask)
            c._valid.upload(bits)
            c.col.null_count = int((~vmask).sum()) if null_count is None else null_count
        return c

    @classmethod
    def adopt(cls, col: GdfColumn, owned: bool):
        c = cls(col.dtype, col.size, _adopt=col)
        c._owned = owned
        return c

    def data(self) -> np.ndarray:
        t = NP_OF[self.col.dtype]
        n = self.col.size
        out = np.empty(n, dtype=t)
        if n:
            _check(lib().blz_memcpy_dtoh(out.ctypes.data, self.col.data, out.nbytes, None), "dtoh")
        return out

    def valid_bits(self) -> Optional[np.ndarray]:
        if not self.col.valid:
            return None
        nb = (self.col.size + 7) // 8
        out = np.empty(nb, dtype=np.uint8)
        if nb:
            _check(lib().blz_memcpy_dtoh(out.ctypes.data, self.col.valid, nb, None), "dtoh")
        return out

    def valid(self) -> Optional[np.ndarray]:
        b = self.valid_bits()
        return None if b is None else unpack_valid(b, self.col.size)

    @property
    def null_count(self):
        return self.col.null_count

    def free(self):
        if not self._owned:
            return
        if self._data is None:
            lib().blz_free_column(C.byref(self.col))
        else:
            self._data.free()
            if self._valid:
                self._valid.free()
        self._owned = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _colpp(cols: Sequence[DeviceColumn]):
    arr = (C.POINTER(GdfColumn) * max(len(cols), 1))()
    for i, c in enumerate(cols):
        arr[i] = C.pointer(c.col)
    return arr


def make_scalar(dtype: int, value=0, is_valid=True) -> GdfScalar:
    s = GdfScalar()
    s.dtype = dtype
    s.is_valid = 1 if is_valid else 0
    field = {GDF_INT8: "si08", GDF_INT16: "si16", GDF_INT32: "si32", GDF_INT64: "si64",
             GDF_FLOAT32: "fp32", GDF_FLOAT64: "fp64", GDF_DATE32: "dt32", GDF_DATE64: "dt64",
             GDF_TIMESTAMP: "tmst", GDF_STRING_CATEGORY: "si32", GDF_BOOL8: "ui08"}.get(dtype)
    if field is not None:
        setattr(s.data, field, value)
    return s


# ---- the reference operator API ----------------------------------------------------

def perform_operation(outputs: Sequence[DeviceColumn], inputs: Sequence[DeviceColumn],
                      left, right, outs, final_positions, operators, unary_operators,
                      left_scalars: Sequence[GdfScalar], right_scalars: Sequence[GdfScalar],
                      scalar_mode=SCALAR_INTENDED, stream=None) -> List[int]:
    """perform_operation (interpreter_cpp.h:25-37). Returns per-output null counts."""
    n = len(left)
    I16 = C.c_int16 * max(n, 1)
    I32 = C.c_int32 * max(n, 1)
    SC = GdfScalar * max(n, 1)
    fp = (C.c_int16 * max(len(final_positions), 1))(*final_positions)
    nulls = (C.c_int32 * max(len(outputs), 1))()
    _call(lib().blz_perform_operation, _colpp(outputs), len(outputs), _colpp(inputs), len(inputs),
          I16(*left), I16(*right), I16(*outs), n, fp, I32(*operators), I32(*unary_operators),
          SC(*left_scalars), SC(*right_scalars), scalar_mode, nulls, stream)
    return list(nulls)[:len(outputs)]


def perform_operation_noalloc(outputs: Sequence[DeviceColumn], inputs: Sequence[DeviceColumn],
                              left, right, outs, final_positions, operators, unary_operators,
                              left_scalars: Sequence[GdfScalar], right_scalars: Sequence[GdfScalar],
                              scalar_mode=SCALAR_INTENDED, stream=None) -> List[int]:
    """perform_operation without device allocation (the library's per-device scratch
    holds the launch buffers). Returns per-output null counts."""
    n = len(left)
    I16 = C.c_int16 * max(n, 1)
    I32 = C.c_int32 * max(n, 1)
    SC = GdfScalar * max(n, 1)
    fp = (C.c_int16 * max(len(final_positions), 1))(*final_positions)
    nulls = (C.c_int32 * max(len(outputs), 1))()
    _call(lib().blz_perform_operation_noalloc, _colpp(outputs), len(outputs), _colpp(inputs), len(inputs),
          I16(*left), I16(*right), I16(*outs), n, fp, I32(*operators), I32(*unary_operators),
          SC(*left_scalars), SC(*right_scalars), scalar_mode, nulls, stream)
    return list(nulls)[:len(outputs)]


def evaluate_expression(frame: Sequence[DeviceColumn], expression: str, output: DeviceColumn,
                        scalar_mode=SCALAR_INTENDED, stream=None) -> int:
    """evaluate_expression (LogicalFilter.h:20-23). Returns the output's null count."""
    nc = C.c_int32(0)
    _call(lib().blz_evaluate_expression, _colpp(frame), len(frame), expression.encode(),
          C.byref(output.col), scalar_mode, C.byref(nc), stream)
    return nc.value


def process_filter(frame: Sequence[DeviceColumn], query_part: str,
                   scalar_mode=SCALAR_INTENDED, stream=None) -> List[DeviceColumn]:
    """process_filter (CalciteInterpreter.cpp:509-573): every column, rows kept in order."""
    out = (GdfColumn * len(frame))()
    kept = C.c_int64(0)
    _call(lib().blz_process_filter, _colpp(frame), len(frame), query_part.encode(), out,
          C.byref(kept), scalar_mode, stream)
    return [DeviceColumn.adopt(out[i], owned=True) for i in range(len(frame))]


def execute_project(frame: Sequence[DeviceColumn], query_part: str,
                    scalar_mode=SCALAR_INTENDED, stream=None, max_out=64) -> List[DeviceColumn]:
    """execute_project_plan (CalciteInterpreter.h:61-63). Aliases share the input's memory."""
    out = (GdfColumn * max_out)()
    alias = (C.c_int32 * max_out)()
    n = C.c_int(0)
    _call(lib().blz_execute_project, _colpp(frame), len(frame), query_part.encode(), out, alias,
          max_out, C.byref(n), scalar_mode, stream)
    res = []
    for i in range(n.value):
        c = DeviceColumn.adopt(out[i], owned=not alias[i])
        if alias[i]:  # shares a frame column's buffers: keep that column alive
            c._base = next((f for f in frame if f.col.data == out[i].data), None)
        res.append(c)
    return res


def filter_project(frame: Sequence[DeviceColumn], filter_query_part: str,
                   project_query_part: Optional[str], scalar_mode=SCALAR_INTENDED, stream=None,
                   max_out=64):
    """LogicalProject over LogicalFilter in one pass. Returns (columns, rows kept)."""
    out = (GdfColumn * max_out)()
    n = C.c_int(0)
    kept = C.c_int64(0)
    _call(lib().blz_filter_project, _colpp(frame), len(frame), filter_query_part.encode(),
          project_query_part.encode() if project_query_part else None, out, max_out, C.byref(n),
          C.byref(kept), scalar_mode, stream)
    return [DeviceColumn.adopt(out[i], owned=True) for i in range(n.value)], kept.value


def process_aggregate(frame: Sequence[DeviceColumn], query_part: str, scalar_mode=SCALAR_INTENDED,
                      stream=None, max_out=64):
    """process_aggregate without GROUP BY (GroupBy.cpp:531-588): one-row columns."""
    out = (GdfColumn * max_out)()
    n = C.c_int(0)
    _call(lib().blz_process_aggregate, _colpp(frame), len(frame), query_part.encode(), out, max_out,
          C.byref(n), scalar_mode, stream)
    return [DeviceColumn.adopt(out[i], owned=True) for i in range(n.value)]


def materialize_column(inp: DeviceColumn, out: DeviceColumn, row_indices: DeviceColumn, stream=None):
    """materialize_column (ColumnManipulation.cuh:16-18). Sets out.null_count."""
    _call(lib().blz_materialize_column, C.byref(inp.col), C.byref(out.col), C.byref(row_indices.col),
          stream)
    return out.col.null_count


def materialize_columns(inputs: Sequence[DeviceColumn], outputs: Sequence[DeviceColumn],
                        row_indices: DeviceColumn, stream=None):
    """materialize_column over several columns sharing the indices, one blocking call."""
    _call(lib().blz_materialize_columns, _colpp(inputs), _colpp(outputs), len(inputs), C.byref(row_indices.col),
          stream)
    return [o.col.null_count for o in outputs]


def gdf_sequence(col: DeviceColumn, init_val=0, step=1, stream=None):
    """gdf_sequence (Utils.cuh:187-190)."""
    _call(lib().blz_gdf_sequence, C.c_void_p(col.col.data), col.col.size, init_val, step, stream)


def apply_boolean_mask(inp: DeviceColumn, stencil: DeviceColumn, stream=None) -> DeviceColumn:
    out = GdfColumn()
    _call(lib().blz_apply_boolean_mask, C.byref(inp.col), C.byref(stencil.col), C.byref(out), stream)
    return DeviceColumn.adopt(out, owned=True)


def update_null_count(col: DeviceColumn, stream=None) -> int:
    _call(lib().blz_update_null_count, C.byref(col.col), stream)
    return col.col.null_count


class Program:
    """A planned and compiled filter+project bound to device buffers (launch many)."""

    def __init__(self, frame: Sequence[DeviceColumn], filter_query_part: Optional[str],
                 project_query_part: Optional[str], scalar_mode=SCALAR_INTENDED, s=0, r=0, t=0,
                 kernel=0, max_out=64):
        self.frame = list(frame)
        self.h = C.c_void_p(0)
        out = (GdfColumn * max_out)()
        n_out = C.c_int(0)
        _call(lib().blz_program_create, C.byref(self.h), _colpp(frame), len(frame),
              filter_query_part.encode() if filter_query_part else None,
              project_query_part.encode() if project_query_part else None, out, max_out,
              C.byref(n_out), scalar_mode, kernel, s, r, t)
        info = [C.c_int(0) for _ in range(6)]
        lib().blz_program_info(self.h, *[C.byref(x) for x in info])
        self.s, self.r, self.n_ops, self.n_vops, self.n_slots, self.grid = [x.value for x in info]
        tr, lb, kk = C.c_int(0), C.c_int(0), C.c_int(0)
        lib().blz_program_tile(self.h, C.byref(tr), C.byref(lb), C.byref(kk))
        self.tile_rows, self.lds_bytes, self.kernel = tr.value, lb.value, kk.value
        pd = [C.c_int(0) for _ in range(3)]
        lib().blz_program_predicate(self.h, *[C.byref(x) for x in pd])
        # register-load predicate: (terms, 8-byte term mask), or None
        self.direct_pred = (pd[1].value, pd[2].value) if pd[0].value else None
        self.outputs = [DeviceColumn.adopt(out[i], owned=False) for i in range(n_out.value)]

    def launch(self, stream=None):
        _call(lib().blz_program_launch, self.h, stream)

    def result(self, stream=None):
        kept = C.c_int64(0)
        nulls = (C.c_int32 * max(len(self.outputs), 1))()
        _call(lib().blz_program_result, self.h, stream, C.byref(kept), nulls)
        for c in self.outputs:
            c.col.size = kept.value
        return kept.value, list(nulls)[:len(self.outputs)]

    def time(self, iters=10, stream=None):
        ms = C.c_double(0)
        _call(lib().blz_program_time, self.h, stream, iters, C.byref(ms))
        return ms.value

    def time_stages(self, iters=10, stream=None):
        """Mean milliseconds of each kernel launch of the program (HIP events on the
        launch stream): [predicate, slot scan, gather] for a tile-VM filter."""
        st = (C.c_double * 3)()
        ns = C.c_int(0)
        _call(lib().blz_program_time_stages, self.h, stream, iters, st, C.byref(ns))
        return [st[i] for i in range(ns.value)]

    def close(self):
        if self.h.value:
            lib().blz_program_destroy(self.h)
            self.h = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def program_source(frame_dtypes: Sequence[int], n_rows: int, filter_query_part=None, project_query_part=None,
                   has_valid=None, null_counts=None, scalar_mode=SCALAR_INTENDED) -> str:
    """The compiled tier's kernel source for a frame of this shape (no device: the
    column descriptors carry stand-in pointers that are never read)."""
    n = len(frame_dtypes)
    has_valid = has_valid or [False] * n
    null_counts = null_counts or [0] * n
    cols = []
    for t, v, nc in zip(frame_dtypes, has_valid, null_counts):
        g = GdfColumn()
        g.data = 4096
        g.valid = 4096 if v else None
        g.size = n_rows
        g.dtype = t
        g.null_count = nc
        cols.append(g)
    arr = (C.POINTER(GdfColumn) * n)(*[C.pointer(g) for g in cols])
    length = C.c_size_t(0)
    f = filter_query_part.encode() if filter_query_part else None
    q = project_query_part.encode() if project_query_part else None
    _call(lib().blz_program_source, arr, n, f, q, scalar_mode, None, 0, C.byref(length))
    buf = C.create_string_buffer(length.value + 1)
    _call(lib().blz_program_source, arr, n, f, q, scalar_mode, buf, len(buf), C.byref(length))
    return buf.value.decode()


def jit_compile(src: str):
    """Compiles a generated kernel source for gfx950 (no device needed); raises
    BlzError with the compiler's log when it does not compile."""
    err = C.create_string_buffer(8192)
    rc = lib().blz_jit_compile(src.encode(), err, len(err))
    if rc != BLZ_OK:
        raise BlzError(rc, err.value.decode(errors="replace"))


def jit_compile_many(srcs):
    """Compiles generated kernel sources not compiled yet as one program."""
    arr = (C.c_char_p * len(srcs))(*[x.encode() for x in srcs])
    _call(lib().blz_jit_compile_many, arr, len(srcs))


def jit_counters():
    """(compilations, cache hits, milliseconds compiling) of this process."""
    c, h, ms = C.c_longlong(0), C.c_longlong(0), C.c_double(0)
    lib().blz_jit_counters(C.byref(c), C.byref(h), C.byref(ms))
    return c.value, h.value, ms.value


def plan_expression(frame_dtypes: Sequence[int], expression: str, cap=256):
    """The planner alone (host): evaluate_expression's plan for a frame of dtypes."""
    I16 = C.c_int16 * cap
    I32 = C.c_int32 * cap
    SC = GdfScalar * cap
    left, right, outs, final = I16(), I16(), I16(), (C.c_int16 * 1)()
    bops, uops, in_cols = I32(), I32(), I32()
    ls, rs = SC(), SC()
    n_ops, n_in = C.c_int(0), C.c_int(0)
    dt = (C.c_int32 * max(len(frame_dtypes), 1))(*frame_dtypes)
    _call(lib().blz_plan_expression, dt, len(frame_dtypes), expression.encode(), cap, left, right,
          outs, bops, uops, ls, rs, C.byref(n_ops), in_cols, C.byref(n_in), final)
    k = n_ops.value
    return dict(left=list(left)[:k], right=list(right)[:k], outputs=list(outs)[:k],
                binops=list(bops)[:k], unops=list(uops)[:k], lscalars=[ls[i] for i in range(k)],
                rscalars=[rs[i] for i in range(k)], input_columns=list(in_cols)[:n_in.value],
                final_position=final[0])


def add_expression_to_plan(expression: str, expression_position: int, num_outputs: int, num_inputs: int,
                           new_input_indices: Sequence[int], cap=256):
    """add_expression_to_plan (LogicalFilter.h:26-42): the ops one expression appends
    to a plan (host only)."""
    I16 = C.c_int16 * cap
    I32 = C.c_int32 * cap
    SC = GdfScalar * cap
    left, right, outs, bops, uops, ls, rs = I16(), I16(), I16(), I32(), I32(), SC(), SC()
    idx = (C.c_int16 * max(len(new_input_indices), 1))(*new_input_indices)
    n = C.c_int(0)
    _call(lib().blz_add_expression_to_plan, expression.encode(), expression_position, num_outputs, num_inputs,
          idx, len(new_input_indices), cap, left, right, outs, bops, uops, ls, rs, C.byref(n))
    k = n.value
    return dict(left=list(left)[:k], right=list(right)[:k], outputs=list(outs)[:k], binops=list(bops)[:k],
                unops=list(uops)[:k], lscalars=[ls[i] for i in range(k)], rscalars=[rs[i] for i in range(k)])


def parse_project_plan(frame_dtypes: Sequence[int], query_part: str, cap=256, max_out=64):
    """parse_project_plan (CalciteInterpreter.h:63) for a frame of dtypes (host only):
    the plan, the frame columns it reads in register order, the evaluated outputs'
    registers and each output's (kind, index, dtype), kind 0 evaluated, 1 "$k"."""
    I16 = C.c_int16 * cap
    I32 = C.c_int32 * cap
    SC = GdfScalar * cap
    left, right, outs, bops, uops, ls, rs = I16(), I16(), I16(), I32(), I32(), SC(), SC()
    in_cols, fin = I32(), (C.c_int16 * max_out)()
    kind, index, odt = (C.c_int32 * max_out)(), (C.c_int32 * max_out)(), (C.c_int32 * max_out)()
    n_ops, n_in, n_eval, n_out = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0)
    dt = (C.c_int32 * max(len(frame_dtypes), 1))(*frame_dtypes)
    _call(lib().blz_parse_project_plan, dt, len(frame_dtypes), query_part.encode(), cap, left, right, outs, bops,
          uops, ls, rs, C.byref(n_ops), in_cols, C.byref(n_in), fin, C.byref(n_eval), kind, index, odt, max_out,
          C.byref(n_out))
    k = n_ops.value
    return dict(left=list(left)[:k], right=list(right)[:k], outputs=list(outs)[:k], binops=list(bops)[:k],
                unops=list(uops)[:k], lscalars=[ls[i] for i in range(k)], rscalars=[rs[i] for i in range(k)],
                input_columns=list(in_cols)[:n_in.value], final_positions=list(fin)[:n_eval.value],
                outputs_kind=list(kind)[:n_out.value], outputs_index=list(index)[:n_out.value],
                outputs_dtype=list(odt)[:n_out.value])


class Batch:
    """Many independent queries (frame, filter, projection) executed by one persistent
    launch per kernel kind; r=0 lets the autotuner choose rows per lane and tile rows."""

    def __init__(self, queries, r=0, t=0, use_graph=True, scalar_mode=SCALAR_INTENDED, max_out=64):
        self.queries = [(list(f), fq, pq) for f, fq, pq in queries]
        self.max_out = max_out
        self._keep = []
        arr = (BlzQuery * len(self.queries))()
        for i, (frame, fq, pq) in enumerate(self.queries):
            fpp = _colpp(frame)
            self._keep.append(fpp)
            arr[i].frame = C.cast(fpp, C.POINTER(C.POINTER(GdfColumn)))
            arr[i].n_frame = len(frame)
            arr[i].filter_query_part = fq.encode() if fq else None
            arr[i].project_query_part = pq.encode() if pq else None
            arr[i].scalar_mode = scalar_mode
        self.h = C.c_void_p(0)
        _call(lib().blz_batch_create, C.byref(self.h), arr, len(self.queries), r, t, 1 if use_graph else 0)
        info = [C.c_int(0) for _ in range(5)]
        lib().blz_batch_info(self.h, *[C.byref(x) for x in info])
        self.r, self.t, self.n_queries, self.graph, self.tuned = [x.value for x in info]
        self.kernels = [self.query_kernel(q) for q in range(self.n_queries)]
        self.n_compiled = sum(1 for k, _ in self.kernels if k == 5)

    def query_kernel(self, q):
        """(kernel kind running query q: 4 tile VM, 5 compiled tier, ...; True when
        it runs inside a shared launch)."""
        k, sh = C.c_int(0), C.c_int(0)
        _check(lib().blz_batch_query_kernel(self.h, q, C.byref(k), C.byref(sh)), "blz_batch_query_kernel")
        return k.value, bool(sh.value)

    def launch(self, stream=None):
        _call(lib().blz_batch_launch, self.h, stream)

    def result(self, q, stream=None):
        """(rows kept, [DeviceColumn views of the outputs], [null counts])."""
        kept = C.c_int64(0)
        nulls = (C.c_int32 * self.max_out)()
        out = (GdfColumn * self.max_out)()
        n_out = C.c_int(0)
        _call(lib().blz_batch_result, self.h, stream, q, C.byref(kept), nulls, out, self.max_out,
              C.byref(n_out))
        cols = [DeviceColumn.adopt(out[i], owned=False) for i in range(n_out.value)]
        return kept.value, cols, [nulls[i] for i in range(n_out.value)]

    def time(self, iters=5, stream=None):
        ms = C.c_double(0)
        _call(lib().blz_batch_time, self.h, stream, iters, C.byref(ms))
        return ms.value

    def close(self):
        if self.h.value:
            lib().blz_batch_destroy(self.h)
            self.h = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _queries(queries, scalar_mode, keep):
    arr = (BlzQuery * len(queries))()
    for i, (frame, fq, pq) in enumerate(queries):
        fpp = _colpp(frame)
        keep.append(fpp)
        arr[i].frame = C.cast(fpp, C.POINTER(C.POINTER(GdfColumn)))
        arr[i].n_frame = len(frame)
        arr[i].filter_query_part = fq.encode() if fq else None
        arr[i].project_query_part = pq.encode() if pq else None
        arr[i].scalar_mode = scalar_mode
    return arr


class Server:
    """Request loop: HBM-budget admission, continuous batching over n_streams."""

    def __init__(self, hbm_budget_bytes, max_batch=64, n_streams=4, r=4):
        self.h = C.c_void_p(0)
        self._keep = {}
        _call(lib().blz_server_create, C.byref(self.h), int(hbm_budget_bytes), max_batch, n_streams, r)

    def submit(self, frame, filter_query_part, project_query_part, scalar_mode=SCALAR_INTENDED):
        keep = []
        arr = _queries([(frame, filter_query_part, project_query_part)], scalar_mode, keep)
        tk = C.c_int64(0)
        _call(lib().blz_server_submit, self.h, arr, C.byref(tk))
        self._keep[tk.value] = (keep, list(frame))
        return tk.value

    def wait(self, ticket, max_out=64):
        """(rows kept, [DeviceColumn views], [null counts], latency ms)."""
        kept = C.c_int64(0)
        nulls = (C.c_int32 * max_out)()
        out = (GdfColumn * max_out)()
        n_out = C.c_int(0)
        lat = C.c_double(0)
        _call(lib().blz_server_wait, self.h, ticket, C.byref(kept), nulls, out, max_out, C.byref(n_out),
              C.byref(lat))
        cols = [DeviceColumn.adopt(out[i], owned=False) for i in range(n_out.value)]
        return kept.value, cols, [nulls[i] for i in range(n_out.value)], lat.value

    def release(self, ticket):
        lib().blz_server_release(self.h, ticket)
        self._keep.pop(ticket, None)

    def replay(self, queries, arrival_s, scalar_mode=SCALAR_INTENDED, keep_results=False):
        """Submits queries at the given offsets (seconds); returns (latency ms array,
        rows kept array, wall seconds), and with keep_results the tickets too: those
        queries stay resident until wait() has read them and release() frees them."""
        n = len(queries)
        keep = []
        arr = _queries(queries, scalar_mode, keep)
        arr_t = (C.c_double * n)(*[float(x) for x in arrival_s])
        lat = (C.c_double * n)()
        kept = (C.c_int64 * n)()
        wall = C.c_double(0)
        tickets = (C.c_int64 * n)() if keep_results else None
        _call(lib().blz_server_replay, self.h, arr, arr_t, n, lat, kept, C.byref(wall), tickets)
        if keep_results:
            for i, (frame, _, _) in enumerate(queries):
                self._keep[tickets[i]] = (keep, list(frame))
            return np.array(lat[:]), np.array(kept[:]), wall.value, list(tickets)
        return np.array(lat[:]), np.array(kept[:]), wall.value

    def submit_token(self, frame, filter_query_part, project_query_part, token, scalar_mode=SCALAR_INTENDED):
        """evaluate_query's asynchronous form: the query's outputs are parked in the
        result repository under `token` (registered) once it has run."""
        keep = []
        arr = _queries([(frame, filter_query_part, project_query_part)], scalar_mode, keep)
        _call(lib().blz_server_submit_token, self.h, arr, C.c_uint64(token))
        self._keep[("token", token)] = (keep, list(frame))

    def stats(self):
        nb, pk, used = C.c_int64(0), C.c_int64(0), C.c_size_t(0)
        lib().blz_server_stats(self.h, C.byref(nb), C.byref(pk), C.byref(used))
        return {"batches": nb.value, "peak_inflight": pk.value, "used_bytes": used.value}

    def close(self):
        if self.h.value:
            lib().blz_server_destroy(self.h)
            self.h = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _csv_args(dtypes, column_indices):
    dt = (C.c_int32 * len(dtypes))(*dtypes)
    if column_indices is None:
        return dt, None, len(dtypes), 0
    ci = (C.c_int32 * max(len(column_indices), 1))(*column_indices)
    return dt, ci, len(column_indices), len(column_indices)


def read_csv(data, dtypes, column_indices=None, delimiter="|", line_terminator="\n", skip_rows=0,
             stream=None, on_device=False):
    """csv_parser::parse over `data` (bytes in host memory, or with on_device a
    DeviceBuffer / device address and `data_len`): ([DeviceColumn], rows)."""
    dt, ci, n_out, n_idx = _csv_args(dtypes, column_indices)
    out = (GdfColumn * max(n_out, 1))()
    rows = C.c_int64(0)
    if on_device:
        ptr, nbytes = data
        _call(lib().blz_csv_parse_device, ptr, nbytes, delimiter.encode(), line_terminator.encode(), skip_rows,
              dt, len(dtypes), ci, n_idx, out, C.byref(rows), stream)
    else:
        buf = bytes(data)
        _call(lib().blz_csv_read, buf, len(buf), delimiter.encode(), line_terminator.encode(), skip_rows, dt,
              len(dtypes), ci, n_idx, out, C.byref(rows), stream)
    return [DeviceColumn.adopt(out[i], owned=True) for i in range(n_out)], rows.value


def parse_csv_device_timed(ptr, nbytes, dtypes, column_indices=None, delimiter="|", line_terminator="\n",
                           skip_rows=0, stream=None):
    """read_csv over device bytes, also returning the device time (ms) of its four
    kernels (row-start counts, scan, row starts, fields)."""
    dt, ci, n_out, n_idx = _csv_args(dtypes, column_indices)
    out = (GdfColumn * max(n_out, 1))()
    rows = C.c_int64(0)
    ms = (C.c_float * 4)()
    _call(lib().blz_csv_parse_device_timed, ptr, nbytes, delimiter.encode(), line_terminator.encode(), skip_rows,
          dt, len(dtypes), ci, n_idx, out, C.byref(rows), ms, stream)
    return [DeviceColumn.adopt(out[i], owned=True) for i in range(n_out)], rows.value, list(ms)


def load_csv_files(paths, dtypes, column_indices=None, delimiter="|", line_terminator="\n", skip_rows=0,
                   stream=None):
    """data_loader::load_data: every file parsed, columns concatenated in file order."""
    dt, ci, n_out, n_idx = _csv_args(dtypes, column_indices)
    ps = (C.c_char_p * len(paths))(*[p.encode() for p in paths])
    out = (GdfColumn * max(n_out, 1))()
    rows = C.c_int64(0)
    _call(lib().blz_csv_load_files, ps, len(paths), delimiter.encode(), line_terminator.encode(), skip_rows, dt,
          len(dtypes), ci, n_idx, out, C.byref(rows), stream)
    return [DeviceColumn.adopt(out[i], owned=True) for i in range(n_out)], rows.value


class ResultSet:
    """result_set_t (ResultSetRepository.h:19-25) as get_result returns it: the
    parked columns (views of the repository's buffers), their names and column
    tokens, the duration (ms) and the error text ("" on success)."""

    def __init__(self, columns, names, column_tokens, duration, error_msg):
        self.columns = columns
        self.names = names
        self.column_tokens = column_tokens
        self.duration = duration
        self.errorMsg = error_msg

    @property
    def is_ready(self):
        return True


class ResultSetRepository:
    """result_set_repository (src/ResultSetRepository.cpp): the process-wide
    repository of asynchronous results behind blz_rs_*."""

    @staticmethod
    def init_session():
        c = C.c_uint64(0)
        _call(lib().blz_rs_init_session, C.byref(c))
        return c.value

    @staticmethod
    def register_query(connection, token):
        _call(lib().blz_rs_register_query, C.c_uint64(connection), C.c_uint64(token))
        return token

    @staticmethod
    def update_token(token, columns: Sequence[DeviceColumn], duration=0.0, error_msg="", names=None):
        """Parks `columns`: the repository takes their device buffers (the Python
        objects no longer free them)."""
        n = len(columns)
        arr = (GdfColumn * max(n, 1))()
        for i, c in enumerate(columns):
            arr[i] = c.col
        nm = (C.c_char_p * max(n, 1))()
        for i in range(n):
            nm[i] = (names[i] if names else "").encode()
        _call(lib().blz_rs_update_token, C.c_uint64(token), arr, nm, n, float(duration),
              error_msg.encode() if error_msg else None)
        for c in columns:  # ownership moved to the repository
            c._owned = False
            for b in (c._data, c._valid):
                if b is not None:
                    b.ptr = C.c_void_p(0)

    @staticmethod
    def get_result(connection, token, max_columns=64) -> ResultSet:
        """Blocks until the token's frame is parked; counts one reference."""
        out = (GdfColumn * max_columns)()
        toks = (C.c_uint64 * max_columns)()
        n = C.c_int(0)
        dur = C.c_double(0)
        msg = C.create_string_buffer(1024)
        _call(lib().blz_rs_get_result, C.c_uint64(connection), C.c_uint64(token), out, toks, max_columns,
              C.byref(n), C.byref(dur), msg, len(msg))
        names = []
        for i in range(n.value):
            b = C.create_string_buffer(256)
            _check(lib().blz_rs_column_name(C.c_uint64(token), i, b, len(b)), "column name")
            names.append(b.value.decode())
        cols = [DeviceColumn.adopt(out[i], owned=False) for i in range(n.value)]
        return ResultSet(cols, names, [toks[i] for i in range(n.value)], dur.value,
                         msg.value.decode(errors="replace"))

    @staticmethod
    def try_free_result(connection, token) -> bool:
        found = C.c_int(0)
        _call(lib().blz_rs_try_free_result, C.c_uint64(connection), C.c_uint64(token), C.byref(found))
        return bool(found.value)

    @staticmethod
    def remove_all_connection_tokens(connection):
        _check(lib().blz_rs_remove_all_connection_tokens(C.c_uint64(connection)), "remove tokens")

    @staticmethod
    def get_column(connection, column_token) -> DeviceColumn:
        out = GdfColumn()
        _call(lib().blz_rs_get_column, C.c_uint64(connection), C.c_uint64(column_token), C.byref(out))
        return DeviceColumn.adopt(out, owned=False)


def synth_column(col: "DeviceColumn", seed: int, dist: int, a: float, b: float, null_rate=0.0, stream=None):
    """Fills an allocated device column with seeded synthetic values in place."""
    _call(lib().blz_synth_column, C.byref(col.col), seed, dist, a, b, null_rate, stream)

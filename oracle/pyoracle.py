"""Python access to the CPU oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

Host-memory restatement of the reference path, used by tests/, smoke() and the
CPU-baseline leg of bench.py as the checker. Columns are HostColumn (numpy data,
optional bool validity). End-to-end helpers compose the restated planner
(planner.py) with the restated kernels exactly as the reference composes them:
process_filter = stencil (INT8, validity preset on) -> ascending indices ->
per-column gather (CalciteInterpreter.cpp:509-573); execute_project =
parse_project_plan + perform_operation + update_null_count
(CalciteInterpreter.cpp:123-336).
"""
import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import planner as P
from .planner import (GDF_INT8, GDF_INT16, GDF_INT32, GDF_INT64, GDF_FLOAT32, GDF_FLOAT64,
                      GDF_BOOL8, GDF_DATE32, GDF_DATE64, GDF_TIMESTAMP, GDF_STRING_CATEGORY)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_build", "liboracle.so")

NP_OF = {GDF_INT8: np.int8, GDF_INT16: np.int16, GDF_INT32: np.int32, GDF_INT64: np.int64,
         GDF_FLOAT32: np.float32, GDF_FLOAT64: np.float64, GDF_BOOL8: np.uint8,
         GDF_DATE32: np.int32, GDF_DATE64: np.int64, GDF_TIMESTAMP: np.int64,
         GDF_STRING_CATEGORY: np.int32}


class _Data(C.Union):
    _fields_ = [("si08", C.c_int8), ("si16", C.c_int16), ("si32", C.c_int32),
                ("si64", C.c_int64), ("ui08", C.c_uint8), ("ui16", C.c_uint16),
                ("ui32", C.c_uint32), ("ui64", C.c_uint64), ("fp32", C.c_float),
                ("fp64", C.c_double), ("dt32", C.c_int32), ("dt64", C.c_int64),
                ("tmst", C.c_int64)]


class _Scalar(C.Structure):
    _fields_ = [("data", _Data), ("dtype", C.c_int32), ("is_valid", C.c_int32)]


class _Col(C.Structure):
    _fields_ = [("data", C.c_void_p), ("valid", C.c_void_p), ("size", C.c_int32),
                ("dtype", C.c_int32), ("null_count", C.c_int32), ("category", C.c_void_p),
                ("col_name", C.c_char_p)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.oracle_perform_operation.restype = C.c_int
        L.oracle_perform_operation.argtypes = [
            C.POINTER(_Col), C.c_int, C.POINTER(_Col), C.c_int, C.POINTER(C.c_int16),
            C.POINTER(C.c_int16), C.POINTER(C.c_int16), C.c_int, C.POINTER(C.c_int16),
            C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(_Scalar), C.POINTER(_Scalar),
            C.c_int, C.POINTER(C.c_int32)]
        L.oracle_stencil_select.restype = C.c_int64
        L.oracle_stencil_select.argtypes = [C.POINTER(_Col), C.POINTER(C.c_int32)]
        L.oracle_gather.restype = C.c_int32
        L.oracle_gather.argtypes = [C.POINTER(_Col), C.POINTER(_Col), C.POINTER(C.c_int32), C.c_int32]
        L.oracle_null_count.restype = C.c_int32
        L.oracle_null_count.argtypes = [C.c_void_p, C.c_int32]
        _lib = L
    return _lib


def _pack(mask, n):
    nb = max(((n + 7) // 8 + 63) // 64 * 64, 8)
    out = np.zeros(nb, dtype=np.uint8)
    if n:
        p = np.packbits(np.asarray(mask, dtype=bool), bitorder="little")
        out[:len(p)] = p
    return out


def _unpack(bits, n):
    return np.unpackbits(bits, bitorder="little")[:n].astype(bool)


@dataclass
class HostColumn:
    data: np.ndarray
    dtype: int
    valid: Optional[np.ndarray] = None      # bool[n] or None (no bitmap)
    null_count: Optional[int] = None         # None: computed from valid

    def __post_init__(self):
        self.data = np.ascontiguousarray(self.data, dtype=NP_OF[self.dtype])
        if self.valid is not None:
            self.valid = np.asarray(self.valid, dtype=bool)
        if self.null_count is None:
            self.null_count = 0 if self.valid is None else int((~self.valid).sum())

    def __len__(self):
        return len(self.data)


def _scalar(s) -> _Scalar:
    r = _Scalar()
    r.dtype = s.dtype
    r.is_valid = 1 if s.is_valid else 0
    f = {GDF_INT8: "si08", GDF_INT16: "si16", GDF_INT32: "si32", GDF_INT64: "si64",
         GDF_FLOAT32: "fp32", GDF_FLOAT64: "fp64", GDF_DATE32: "dt32", GDF_DATE64: "dt64",
         GDF_TIMESTAMP: "tmst", GDF_STRING_CATEGORY: "si32"}.get(s.dtype)
    if f and s.is_valid:
        setattr(r.data, f, s.value)
    return r


def perform_operation(inputs: List[HostColumn], plan: P.Plan, out_dtypes: List[int],
                      out_has_valid: List[bool], scalar_mode=0):
    """perform_operation restated; returns [HostColumn] with null counts set."""
    n = len(inputs[0]) if inputs else 0
    keep = []
    cin = (_Col * max(len(inputs), 1))()
    for i, h in enumerate(inputs):
        cin[i].data = h.data.ctypes.data if len(h.data) else None
        vb = _pack(h.valid, n) if h.valid is not None else None
        keep.append(vb)
        cin[i].valid = vb.ctypes.data if vb is not None else None
        cin[i].size = n
        cin[i].dtype = h.dtype
        cin[i].null_count = h.null_count
    outs, cout = [], (_Col * max(len(out_dtypes), 1))()
    for k, (dt, hv) in enumerate(zip(out_dtypes, out_has_valid)):
        d = np.zeros(n + 8, dtype=NP_OF.get(dt, np.int64))
        vb = np.full(max(((n + 7) // 8 + 63) // 64 * 64, 8), 0xFF, dtype=np.uint8) if hv else None
        outs.append((d, vb))
        cout[k].data = d.ctypes.data
        cout[k].valid = vb.ctypes.data if vb is not None else None
        cout[k].size = n
        cout[k].dtype = dt
    m = len(plan.left)
    I16 = C.c_int16 * max(m, 1)
    I32 = C.c_int32 * max(m, 1)
    SC = _Scalar * max(m, 1)
    nulls = (C.c_int32 * max(len(out_dtypes), 1))()
    rc = lib().oracle_perform_operation(
        cout, len(out_dtypes), cin, len(inputs), I16(*plan.left), I16(*plan.right),
        I16(*plan.outputs), m, (C.c_int16 * max(len(plan.final_positions), 1))(*plan.final_positions),
        I32(*plan.binops), I32(*plan.unops), SC(*[_scalar(s) for s in plan.lscalars]),
        SC(*[_scalar(s) for s in plan.rscalars]), scalar_mode, nulls)
    if rc != 0:
        raise ValueError(f"oracle_perform_operation failed ({rc})")
    res = []
    for k, (d, vb) in enumerate(outs):
        res.append(HostColumn(d[:n].copy(), out_dtypes[k],
                              _unpack(vb, n) if vb is not None else None, nulls[k]))
    return res


def stencil_select(stencil: HostColumn) -> np.ndarray:
    n = len(stencil)
    c = _Col()
    c.data = stencil.data.ctypes.data if n else None
    vb = _pack(stencil.valid, n) if stencil.valid is not None else None
    c.valid = vb.ctypes.data if vb is not None else None
    c.size = n
    c.dtype = stencil.dtype
    idx = np.zeros(max(n, 1), dtype=np.int32)
    m = lib().oracle_stencil_select(C.byref(c), idx.ctypes.data_as(C.POINTER(C.c_int32)))
    return idx[:m].copy()


def gather(col: HostColumn, idx: np.ndarray, with_valid: bool) -> HostColumn:
    n = len(col)
    m = len(idx)
    cin, cout = _Col(), _Col()
    cin.data = col.data.ctypes.data if n else None
    vin = _pack(col.valid, n) if col.valid is not None else None
    cin.valid = vin.ctypes.data if vin is not None else None
    cin.size, cin.dtype = n, col.dtype
    d = np.zeros(m + 8, dtype=NP_OF[col.dtype])
    vb = np.zeros(max((m + 7) // 8 + 8, 8), dtype=np.uint8) if with_valid else None
    cout.data = d.ctypes.data
    cout.valid = vb.ctypes.data if vb is not None else None
    cout.size, cout.dtype = m, col.dtype
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    nulls = lib().oracle_gather(C.byref(cin), C.byref(cout), idx.ctypes.data_as(C.POINTER(C.c_int32)), m)
    return HostColumn(d[:m].copy(), col.dtype, _unpack(vb, m) if vb is not None else None, nulls)


def null_count(col: HostColumn) -> int:
    if col.valid is None or len(col) == 0:
        return 0
    vb = _pack(col.valid, len(col))
    return lib().oracle_null_count(vb.ctypes.data, len(col))


# ---- end-to-end compositions ---------------------------------------------------------

def evaluate_expression(frame: List[HostColumn], expression: str, out_dtype=GDF_INT8,
                        out_has_valid=True, scalar_mode=0) -> HostColumn:
    """evaluate_expression (LogicalFilter.cpp:271-371) into a new column."""
    plan = P.plan_expression([c.dtype for c in frame], expression, out_dtype)
    if plan is None:  # bare "$k": a clone of the column
        k = P.get_index(expression)
        c = frame[k]
        return HostColumn(c.data.copy(), c.dtype, None if c.valid is None else c.valid.copy(),
                          c.null_count)
    ins = [frame[i] for i in plan.input_columns]
    return perform_operation(ins, plan, [out_dtype], [out_has_valid], scalar_mode)[0]


def process_filter(frame: List[HostColumn], query_part: str, scalar_mode=0) -> List[HostColumn]:
    n = len(frame[0])
    if n == 0:
        return [HostColumn(c.data[:0], c.dtype, None if c.valid is None else c.valid[:0]) for c in frame]
    cond = P.named_expression(query_part, "condition") if "condition=[" in query_part else query_part
    stencil = evaluate_expression(frame, cond, GDF_INT8, True, scalar_mode)
    idx = stencil_select(stencil)
    return [gather(c, idx, c.valid is not None) for c in frame]


def execute_project(frame: List[HostColumn], query_part: str, scalar_mode=0) -> List[HostColumn]:
    dts = [c.dtype for c in frame]
    plan, outs, _names = P.parse_project(dts, query_part)
    evald = []
    if plan is not None and len(frame[0]) > 0:
        ins = [frame[i] for i in plan.input_columns]
        evald = perform_operation(ins, plan, plan.output_dtypes, [True] * len(plan.output_dtypes),
                                  scalar_mode)
    elif plan is not None:
        evald = [HostColumn(np.zeros(0, NP_OF[d]), d, np.zeros(0, bool)) for d in plan.output_dtypes]
    res = []
    for kind, idx in outs:
        if kind == "expr":
            res.append(evald[idx])
        else:
            c = frame[idx]
            res.append(HostColumn(c.data, c.dtype, c.valid, null_count(c)))
    return res


def run_logical_plan(tables: List[HostColumn], plan_text: str, scalar_mode=0) -> List[HostColumn]:
    """Bottom-up walk of a scan / filter / project plan (CalciteInterpreter.cpp:731-870)."""
    lines = [l.strip() for l in plan_text.strip().split("\n") if l.strip()]
    frame = list(tables)
    for line in reversed(lines):
        if "TableScan" in line:
            continue
        if line.startswith("LogicalFilter"):
            frame = process_filter(frame, line, scalar_mode)
        elif line.startswith("LogicalProject"):
            frame = execute_project(frame, line, scalar_mode)
        else:
            raise ValueError(f"unsupported plan line: {line}")
    return frame


def process_aggregate(frame: List[HostColumn], query_part: str, scalar_mode=0) -> List[HostColumn]:
    """process_aggregate without GROUP BY (GroupBy.cpp:531-588 -> compute_aggregations
    :292-336 -> aggregations_without_groupby :230-290): each aggregation's input is a
    frame column ("$k"; empty = column 0, get_index CalciteExpressionParsing.cpp:600-608)
    or an expression evaluated first; the result is a one-row column. COUNT = rows -
    null_count; SUM / MIN / MAX over the valid rows (cudf::reduce skips nulls; no
    valid row -> null); AVG = SUM / (rows - null_count) in binary64, null when no row
    is valid. An evaluated input's null_count stays 0 (evaluate_expression never
    updates it, LogicalFilter.cpp:368), and a column whose null_count is 0 is read
    as all valid (the interpreter's input rule, interpreter_ops.cuh:1312-1327), so
    COUNT(expr) counts every row and AVG(expr) divides the sum of every row by it
    (GroupBy.cpp:263, :284). An integer SUM wraps in int64; a float SUM here is the
    exact sum rounded once (math.fsum): the reference's summation order is cuDF's,
    not pinned."""
    import math
    groups, aggs = P.parse_aggregate(query_part)
    if groups:
        raise P.PlanError("aggregation with GROUP BY is not restated")
    dts = [c.dtype for c in frame]
    out = []
    for op, inp in aggs:
        if P.contains_evaluation(inp):
            col = evaluate_expression(frame, inp, P.get_output_type_expression(dts, inp), True, scalar_mode)
            col = HostColumn(col.data, col.dtype, col.valid, 0)
        else:
            col = frame[P.get_index(inp) if inp else 0]
        ot = P.aggregation_output_type(col.dtype, op)
        n = len(col)
        valid_rows = n - (col.null_count if col.valid is not None else 0)
        vals = col.data if (col.valid is None or col.null_count == 0) else col.data[col.valid]
        is_f = col.dtype in (GDF_FLOAT32, GDF_FLOAT64)
        ok = True
        if op == "COUNT":
            v = valid_rows
        elif op == "AVG":
            ok = n > 0 and valid_rows > 0
            if ok:
                s = math.fsum(vals.astype(np.float64)) if is_f else float(_wrap_sum(vals))
                v = s / float(valid_rows)
        else:
            ok = n > 0 and len(vals) > 0
            if ok and op == "SUM":
                v = math.fsum(vals.astype(np.float64)) if is_f else _wrap_sum(vals)
            elif ok and is_f:
                finite = vals[~np.isnan(vals)]   # a NaN never replaces the running min / max
                v = (finite.min() if op == "MIN" else finite.max()) if len(finite) else \
                    (np.inf if op == "MIN" else -np.inf)
            elif ok:
                v = vals.astype(np.int64).min() if op == "MIN" else vals.astype(np.int64).max()
        data = np.zeros(1, NP_OF[ot])
        if ok:
            data[0] = v
        out.append(HostColumn(data, ot, np.array([ok])))
    return out


def _wrap_sum(vals):
    """Sum of integers in int64 with wrap-around (two's complement)."""
    return int(((int(np.sum(vals.astype(object))) + 2**63) % 2**64) - 2**63)

"""Seeded synthetic workloads of the batched and served benchmarks.

Tables are generated in place on the device (blz_synth_column: row i depends only
on (seed, i)), so no host copy of the tens of gigabytes involved is made.

Wide-expression operator set (arithmetic, integer modulo by a non-zero literal,
CAST of int64 to float64, comparisons, AND / OR / NOT, and the nested conditional
CASE(IS NOT NULL(x), x, y) that the planner rewrites to COALESCE):
  rand_expr(rng, dtypes, n_ops, nullable) builds an expression with exactly n_ops
  operators over the columns of a table.
"""
import math

import numpy as np

from . import (GDF_FLOAT32, GDF_FLOAT64, GDF_INT32, GDF_INT64, DeviceColumn, synth_column,
               update_null_count)

COLUMN_KINDS = {
    # dtype -> (distribution, a, b) for blz_synth_column
    GDF_INT32: (0, -1_000_000, 1_000_000),
    GDF_INT64: (0, 0, 2 ** 40),
    GDF_FLOAT32: (2, 0.0, 1.0),
    GDF_FLOAT64: (3, 0.0, 2.0),
}
# The wide-expression workload: 8 columns (2 int32 uniform [-1e6, 1e6], 2 int64
# uniform [0, 2^40], 2 float32 normal(0, 1), 2 float64 lognormal(0, 2)), 5 % nulls
# each, as (dtype, blz_synth_column distribution, a, b); and one 24-op projection
# with MOD, CAST int64 -> float64, comparisons, AND / OR / NOT and a nested
# conditional (CASE -> COALESCE), four outputs.
WIDE_COLUMNS = [(GDF_INT32, 0, -1e6, 1e6), (GDF_INT32, 0, -1e6, 1e6), (GDF_INT64, 0, 0, 2.0 ** 40),
                (GDF_INT64, 0, 0, 2.0 ** 40), (GDF_FLOAT32, 2, 0.0, 1.0), (GDF_FLOAT32, 2, 0.0, 1.0),
                (GDF_FLOAT64, 3, 0.0, 2.0), (GDF_FLOAT64, 3, 0.0, 2.0)]
WIDE_NULL_RATE = 0.05
WIDE_PROJECT = ("LogicalProject(E0=[+(*($0, 3), MOD(-($1, $0), 7))], "
                "E1=[CASE(IS NOT NULL($6), $6, +(BL_CAST_DOUBLE($2), -($7, 1.0)))], "
                "E2=[AND(<($4, 0.5), OR(>($5, -0.5), NOT(=($0, $1))), <>($2, 0))], "
                "E3=[-(*(+($2, MOD($3, 1000)), 2), /(*($6, $7), +(*($4, $5), 1.5)))])")

ARITH = ["+", "-", "*"]
CMP = ["=", "<>", "<", ">", "<=", ">="]


def _lit(rng, is_int):
    return str(int(rng.integers(1, 100))) if is_int else repr(round(float(rng.uniform(0.5, 50)), 2))


def rand_expr(rng, dtypes, n_ops, nullable=(), case_left=None):
    """Expression string with exactly n_ops operators (n_ops >= 1). case_left: a
    one-element list, the CASE conditionals still allowed (shared by the
    expressions of one query)."""
    k = len(dtypes)
    ints = [i for i, d in enumerate(dtypes) if d in (GDF_INT32, GDF_INT64)]
    i64 = [i for i, d in enumerate(dtypes) if d == GDF_INT64]

    def col():
        return f"${int(rng.integers(0, k))}"

    # the planner rewrites one CASE(IS NOT NULL(x), x, y) per query part
    # (clean_calcite_expression): at most one, not nested
    if case_left is None:
        case_left = [1]

    def build(ops):
        if ops == 0:
            return col()
        r = rng.random()
        if r < 0.08:
            return f"NOT({build(ops - 1)})"
        if r < 0.14 and i64:
            return f"BL_CAST_DOUBLE(${int(rng.choice(i64))})" if ops == 1 else \
                f"+(BL_CAST_DOUBLE(${int(rng.choice(i64))}), {build(ops - 2)})"
        if r < 0.22 and ints:
            return f"MOD({build(ops - 1)}, {_lit(rng, True)})"
        if r < 0.30 and nullable and case_left[0]:
            case_left[0] = 0
            c = int(rng.choice(nullable))
            return f"CASE(IS NOT NULL(${c}), ${c}, {build(ops - 1)})"
        if r < 0.45:
            op = CMP[int(rng.integers(0, len(CMP)))]
        elif r < 0.55:
            op = "AND" if rng.random() < 0.5 else "OR"
        else:
            op = ARITH[int(rng.integers(0, len(ARITH)))]
        rest = ops - 1
        if rest == 0 or rng.random() < 0.3:
            return f"{op}({build(rest)}, {_lit(rng, rng.random() < 0.5)})"
        left = int(rng.integers(0, rest + 1))
        return f"{op}({build(left)}, {build(rest - left)})"

    return build(n_ops)


def count_ops(expr):
    """Operators of a Calcite expression string (the planner's tokens)."""
    n = 0
    for tok in ("+(", "-(", "*(", "/(", "MOD(", "NOT(", "BL_CAST_DOUBLE(", "AND(", "OR(", "=(", "<>(", "<(",
                ">(", "<=(", ">=(", "CASE("):
        n += expr.count(tok)
    # "<(" also matches inside "<=(" and "<>(", ">(" inside ">=(": undo the double count
    n -= expr.count("<=(") + expr.count("<>(") + expr.count(">=(")
    return n


def table_spec(rng, n_rows, n_cols=None, null_max=0.10, dtypes=None):
    if dtypes is None:
        n_cols = n_cols or int(rng.integers(4, 9))
        pool = [GDF_INT32, GDF_INT64, GDF_FLOAT32, GDF_FLOAT64]
        dtypes = [pool[int(rng.integers(0, 4))] for _ in range(n_cols)]
    nulls = [float(rng.uniform(0, null_max)) for _ in dtypes]
    seeds = [int(rng.integers(1, 2 ** 62)) for _ in dtypes]
    return {"rows": int(n_rows), "dtypes": dtypes, "nulls": nulls, "seeds": seeds}


def table_bytes(spec):
    w = {GDF_INT32: 4, GDF_INT64: 8, GDF_FLOAT32: 4, GDF_FLOAT64: 8}
    return sum(spec["rows"] * w[d] + (spec["rows"] + 7) // 8 * (nr > 0) for d, nr in zip(spec["dtypes"],
                                                                                           spec["nulls"]))


def make_table(spec):
    cols = []
    for d, nr, sd in zip(spec["dtypes"], spec["nulls"], spec["seeds"]):
        c = DeviceColumn(d, spec["rows"], with_valid=nr > 0)
        dist, a, b = COLUMN_KINDS[d]
        synth_column(c, sd, dist, a, b, nr)
        if nr > 0:
            update_null_count(c)  # a bitmap counts only with a non-zero null count
        cols.append(c)
    return cols


def log_uniform_rows(rng, lo, hi):
    return int(math.exp(rng.uniform(math.log(lo), math.log(hi))))


def rand_query(rng, spec, with_filter, ops_lo=4, ops_hi=40):
    """(filter or None, projection) with ops_lo..ops_hi operators in total over the
    table of `spec`; 1-4 projected expressions."""
    dtypes = spec["dtypes"]
    nullable = [i for i, nr in enumerate(spec["nulls"]) if nr > 0]
    case_left = [1]
    total = int(rng.integers(ops_lo, ops_hi + 1))
    n_out = int(rng.integers(1, 5))
    filt = None
    if with_filter:
        f_ops = max(1, total // 4)
        c = int(rng.integers(0, len(dtypes)))
        lit = _lit(rng, dtypes[c] in (GDF_INT32, GDF_INT64))
        inner = rand_expr(rng, dtypes, f_ops - 1, nullable, [0]) if f_ops > 1 else f"${c}"
        filt = f"LogicalFilter(condition=[<({inner}, {lit})])"
        total -= f_ops
    n_out = max(1, min(n_out, total))
    per = [total // n_out + (1 if i < total % n_out else 0) for i in range(n_out)]
    items = [f"E{i}=[{rand_expr(rng, dtypes, max(1, p), nullable, case_left)}]" for i, p in enumerate(per)]
    return filt, "LogicalProject(" + ", ".join(items) + ")"

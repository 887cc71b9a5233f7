"""The oracle's aggregation without GROUP BY (CPU only), pinned on the reference's
own test (tests/integration/aggregation-without-group-by/
test-aggregation-without-group-by.cpp) and on closed forms."""
import math

import numpy as np
import pytest

from oracle import planner as P
from oracle import pyoracle as O

ALL5 = ("LogicalAggregate(group=[{}], EXPR$0=[COUNT()], EXPR$1=[SUM($0)], EXPR$2=[AVG($0)], "
        "EXPR$3=[MAX($0)], EXPR$4=[MIN($0)])")


def _vals(cols):
    return [(c.dtype, c.data[0].item(), bool(c.valid[0])) for c in cols]


def test_reference_test_01_count_sum_avg_max_min():
    # TEST_01 (:66-112): p_partkey = 1..200 (INT32) -> 200, 20100, 100.5, 200, 1
    part = O.HostColumn(np.arange(1, 201, dtype=np.int32), P.GDF_INT32)
    got = _vals(O.process_aggregate([part], ALL5))
    assert got == [(P.GDF_INT64, 200, True), (P.GDF_INT64, 20100, True), (P.GDF_FLOAT64, 100.5, True),
                   (P.GDF_INT32, 200, True), (P.GDF_INT32, 1, True)]


def test_reference_test_03_no_rows():
    # TEST_03 (:116-175): after a filter keeping no row: COUNT 0 (valid), the rest null
    part = O.HostColumn(np.arange(1, 201, dtype=np.int32), P.GDF_INT32)
    none = O.process_filter([part], "LogicalFilter(condition=[<($0, 0)])")
    got = O.process_aggregate(none, ALL5)
    assert [bool(c.valid[0]) for c in got] == [True, False, False, False, False]
    assert got[0].data[0] == 0


def test_parse_aggregate():
    g, aggs = P.parse_aggregate("LogicalAggregate(group=[{}], revenue=[SUM($0)], n=[COUNT()])")
    assert g == [] and aggs == [("SUM", "$0"), ("COUNT", "")]
    g, _ = P.parse_aggregate("LogicalAggregate(group=[{0, 2}], EXPR$0=[SUM($1)])")
    assert g == [0, 2]
    with pytest.raises(P.PlanError):
        P.parse_aggregate("LogicalAggregate(group=[{}], EXPR$0=[MEDIAN($0)])")


def test_closed_forms_with_nulls():
    rng = np.random.default_rng(5)
    x = rng.integers(-10**6, 10**6, 5000).astype(np.int64)
    v = rng.random(5000) > 0.2
    col = O.HostColumn(x, P.GDF_INT64, v)
    got = O.process_aggregate([col], ALL5)
    kept = x[v]
    assert got[0].data[0] == 5000 - int((~v).sum())          # COUNT() = rows - null_count
    assert got[1].data[0] == int(kept.sum())
    assert got[2].data[0] == float(int(kept.sum())) / len(kept)
    assert got[3].data[0] == kept.max() and got[4].data[0] == kept.min()


def test_integer_sum_wraps_in_int64_and_float_sum_is_exact():
    big = O.HostColumn(np.array([2**62, 2**62, 5], dtype=np.int64), P.GDF_INT64)
    s = O.process_aggregate([big], "LogicalAggregate(group=[{}], EXPR$0=[SUM($0)])")[0]
    assert s.data[0] == ((2**63 + 5 + 2**63) % 2**64) - 2**63   # wrapped
    f = O.HostColumn(np.array([1e16, 1.0, -1e16, 1.0]), P.GDF_FLOAT64)
    s = O.process_aggregate([f], "LogicalAggregate(group=[{}], EXPR$0=[SUM($0)])")[0]
    assert s.data[0] == math.fsum([1e16, 1.0, -1e16, 1.0]) == 2.0


def test_all_null_input_and_expression_input():
    col = O.HostColumn(np.arange(10, dtype=np.int32), P.GDF_INT32, np.zeros(10, bool))
    got = O.process_aggregate([col], ALL5)
    assert [bool(c.valid[0]) for c in got] == [True, False, False, False, False]
    assert got[0].data[0] == 0
    col = O.HostColumn(np.arange(10, dtype=np.int32), P.GDF_INT32)
    s = O.process_aggregate([col], "LogicalAggregate(group=[{}], EXPR$0=[SUM(*($0, 2))])")[0]
    assert s.dtype == P.GDF_INT64 and s.data[0] == 90


def test_count_and_avg_over_a_nullable_expression_count_every_row():
    """compute_aggregations evaluates an expression input into a new column
    (GroupBy.cpp:318-319) whose null_count evaluate_expression never updates
    (LogicalFilter.cpp:368), so COUNT = size - 0 (:284) and AVG's divisor is the
    row count (:263): the expression's nulls are counted, and its sum reads every
    row, as for any column whose null_count is 0."""
    rng = np.random.default_rng(6)
    n = 3000
    x = rng.integers(-1000, 1000, n).astype(np.int32)
    v = rng.random(n) > 0.3
    col = O.HostColumn(x, P.GDF_INT32, v)
    qp = ("LogicalAggregate(group=[{}], c=[COUNT(+($0, 1))], a=[AVG(+($0, 1))], s=[SUM(+($0, 1))], "
          "cc=[COUNT($0)], ac=[AVG($0)])")
    got = O.process_aggregate([col], qp)
    y = x.astype(np.int64) + 1                      # every row, null ones included
    assert got[0].data[0] == n
    assert got[1].data[0] == float(int(y.sum())) / n
    assert got[2].data[0] == int(y.sum())
    # the column itself keeps its null count: its nulls are not counted
    assert got[3].data[0] == int(v.sum())
    assert got[4].data[0] == float(int(x[v].astype(np.int64).sum())) / int(v.sum())


def test_null_count_zero_bitmap_is_read_as_all_valid():
    x = np.arange(8, dtype=np.int64)
    col = O.HostColumn(x, P.GDF_INT64, np.array([True, False] * 4), 0)   # inconsistent: null_count 0
    got = O.process_aggregate([col], "LogicalAggregate(group=[{}], s=[SUM($0)], c=[COUNT($0)])")
    assert got[0].data[0] == 28 and got[1].data[0] == 8

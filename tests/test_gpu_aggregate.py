"""GPU parity of blz_process_aggregate (aggregation without GROUP BY) against the
oracle: COUNT, MIN, MAX and integer SUM / AVG bit-exact; a float SUM (and the AVG
derived from it) within the bound of a summation in any order,
|gpu - exact| <= (n + 2) * 2^-53 * sum(|x|) + |exact| * 2^-52, since the reference's
order (cuDF's) is not pinned."""
import numpy as np
import pytest

from oracle import planner as P
from oracle import pyoracle as O
from tests.helpers import C1_FILTER, C1_PROJECT, c1_frame, gen_column, to_device

pytestmark = pytest.mark.gpu

ALL6 = ("LogicalAggregate(group=[{}], EXPR$0=[COUNT()], EXPR$1=[SUM($0)], EXPR$2=[AVG($0)], "
        "EXPR$3=[MAX($0)], EXPR$4=[MIN($0)], c=[COUNT($0)])")


@pytest.fixture(scope="module")
def B():
    import blz_ral
    if blz_ral.device_count() < 1:
        pytest.fail("no HIP device visible")
    return blz_ral


def _case(**kw):
    print("CASE", " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


def _check(got, ref, inputs, what):
    assert len(got) == len(ref)
    for k, (g, r) in enumerate(zip(got, ref)):
        assert g.col.dtype == r.dtype and g.col.size == 1, f"{what}: output {k} dtype / size"
        gv = bool(g.valid()[0])
        assert gv == bool(r.valid[0]), f"{what}: output {k} validity"
        assert g.col.null_count == (0 if gv else 1)
        if not gv:
            continue
        gd, rd = g.data()[0], r.data[0]
        src = inputs[k]
        if r.dtype == P.GDF_FLOAT64 and src is not None and src.dtype in (P.GDF_FLOAT32, P.GDF_FLOAT64):
            x = src.data if src.valid is None else src.data[src.valid]
            x = np.abs(x.astype(np.float64))
            n = len(x)
            scale = float(x.sum()) if r.dtype == P.GDF_FLOAT64 else 0.0
            if "AVG" in what.split("|")[k]:
                scale /= max(n, 1)
            bound = (n + 2) * 2.0**-53 * scale + abs(float(rd)) * 2.0**-52
            assert abs(float(gd) - float(rd)) <= bound, f"{what}: output {k} {gd!r} vs {rd!r} (bound {bound})"
        else:
            assert gd.tobytes() == np.asarray(rd).tobytes(), f"{what}: output {k} {gd!r} vs {rd!r}"


@pytest.mark.parametrize("dtype", [P.GDF_INT8, P.GDF_INT16, P.GDF_INT32, P.GDF_INT64, P.GDF_FLOAT32,
                                   P.GDF_FLOAT64, P.GDF_DATE32, P.GDF_DATE64])
@pytest.mark.parametrize("n", [1, 1000, 262_145, 3_000_000])
def test_aggregates_every_dtype(B, dtype, n):
    rng = np.random.default_rng(8800 + 7 * dtype + n % 1000)
    host = [gen_column(rng, dtype, n, 0.1)]
    _case(test="aggregates", dtype=dtype, n=n)
    ref = O.process_aggregate(host, ALL6)
    dev = [to_device(B, h) for h in host]
    got = B.process_aggregate(dev, ALL6)
    labels = "COUNT|SUM|AVG|MAX|MIN|COUNT"
    _check(got, ref, [None] + [host[0]] * 4 + [None], labels)
    for c in got + dev:
        c.free()


def test_reference_test_01_and_03(B):
    # test-aggregation-without-group-by.cpp TEST_01 (:66-112) and TEST_03 (:116-175)
    part = O.HostColumn(np.arange(1, 201, dtype=np.int32), P.GDF_INT32)
    q = ALL6.replace(", c=[COUNT($0)]", "")
    dev = [to_device(B, part)]
    got = B.process_aggregate(dev, q)
    vals = [(g.col.dtype, g.data()[0].item(), bool(g.valid()[0])) for g in got]
    assert vals == [(P.GDF_INT64, 200, True), (P.GDF_INT64, 20100, True), (P.GDF_FLOAT64, 100.5, True),
                    (P.GDF_INT32, 200, True), (P.GDF_INT32, 1, True)]
    none, kept = B.filter_project(dev, "LogicalFilter(condition=[<($0, 0)])", None)
    assert kept == 0
    got2 = B.process_aggregate(none, q)
    assert [bool(g.valid()[0]) for g in got2] == [True, False, False, False, False]
    assert got2[0].data()[0] == 0
    for c in got + got2 + none + dev:
        c.free()


def test_all_null_wrapping_sum_and_expression_input(B):
    n = 5000
    allnull = O.HostColumn(np.arange(n, dtype=np.int32), P.GDF_INT32, np.zeros(n, bool))
    big = O.HostColumn(np.full(n, 2**62 - 1, dtype=np.int64), P.GDF_INT64)
    f = gen_column(np.random.default_rng(3), P.GDF_FLOAT64, n, 0.0)
    host = [allnull, big, f]
    q = ("LogicalAggregate(group=[{}], a=[COUNT($0)], b=[SUM($0)], c=[MIN($0)], d=[AVG($0)], "
         "e=[SUM($1)], g=[SUM(*($1, 3))], h=[MAX(+($2, 1.5))], k=[MIN(*($2, -2.0))])")
    _case(test="aggregate_edges")
    ref = O.process_aggregate(host, q)
    dev = [to_device(B, h) for h in host]
    got = B.process_aggregate(dev, q)
    _check(got, ref, [None] * 8, "|".join(["x"] * 8))
    assert [bool(g.valid()[0]) for g in got[:4]] == [True, False, False, False]
    assert got[4].data()[0] == np.int64(((n * (2**62 - 1) + 2**63) % 2**64) - 2**63)
    for c in got + dev:
        c.free()


def test_group_by_is_unsupported(B):
    dev = [to_device(B, O.HostColumn(np.arange(10, dtype=np.int32), P.GDF_INT32))]
    with pytest.raises(B.BlzError) as e:
        B.process_aggregate(dev, "LogicalAggregate(group=[{0}], EXPR$0=[SUM($0)])")
    assert e.value.code == B.BLZ_ERR_UNSUPPORTED
    dev[0].free()


def test_q6_filter_project_sum(B):
    """TPC-H Q6's shape on the filter+project workload: SUM over the projected
    b*(1-c) of the kept rows, twice (the same bits: the reduction order is fixed)."""
    n = 1_000_003
    host = c1_frame(n, seed=66)
    _case(test="q6", n=n)
    mid = O.execute_project(O.process_filter(host, C1_FILTER), C1_PROJECT)
    q = "LogicalAggregate(group=[{}], revenue=[SUM($0)], n=[COUNT()], m=[MAX($1)])"
    ref = O.process_aggregate(mid, q)
    dev = [to_device(B, h) for h in host]
    outs, kept = B.filter_project(dev, C1_FILTER, C1_PROJECT)
    assert kept == len(mid[0])
    got = B.process_aggregate(outs, q)
    _check(got, ref, [mid[0], None, None], "SUM|COUNT|MAX")
    again = B.process_aggregate(outs, q)
    assert again[0].data().tobytes() == got[0].data().tobytes()
    for c in got + again + outs + dev:
        c.free()


def test_count_avg_sum_over_a_nullable_expression(B):
    """An expression input keeps null_count 0 after evaluate_expression
    (LogicalFilter.cpp:368): COUNT(expr) is the row count, AVG(expr) divides the sum
    of every row by it (GroupBy.cpp:263, :284), SUM / MIN / MAX read every row; the
    nullable column itself counts only its valid rows. Integer inputs: bit-exact."""
    rng = np.random.default_rng(91)
    n = 200_003
    x = gen_column(rng, P.GDF_INT32, n, 0.3)
    y = gen_column(rng, P.GDF_INT64, n, 0.1)
    host = [x, y]
    q = ("LogicalAggregate(group=[{}], c=[COUNT(+($0, 1))], a=[AVG(+($0, 1))], s=[SUM(*($0, $1))], "
         "mx=[MAX(-($1, $0))], mn=[MIN(+($0, 1))], cc=[COUNT($0)], ac=[AVG($0)], sc=[SUM($1)])")
    _case(test="aggregate_nullable_expression", n=n)
    ref = O.process_aggregate(host, q)
    assert ref[0].data[0] == n                      # every row counted
    dev = [to_device(B, h) for h in host]
    got = B.process_aggregate(dev, q)
    _check(got, ref, [None] * 8, "|".join(["x"] * 8))
    for c in got + dev:
        c.free()


@pytest.mark.parametrize("dtype", [P.GDF_INT8, P.GDF_INT16, P.GDF_INT32, P.GDF_INT64, P.GDF_FLOAT32,
                                   P.GDF_FLOAT64])
@pytest.mark.parametrize("nulls", [0.0, 0.2])
def test_aggregates_of_a_column_view_not_16_byte_aligned(B, dtype, nulls):
    """A column whose data starts one element into an allocation (a view, as a
    slice of another column would be) is read element by element, in the same
    order as an aligned one: the same results, integer bit-exact."""
    rng = np.random.default_rng(515 + dtype)
    n = 300_001
    full = gen_column(rng, dtype, n + 1, 0.0)
    sub = O.HostColumn(full.data[1:].copy(), dtype, (rng.random(n) >= nulls) if nulls else None)
    _case(test="aggregate_view", dtype=dtype, nulls=nulls, n=n)
    ref = O.process_aggregate([sub], ALL6)
    base = to_device(B, full)
    bits = to_device(B, sub)              # owns the view's bitmap (and an aligned copy)
    view = B.GdfColumn()
    w = np.dtype(full.data.dtype).itemsize
    view.data = base.col.data + w
    view.valid = bits.col.valid
    view.size = n
    view.dtype = dtype
    view.null_count = bits.col.null_count
    vcol = B.DeviceColumn.adopt(view, owned=False)
    got = B.process_aggregate([vcol], ALL6)
    labels = "COUNT|SUM|AVG|MAX|MIN|COUNT"
    _check(got, ref, [None] + [sub] * 4 + [None], labels)
    aligned = B.process_aggregate([bits], ALL6)
    for g, a in zip(got, aligned):
        assert g.data().tobytes() == a.data().tobytes()
    for c in got + aligned + [base, bits]:
        c.free()

"""The allocation-free perform_operation (blz_perform_operation_noalloc) against
blz_perform_operation through the C ABI: the same plans, the same columns, the
same output bytes, validity and null counts, over two calls in a row (the second
with a smaller input and a different plan), and again after the other entry
point has run."""
import numpy as np
import pytest

from oracle import planner as P
from tests.helpers import gen_column, to_device

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def B():
    import blz_ral
    if blz_ral.device_count() < 1:
        pytest.fail("no HIP device visible")
    return blz_ral


def _plan(B, dtypes, expr):
    pl = B.plan_expression(dtypes, expr)
    return pl


def _run(B, fn, dev_in, pl, out_dtype, n):
    out = B.DeviceColumn(out_dtype, n, with_valid=True)
    ins = [dev_in[c] for c in pl["input_columns"]]
    nulls = fn([out], ins, pl["left"], pl["right"], pl["outputs"], [pl["final_position"]], pl["binops"],
               pl["unops"], pl["lscalars"], pl["rscalars"])
    return out, nulls[0]


CALLS = [
    # (rows, expression, output dtype): the second call is smaller, the third has a
    # transcendental (the general op set)
    (300_017, "+(*($0, $1), CASE(IS NOT NULL($2), $2, -($3, 7)))", P.GDF_FLOAT64),
    (5_003, "AND(>($0, 10), <(MOD($1, 7), 3))", P.GDF_INT8),
    (70_001, "*(SIN($2), +($3, 2.5))", P.GDF_FLOAT64),
]


def test_noalloc_matches_perform_operation(B):
    rng = np.random.default_rng(77_000)
    dts = [P.GDF_INT32, P.GDF_INT64, P.GDF_FLOAT64, P.GDF_FLOAT32]
    for rows, expr, out_dt in CALLS:
        print(f"CASE noalloc rows={rows} expr={expr}", flush=True)
        host = [gen_column(rng, dt, rows, 0.1) for dt in dts]
        dev = [to_device(B, h) for h in host]
        try:
            pl = _plan(B, dts, expr)
        except B.BlzError as e:  # the planner rejects what it must reject for both entry points
            pytest.fail(f"plan failed: {e}")
        a, na = _run(B, B.perform_operation_noalloc, dev, pl, out_dt, rows)
        b, nb = _run(B, B.perform_operation, dev, pl, out_dt, rows)
        assert a.data().tobytes() == b.data().tobytes(), expr
        assert np.array_equal(a.valid(), b.valid()), expr
        assert na == nb, expr
        # and once more after the allocating entry point ran: the scratch still serves
        c, nc = _run(B, B.perform_operation_noalloc, dev, pl, out_dt, rows)
        assert c.data().tobytes() == b.data().tobytes() and nc == nb
        for x in (a, b, c):
            x.free()
        for d in dev:
            d.free()


def test_noalloc_empty_input(B):
    dev = [B.DeviceColumn.from_numpy(np.zeros(0, np.int32))]
    pl = B.plan_expression([P.GDF_INT32], "+($0, 1)")
    out = B.DeviceColumn(P.GDF_INT32, 0, with_valid=True)
    nulls = B.perform_operation_noalloc([out], dev, pl["left"], pl["right"], pl["outputs"], [pl["final_position"]],
                                        pl["binops"], pl["unops"], pl["lscalars"], pl["rscalars"])
    assert nulls == [0]

"""GPU parity at the full sizes of the wide-expression and filter + gather workloads.

The oracle is scalar and cannot evaluate 100 M or 500 M rows in seconds, so these
tests check what holds independently of size:
  * row locality — every expression output row depends on that input row only, so
    the oracle evaluated on a sample of rows (the first and last 64 Ki rows and
    a random million) must equal the GPU's output at those rows, bit-exact;
  * null counts — each output's count equals the zero bits of its bitmap;
  * compaction — with a payload column holding the row index, the filter's
    output must be exactly the indices of the rows whose key passes, in order
    (count, selection and row order at once), and every other payload column's
    output row j must be its input at that index.
Inputs are generated on the device (row i depends only on the seed and i) and
downloaded for the checks.
"""
import numpy as np
import pytest

from oracle import planner as P
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def B():
    import blz_ral
    if blz_ral.device_count() < 1:
        pytest.fail("no HIP device visible")
    return blz_ral


def _say(*a):
    print(*a, flush=True)


def _synth(B, dtype, n, seed, dist, a, b, null_rate=0.0):
    c = B.DeviceColumn(dtype, n, with_valid=null_rate > 0)
    B.synth_column(c, seed, dist, a, b, null_rate)
    if null_rate > 0:
        B.update_null_count(c)
    return c


def _host(c):
    return O.HostColumn(c.data(), c.col.dtype, c.valid())


# 100 M rows x 8 columns (2 int32 uniform [-1e6, 1e6], 2 int64 uniform [0, 2^40],
# 2 float32 normal(0, 1), 2 float64 lognormal(0, 2)), 5 % nulls each; one 24-op
# plan with MOD, CAST int64 -> float64, comparisons, AND / OR / NOT and a nested
# conditional (CASE -> COALESCE); four outputs.
def _c2_project():
    from blz_ral import workloads as W
    return W.WIDE_PROJECT


C2_PROJECT = _c2_project()


@pytest.mark.parametrize("kernel", [0, 4], ids=["compiled", "tile_vm"])
def test_wide_expression_100M_rows(B, kernel):
    n = 100_000_000
    _say("CASE test=wide_expression_100M kernel", kernel, "n", n)
    kinds = [(P.GDF_INT32, 0, -1e6, 1e6), (P.GDF_INT32, 0, -1e6, 1e6), (P.GDF_INT64, 0, 0, 2.0**40),
             (P.GDF_INT64, 0, 0, 2.0**40), (P.GDF_FLOAT32, 2, 0.0, 1.0), (P.GDF_FLOAT32, 2, 0.0, 1.0),
             (P.GDF_FLOAT64, 3, 0.0, 2.0), (P.GDF_FLOAT64, 3, 0.0, 2.0)]
    dev = [_synth(B, dt, n, 7001 + i, dist, a, b, 0.05) for i, (dt, dist, a, b) in enumerate(kinds)]
    plan, _, _ = P.parse_project([c.col.dtype for c in dev], C2_PROJECT)
    assert len(plan.binops) == 24
    prog = B.Program(dev, None, C2_PROJECT, kernel=kernel)
    _say("program kernel", prog.kernel, "R", prog.r, "ops", prog.n_ops)
    assert prog.kernel == (5 if kernel == 0 else 4)  # the compiled tier by default, or the tile VM
    prog.launch()
    _, nulls = prog.result()
    _say("evaluated; checking")
    rng = np.random.default_rng(7100)
    rows = np.unique(np.concatenate([np.arange(65536), np.arange(n - 65536, n),
                                     rng.integers(0, n, 1_000_000)]))
    host = []
    for c in dev:
        h = _host(c)
        host.append(O.HostColumn(h.data[rows], h.dtype, h.valid[rows]))
    ref = O.execute_project(host, C2_PROJECT)
    for k, (g, r, nc) in enumerate(zip(prog.outputs, ref, nulls)):
        gd, gv = g.data(), g.valid()
        assert nc == int(np.count_nonzero(~gv)), f"output {k}: null count"
        assert gd[rows].tobytes() == r.data.tobytes(), f"output {k}: data at the sampled rows"
        assert np.array_equal(gv[rows], r.valid), f"output {k}: validity at the sampled rows"
    prog.close()
    for c in dev:
        c.free()


@pytest.fixture(scope="module")
def c3_frame(B):
    n = 500_000_000
    _say("CASE c3_frame n", n)
    frame = [_synth(B, P.GDF_INT32, n, 7201, 0, 0, 2.0**31 - 1),   # key uniform [0, 2^31)
             _synth(B, P.GDF_INT64, n, 0, 5, 0, 0),                 # row index
             _synth(B, P.GDF_INT64, n, 0, 5, 1e12, 0),              # 10^12 + row index
             _synth(B, P.GDF_INT64, n, 0, 5, -5, 0),                # row index - 5
             _synth(B, P.GDF_FLOAT64, n, 7202, 1, 0.0, 1.0),
             _synth(B, P.GDF_FLOAT64, n, 7203, 1, 0.0, 1.0),
             _synth(B, P.GDF_FLOAT64, n, 7204, 1, 0.0, 1.0)]
    key = frame[0].data()
    f4 = frame[4].data()
    yield frame, key, f4
    for c in frame:
        c.free()


@pytest.mark.parametrize("s", [0.001, 0.1, 0.5, 0.9])
def test_filter_gather_500M_rows(B, c3_frame, s):
    frame, key, f4 = c3_frame
    thr = int(s * 2**31)
    filt = f"LogicalFilter(condition=[<($0, {thr})])"
    proj = "LogicalProject(p1=[$1], p2=[$2], p3=[$3], f1=[$4], f2=[$5], f3=[$6])"
    _say("CASE test=filter_gather_500M s", s)
    prog = B.Program(frame, filt, proj)
    prog.launch()
    kept, _ = prog.result()
    want = np.flatnonzero(key < thr)
    assert kept == len(want)
    out = [c.data() for c in prog.outputs[:4]]
    assert np.array_equal(out[0], want)                        # selection and row order
    assert np.array_equal(out[1] - 10**12, want)
    assert np.array_equal(out[2] + 5, want)
    rng = np.random.default_rng(7300)
    j = rng.integers(0, kept, min(kept, 1_000_000)) if kept else np.zeros(0, np.int64)
    assert out[3][j].tobytes() == f4[want[j]].tobytes()       # the float payload, bit-exact
    prog.close()

"""The rest of the drop-in boundary on the GPU: the reference's interops benchmark
sequence from C++ (parse_project_plan, then perform_operation on the params'
vectors, benchmarks/interops/interops_benchmark.cpp:89-95, and
add_expression_to_plan) through include/blz/ral.hpp, and materialize_column over
index columns of every width (ColumnManipulation.cu:120-134)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import planner as P
from oracle import pyoracle as O
from tests.helpers import gen_column, to_device, ulp_diff
from tests.test_host import INTEROP_PLANS

pytestmark = pytest.mark.gpu

SHIM = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "ral_shim_test")


@pytest.fixture(scope="module")
def B():
    import blz_ral
    if blz_ral.device_count() < 1:
        pytest.fail("no HIP device visible")
    return blz_ral


def _interop_frame(n):
    i = np.arange(n, dtype=np.int64)
    x = ((i * 7919 + 13) % 165191049).astype(np.int32)
    y = (1 + (i * 104729 + 7) % 2147483646).astype(np.float64)
    z = ((i * 40503 + 3) % 1073741823).astype(np.int32)
    return [O.HostColumn(x, P.GDF_INT32), O.HostColumn(y, P.GDF_FLOAT64), O.HostColumn(z, P.GDF_INT32)]


def _parse(lines, dtypes):
    outs = {}
    for ln in lines:
        if not ln.startswith("out "):
            continue
        _, k, o, i, v = ln.split()
        outs.setdefault((int(k), int(o)), []).append(float.fromhex(v) if ("p" in v or "nan" in v or "inf" in v)
                                                     else int(v))
    return outs


def _near_integer(v, ulps=4):
    """|v - round(v)| within `ulps` units in the last place of v."""
    r = np.round(v)
    return np.abs(v - r) <= ulps * np.spacing(np.maximum(np.abs(v), np.abs(r)))


def test_interops_benchmark_sequence_from_cpp(B):
    """The interops benchmark's four plans, parsed by parse_project_plan and run by
    perform_operation from C++, against the oracle. Integer results are bit-exact.
    Products of a sine and a power are within 4 ulp (two library results, each
    within 1 ulp, multiplied). FLOOR / CEIL of sin^2 + cos^2 may land on either side
    of an integer: a row may differ only where the oracle's sum is within 4 ulp of
    that integer (by 2 where both sums are). add_expression_to_plan's plan gives
    evaluate_expression's bytes."""
    n = 3001
    print(f"CASE interops n={n}", flush=True)
    r = subprocess.run([SHIM, "interops", str(n)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    lines = r.stdout.splitlines()
    host = _interop_frame(n)
    got = _parse(lines, None)
    inner = {}
    for name, e in (("sc", "+(POWER(SIN($1), 2), POWER(COS($1), 2))"),
                    ("sc2", "+(POWER(SIN(+($1, 53.42)), 2), POWER(COS(+($1, 53.42)), 2))")):
        inner[name] = O.evaluate_expression(host, e, P.GDF_FLOAT64).data
    flips = 0
    for k, qp in enumerate(INTEROP_PLANS[:4]):
        ref = O.execute_project(host, qp)
        for o, rc in enumerate(ref):
            g = np.array(got[(k, o)], dtype=rc.data.dtype)
            assert len(g) == n
            body = qp.split("=[", o + 1)[o + 1]
            if rc.data.dtype.kind != "f":
                assert g.tobytes() == rc.data.tobytes(), (k, o)
            elif body.startswith("-(FLOOR("):
                diff = np.nonzero(g != rc.data)[0]
                n1, n2 = _near_integer(inner["sc"][diff]), _near_integer(inner["sc2"][diff])
                step = np.abs(g[diff] - rc.data[diff])
                ok = ((step == 1) & (n1 | n2)) | ((step == 2) & n1 & n2)
                assert ok.all(), (k, o, diff[~ok][:5])
                flips += len(diff)
            else:
                d = ulp_diff(g, rc.data)
                print(f"ULP interops plan={k} out={o} max={d.max():g}", flush=True)
                assert (d <= 4).all(), (k, o, float(d.max()))
    print(f"interops FLOOR/CEIL rows on the other side of an integer: {flips}", flush=True)
    line = next(ln for ln in lines if ln.startswith("add_expression_to_plan"))
    assert line.endswith("equal 1"), line
    a = np.array(got[(9, 0)], dtype=np.int64)
    ref = O.evaluate_expression(host, "+(*($0, 3), MOD($2, 7))", P.GDF_INT64)
    assert a.tobytes() == ref.data.tobytes()
    ops = [ln.split()[1:] for ln in lines if ln.startswith("op ")]
    want = P.Plan()
    P.add_expression_to_plan(want, "+(*($0, 3), MOD($2, 7))", 0, 1, 2, [0, -1, 1])
    assert [[int(v) for v in op[1:4]] for op in ops] == [list(t) for t in zip(want.left, want.right, want.outputs)]


@pytest.mark.parametrize("idx_dtype", [P.GDF_INT8, P.GDF_INT16, P.GDF_INT32, P.GDF_INT64])
@pytest.mark.parametrize("dtype", [P.GDF_INT8, P.GDF_INT32, P.GDF_FLOAT64])
def test_materialize_column_index_widths(B, dtype, idx_dtype):
    """Index columns of 1, 2, 4 and 8 bytes, read as signed integers of their width:
    negative and past-the-end indices leave the row untouched; the validity bits
    and null count follow the gathered rows."""
    rng = np.random.default_rng(100 * dtype + idx_dtype)
    info = np.iinfo(O.NP_OF[idx_dtype])
    n = min(int(info.max), 20_000)
    m = 9_001
    col = gen_column(rng, dtype, n, 0.2)
    idx = rng.integers(max(int(info.min), -5), n + 3, m).astype(np.int64)
    if idx_dtype == P.GDF_INT64:
        idx[:4] = [2 ** 40, -(2 ** 40), 2 ** 31, n - 1]          # past the end beyond 32 bits
    idx = idx.astype(O.NP_OF[idx_dtype])
    print(f"CASE gather_index_width dtype={dtype} idx={idx_dtype}", flush=True)
    dcol = to_device(B, col)
    didx = B.DeviceColumn.from_numpy(idx, idx_dtype)
    out = B.DeviceColumn(dtype, m, with_valid=True)
    B.materialize_column(dcol, out, didx)
    j = idx.astype(np.int64)
    ok = (j >= 0) & (j < n)
    got = out.data()
    assert np.array_equal(got[ok], col.data[j[ok]])
    want_valid = np.zeros(m, bool)
    want_valid[ok] = col.valid[j[ok]]
    assert np.array_equal(out.valid(), want_valid)
    assert out.null_count == int(np.count_nonzero(~want_valid))
    # the int32 path against the oracle's gather, on the indices that fit
    ref = O.gather(col, np.where(ok, j, -1).astype(np.int32), True)
    assert np.array_equal(out.valid(), ref.valid) and out.null_count == ref.null_count

"""The oracle against closed forms and library routines (CPU only).

The reference's own tests (golden vectors, test_oracle_golden.py) leave some of the
path's functions without an expected value: MOD and integer division, POWER,
COALESCE, IS [NOT] NULL on data, NOT, ABS, LOG10, the int64 -> binary64 CAST
extension, the SQL date parts, validity propagation and stencil -> gather. Each
is checked here against what the mathematics or a standard routine fixes,
independently of the oracle's code: C truncating division and remainder, exact
integer powers with int64 wrap-around, numpy.where on the validity, Python's
proleptic Gregorian calendar, the IEEE-754 binary64 conversion of an integer, and
brute-force bit loops.
"""
import datetime
import math

import numpy as np
import pytest

from oracle import planner as P
from oracle import pyoracle as O

N = 2000


def _rng(seed):
    return np.random.default_rng(90_000 + seed)


def _col(data, dtype, valid=None):
    return O.HostColumn(np.ascontiguousarray(data), dtype, valid)


def _c_trunc_div(a, b):
    """C99 a / b for int64 (truncation toward zero); 0 where the reference is
    undefined (b == 0, INT64_MIN / -1)."""
    out = []
    for x, y in zip(a.tolist(), b.tolist()):
        if y == 0 or (x == -2**63 and y == -1):
            out.append(0)
            continue
        q = abs(x) // abs(y)
        out.append(q if (x >= 0) == (y > 0) else -q)
    return np.array(out, dtype=np.int64)


def _c_rem(a, b):
    """C99 a % b for int64 (sign of the dividend)."""
    q = _c_trunc_div(a, b)
    return np.array([0 if y == 0 or (x == -2**63 and y == -1) else x - qq * y
                     for x, y, qq in zip(a.tolist(), b.tolist(), q.tolist())], dtype=np.int64)


def test_integer_division_and_modulo_truncate_like_c():
    rng = _rng(1)
    a = rng.integers(-10**12, 10**12, N).astype(np.int64)
    b = rng.integers(-1000, 1000, N).astype(np.int64)
    b[:10] = 0                       # defined as 0
    a[10], b[10] = -2**63, -1        # defined as 0
    frame = [_col(a, P.GDF_INT64), _col(b, P.GDF_INT64)]
    q = O.evaluate_expression(frame, "/($0, $1)", P.GDF_INT64)
    r = O.evaluate_expression(frame, "MOD($0, $1)", P.GDF_INT64)
    assert np.array_equal(q.data, _c_trunc_div(a, b))
    assert np.array_equal(r.data, _c_rem(a, b))
    # the identity the two must satisfy wherever both are defined
    ok = (b != 0) & ~((a == -2**63) & (b == -1))
    assert np.array_equal(q.data[ok] * b[ok] + r.data[ok], a[ok])


def test_float_division_and_modulo_of_floats():
    rng = _rng(2)
    x = rng.uniform(-1e6, 1e6, N)
    y = rng.uniform(-1e3, 1e3, N)
    frame = [_col(x, P.GDF_FLOAT64), _col(y, P.GDF_FLOAT64)]
    q = O.evaluate_expression(frame, "/($0, $1)", P.GDF_FLOAT64)
    assert q.data.tobytes() == (x / y).tobytes()       # one IEEE division
    # MOD truncates both operands to int64 first (interpreter_ops.cuh:650-656)
    r = O.evaluate_expression(frame, "MOD($0, $1)", P.GDF_FLOAT64)
    want = _c_rem(np.trunc(x).astype(np.int64), np.trunc(y).astype(np.int64)).astype(np.float64)
    assert r.data.tobytes() == want.tobytes()


def test_power_integer_loop_and_float_pow():
    rng = _rng(3)
    base = rng.integers(-40, 40, N).astype(np.int64)
    exp = rng.integers(0, 14, N).astype(np.int64)
    frame = [_col(base, P.GDF_INT64), _col(exp, P.GDF_INT64)]
    got = O.evaluate_expression(frame, "POWER($0, $1)", P.GDF_INT64)
    # d = 1; for i < R: d *= L, in int64 (wrap-around): the exact power mod 2^64
    want = np.array([((int(b) ** int(e) + 2**63) % 2**64) - 2**63 for b, e in zip(base, exp)], dtype=np.int64)
    assert np.array_equal(got.data, want)
    xf = rng.uniform(0.1, 10.0, N)
    frame = [_col(xf, P.GDF_FLOAT64), _col(exp.astype(np.float64), P.GDF_FLOAT64)]
    got = O.evaluate_expression(frame, "POWER($0, $1)", P.GDF_FLOAT64)
    want = np.array([math.pow(a, b) for a, b in zip(xf.tolist(), exp.astype(float).tolist())])
    assert got.data.tobytes() == want.tobytes()


def test_coalesce_is_null_and_validity_propagation():
    rng = _rng(4)
    x = rng.integers(-100, 100, N).astype(np.int64)
    y = rng.integers(-100, 100, N).astype(np.int64)
    vx = rng.random(N) > 0.3
    vy = rng.random(N) > 0.3
    frame = [_col(x, P.GDF_INT64, vx), _col(y, P.GDF_INT64, vy)]
    c = O.evaluate_expression(frame, "CASE(IS NOT NULL($0), $0, $1)", P.GDF_INT64)
    assert np.array_equal(c.data, np.where(vx, x, y))
    assert np.array_equal(c.valid, vx | vy)                     # COALESCE: OR of the words
    s = O.evaluate_expression(frame, "+($0, $1)", P.GDF_INT64)
    assert np.array_equal(s.data, x + y)
    assert np.array_equal(s.valid, vx & vy)                     # binary op: AND of the words
    assert s.null_count == int(np.count_nonzero(~(vx & vy)))
    isn = O.evaluate_expression(frame, "IS NULL($0)", P.GDF_INT8)
    isnn = O.evaluate_expression(frame, "IS NOT NULL($0)", P.GDF_INT8)
    assert np.array_equal(isn.data, (~vx).astype(np.int8))
    assert np.array_equal(isnn.data, vx.astype(np.int8))


def test_not_abs_log10():
    rng = _rng(5)
    i = rng.integers(-3, 3, N).astype(np.int64)
    f = rng.uniform(-50.0, 50.0, N)
    f[:5] = [0.0, -0.0, 1e-300, -1e300, 2.5]
    frame = [_col(i, P.GDF_INT64), _col(f, P.GDF_FLOAT64)]
    n_i = O.evaluate_expression(frame, "NOT($0)", P.GDF_INT64)
    assert np.array_equal(n_i.data, (i == 0).astype(np.int64))
    n_f = O.evaluate_expression(frame, "NOT($1)", P.GDF_FLOAT64)
    assert np.array_equal(n_f.data, (f == 0).astype(np.float64))
    a_f = O.evaluate_expression(frame, "ABS($1)", P.GDF_FLOAT64)
    assert a_f.data.tobytes() == np.fabs(f).tobytes()
    a_i = O.evaluate_expression(frame, "ABS($0)", P.GDF_INT64)   # through binary64 (exact here)
    assert np.array_equal(a_i.data, np.abs(i))
    pos = np.abs(f) + 1.0
    frame = [_col(pos, P.GDF_FLOAT64)]
    lg = O.evaluate_expression(frame, "LOG10($0)", P.GDF_FLOAT64)
    assert lg.data.tobytes() == np.array([math.log10(v) for v in pos.tolist()]).tobytes()


def test_cast_int64_to_binary64_rounds_to_nearest_even():
    rng = _rng(6)
    v = rng.integers(-2**62, 2**62, N).astype(np.int64)
    v[:6] = [0, 1, -1, 2**53 + 1, 2**53 + 3, -(2**53 + 1)]    # halfway cases round to even
    frame = [_col(v, P.GDF_INT64)]
    got = O.evaluate_expression(frame, "BL_CAST_DOUBLE($0)", P.GDF_FLOAT64)
    want = np.array([float(x) for x in v.tolist()])           # Python int -> float is RNE
    assert got.data.tobytes() == want.tobytes()


@pytest.mark.parametrize("part", ["YEAR", "MONTH", "DAY"])
def test_date_parts_of_date32_follow_the_gregorian_calendar(part):
    rng = _rng(7)
    days = rng.integers(-200_000, 200_000, N).astype(np.int32)   # ~ years 1422 .. 2517
    days[:4] = [0, -1, 59, 11016]                               # 1970-01-01, 1969-12-31, 1970-03-01, 2000-02-29
    frame = [_col(days, P.GDF_DATE32)]
    got = O.evaluate_expression(frame, f"EXTRACT(FLAG({part}), $0)", P.GDF_INT16)
    epoch = datetime.date(1970, 1, 1)
    want = [getattr(epoch + datetime.timedelta(days=int(d)), part.lower()) for d in days]
    assert np.array_equal(got.data.astype(np.int64), np.array(want, dtype=np.int64))


@pytest.mark.parametrize("part", ["YEAR", "MONTH", "DAY", "HOUR", "MINUTE", "SECOND"])
def test_date_parts_of_date64_milliseconds(part):
    rng = _rng(8)
    ms = rng.integers(-5 * 10**12, 5 * 10**12, N).astype(np.int64)
    ms[:3] = [0, -1, 86_399_999]
    frame = [_col(ms, P.GDF_DATE64)]
    got = O.evaluate_expression(frame, f"EXTRACT(FLAG({part}), $0)", P.GDF_INT16)
    epoch = datetime.datetime(1970, 1, 1)
    want = [getattr(epoch + datetime.timedelta(milliseconds=int(t)), part.lower()) for t in ms]
    assert np.array_equal(got.data.astype(np.int64), np.array(want, dtype=np.int64))


def test_stencil_to_gather_is_boolean_indexing_in_row_order():
    rng = _rng(9)
    key = rng.integers(-5, 5, N).astype(np.int8)
    vk = rng.random(N) > 0.2
    pay = rng.uniform(0, 1, N)
    vp = rng.random(N) > 0.1
    frame = [_col(key, P.GDF_INT8, vk), _col(pay, P.GDF_FLOAT64, vp)]
    out = O.process_filter(frame, "LogicalFilter(condition=[$0])")   # bare column: non-zero and valid
    keep = (key != 0) & vk
    assert np.array_equal(out[1].data, pay[keep])
    assert np.array_equal(out[1].valid, vp[keep])
    assert out[1].null_count == int(np.count_nonzero(~vp[keep]))
    out = O.process_filter(frame, "LogicalFilter(condition=[>($1, 0.5)])")
    keep = (pay > 0.5) & vp
    assert np.array_equal(out[0].data, key[keep])
    assert np.array_equal(out[0].valid, vk[keep])

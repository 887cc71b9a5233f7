"""GPU parity for the operator table's out-of-line arm (interpreter_ops.cuh:657-676
POW, :753-779 FLOOR ... LOG10, :52-153 date parts): SIN, COS, TAN, ASIN, ACOS,
ATAN, LN, LOG10, COTAN, POWER (binary64 and the integer loop) and YEAR ... SECOND
on DATE32 / DATE64 / TIMESTAMP, through the tile VM, the direct register-file
kernel, the gather of a filter's kept rows and the register-plan ABI, against the
oracle on the same seeded inputs.

Integer results (POWER of integers, date parts) are bit-exact. Floating-point
results are within 1 ulp of the oracle (the device evaluates the functions with
the ROCm device math library, the oracle with the host C library), NaN exactly
where the oracle has NaN, infinities and signed zeros equal. Every test prints one
line per op with its largest ulp distance and the number of rows that differ.
"""
import math

import numpy as np
import pytest

from oracle import planner as P
from oracle import pyoracle as O
from tests.helpers import assert_column_equal, to_device, ulp_diff

pytestmark = pytest.mark.gpu

ULP_TOL = 1   # transcendental and POWER results: at most 1 ulp from the oracle
# COTAN is cos(x) / sin(x) (interpreter_ops.cuh:765-767): a quotient of two library
# results, each within 1 ulp between the two libraries, rounded once more: up to
# 2 ulp of relative difference, which is 4 ulp of the quotient where the quotient
# sits at the bottom of its binade and the operands at the top of theirs
ULP_TOL_OP = {"COTAN": 4}


@pytest.fixture(scope="module")
def B():
    import blz_ral
    if blz_ral.device_count() < 1:
        pytest.fail("no HIP device visible")
    return blz_ral


def _print_case(**kw):
    print("CASE", " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


F64_SPECIALS = [0.0, -0.0, np.nan, np.inf, -np.inf, 1.0, -1.0, 0.5, -0.5, np.pi / 2, np.pi, -np.pi / 4,
                1e-8, -1e-8, 1e22, -1e22, 1e300, -1e300, 5e-324, -5e-324, 2.2250738585072014e-308,
                1.7976931348623157e308, 0.9999999999999999, -0.9999999999999999, 1.0000000000000002,
                710.0, -745.0, 100.0, 10.0, 0.1, 3.0, 1e-3, 1e3]


def _float_inputs(rng, n, np_dtype):
    """Values spread over the functions' domains: [-10, 10], around [-1, 1] (ASIN /
    ACOS and outside), magnitudes from the denormals to the largest finite, and the
    special values (signed zeros, NaN, infinities, pi/2, huge arguments)."""
    k = n // 8
    parts = [rng.uniform(-10.0, 10.0, n - 3 * k - len(F64_SPECIALS)),
             rng.uniform(-1.2, 1.2, k),
             np.abs(rng.normal(0.0, 1.0, k)),
             rng.choice([-1.0, 1.0], k) * 10.0 ** rng.uniform(-320.0, 308.0, k),
             np.array(F64_SPECIALS)]
    x = np.concatenate(parts)
    with np.errstate(over="ignore", under="ignore"):
        x = x.astype(np_dtype)
    return x[rng.permutation(len(x))]


def _int_inputs(rng, n, np_dtype):
    info = np.iinfo(np_dtype)
    x = rng.integers(-1000, 1001, n).astype(np.int64)
    x[:8] = [0, 1, -1, info.max, info.min, info.max - 1, 2 ** 20, -(2 ** 20)]
    if np_dtype == np.int64:
        x[8:400] = rng.integers(-(2 ** 62), 2 ** 62, 392)
    return x.astype(np_dtype)


def _exact(op, x, y=None):
    """The function's exact value (60 significant digits, Python's decimal module)
    where decimal computes it: LN / LOG10 of a positive finite x, POWER of a positive
    finite base or a negative one with an integral exponent; else None."""
    from decimal import Decimal, localcontext
    x = float(x)
    if not math.isfinite(x) or (y is not None and not math.isfinite(float(y))):
        return None
    with localcontext() as ctx:
        ctx.prec = 60
        if op in ("LN", "LOG10", "LOG"):
            if x <= 0:
                return None
            return Decimal(x).ln() if op == "LN" else Decimal(x).log10()
        if op == "POWER":
            y = float(y)
            if x == 0:
                return None
            if x < 0:
                if y != math.floor(y):
                    return None
                r = Decimal(-x) ** Decimal(y)
                return -r if int(y) % 2 else r
            return Decimal(x) ** Decimal(y)
    return None


def _check_float(got, ref, what, dtype_name, tol=ULP_TOL, op=None, x=None, y=None):
    """At most `tol` ulp from the oracle; where the two libraries differ by more,
    and the function's exact value is computable (LN, LOG10, POWER), the device's
    result must be within 1 ulp of the exact value (the host library is then the
    less accurate of the two: the C library's log10 and pow are documented to
    2 ulp)."""
    from decimal import Decimal
    d = ulp_diff(got, ref)
    worst = float(d.max()) if len(d) else 0.0
    n_diff = int(np.count_nonzero(d))
    over = np.nonzero(d > tol)[0]
    worst_exact = 0.0
    for i in over.tolist():
        e = _exact(op, x[i], None if y is None else y[i]) if (op is not None and x is not None) else None
        if e is None:
            raise AssertionError(f"{what}: {d[i]:g} ulp at row {i}: gpu={got[i]!r} oracle={ref[i]!r}")
        g = got[i]
        err = abs(Decimal(float(g)) - e) / Decimal(float(np.spacing(np.abs(g))))
        worst_exact = max(worst_exact, float(err))
        if err > 1:
            raise AssertionError(f"{what}: {d[i]:g} ulp from the oracle and {float(err):.3f} ulp from the exact "
                                 f"value at row {i}: gpu={got[i]!r} oracle={ref[i]!r}")
    print(f"ULP {what} dtype={dtype_name} max={worst:g} rows_differing={n_diff} of {len(d)} "
          f"rows_over_{tol}={len(over)} (device within {worst_exact:.3f} ulp of exact there)", flush=True)


def _run_program(B, host, filt, proj, kernel):
    dev = [to_device(B, h) for h in host]
    prog = B.Program(dev, filt, proj, kernel=kernel)
    assert prog.kernel == kernel
    prog.launch()
    kept, nulls = prog.result()
    outs = []
    for g, nc in zip(prog.outputs, nulls):
        g.col.null_count = nc
        outs.append((g.data(), g.valid(), nc))
    prog.close()
    for d in dev:
        d.free()
    return kept, outs


TRIG = ["SIN", "COS", "TAN", "ASIN", "ACOS", "ATAN", "LN", "LOG10"]
KEY = "LogicalFilter(condition=[<($4, 700)])"   # keeps about 70% of the rows


def _trig_frame(rng, n):
    return [O.HostColumn(_float_inputs(rng, n, np.float64), P.GDF_FLOAT64, rng.random(n) > 0.05),
            O.HostColumn(_float_inputs(rng, n, np.float32), P.GDF_FLOAT32),
            O.HostColumn(_int_inputs(rng, n, np.int32), P.GDF_INT32),
            O.HostColumn(_int_inputs(rng, n, np.int64), P.GDF_INT64, rng.random(n) > 0.1),
            O.HostColumn(rng.integers(0, 1000, n).astype(np.int32), P.GDF_INT32)]


@pytest.mark.parametrize("mode", ["project", "filter"])
@pytest.mark.parametrize("kernel", [4, 3, 5])
@pytest.mark.parametrize("col", [0, 1, 2, 3], ids=["f64", "f32", "i32", "i64"])
def test_transcendentals_within_one_ulp(B, col, kernel, mode):
    """SIN ... LOG10 of a column (binary64 through the register slot, narrowed to the
    output dtype: the input's float type, FLOAT64 for integers), projected over every
    row or over the kept rows of a filter."""
    rng = np.random.default_rng(11_000 + col)
    n = 24_011
    host = _trig_frame(rng, n)
    proj = "LogicalProject(" + ", ".join(f"{f}=[{f}(${col})]" for f in TRIG) + ")"
    filt = KEY if mode == "filter" else None
    _print_case(test="transcendentals", col=col, kernel=kernel, mode=mode, n=n)
    mid = O.process_filter(host, filt) if filt else host
    ref = O.execute_project(mid, proj)
    kept, outs = _run_program(B, host, filt, proj, kernel)
    assert kept == len(ref[0])
    xin = mid[col].data.astype(np.float64)
    for f, (gd, gv, nc), r in zip(TRIG, outs, ref):
        assert gd.dtype == r.data.dtype
        _check_float(gd, r.data, f"{f} kernel={kernel} {mode}", str(gd.dtype), op=f, x=xin)
        rv = r.valid if r.valid is not None else np.ones(len(r.data), bool)
        assert np.array_equal(gv, rv), f"{f}: validity differs"
        assert nc == r.null_count


def _pow_frame(rng, n):
    base = rng.uniform(-10.0, 10.0, n)
    base[: n // 4] = np.round(base[: n // 4])                 # integral bases (negative ones too)
    expo = rng.uniform(-5.0, 5.0, n)
    expo[: n // 2] = np.round(expo[: n // 2])                 # integral exponents
    sp_b = [0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 2.0, -2.0, 1e-300, 1e300, 0.5, 10.0]
    sp_e = [0.0, -0.0, 1.0, -1.0, 0.5, -0.5, np.inf, -np.inf, np.nan, 400.0, -400.0, 1100.0, 3.0]
    m = len(sp_b) * len(sp_e)
    bb, ee = np.meshgrid(sp_b, sp_e)
    base[-m:] = bb.ravel()
    expo[-m:] = ee.ravel()
    ib = rng.integers(-5, 6, n).astype(np.int64)
    ie = rng.integers(-3, 71, n).astype(np.int64)
    ib[:6] = [0, 1, -1, 2 ** 31 - 1, -(2 ** 31), 3]
    ie[:6] = [0, 5, 63, 2, 3, 40]
    with np.errstate(over="ignore"):
        b32, e32 = base.astype(np.float32), expo.astype(np.float32)
    return [O.HostColumn(base, P.GDF_FLOAT64, rng.random(n) > 0.05), O.HostColumn(expo, P.GDF_FLOAT64),
            O.HostColumn(b32, P.GDF_FLOAT32), O.HostColumn(e32, P.GDF_FLOAT32),
            O.HostColumn(ib.astype(np.int32), P.GDF_INT32), O.HostColumn(ie, P.GDF_INT64, rng.random(n) > 0.1),
            O.HostColumn(rng.integers(0, 1000, n).astype(np.int32), P.GDF_INT32)]


POW_FLOAT = ["POWER($0, $1)", "POWER($2, $3)", "POWER($0, 2)", "POWER($0, 0.5)", "POWER($1, $4)",
             "POWER($2, 3)"]
POW_INT = ["POWER($4, $5)", "POWER($5, 3)", "POWER($4, 2)", "POWER($5, $4)"]


@pytest.mark.parametrize("mode", ["project", "filter"])
@pytest.mark.parametrize("kernel", [4, 3, 5])
def test_power(B, kernel, mode):
    """POWER: pow((double)L, (double)R) when an operand is floating point (within 1
    ulp), else the reference's loop `d = 1; for (i < R) d *= L` in wrapping int64
    (bit-exact), with negative, zero and large exponents, negative bases, zeros,
    infinities and NaN."""
    rng = np.random.default_rng(12_000)
    n = 20_000
    host = _pow_frame(rng, n)
    exprs = POW_FLOAT + POW_INT
    proj = "LogicalProject(" + ", ".join(f"P{i}=[{e}]" for i, e in enumerate(exprs)) + ")"
    filt = "LogicalFilter(condition=[<($6, 700)])" if mode == "filter" else None
    _print_case(test="power", kernel=kernel, mode=mode, n=n)
    mid = O.process_filter(host, filt) if filt else host
    ref = O.execute_project(mid, proj)
    kept, outs = _run_program(B, host, filt, proj, kernel)
    assert kept == len(ref[0])
    for e, (gd, gv, nc), r in zip(exprs, outs, ref):
        assert gd.dtype == r.data.dtype, e
        if e in POW_INT:
            assert gd.dtype == np.int64
            assert gd.tobytes() == r.data.tobytes(), f"{e}: integer POWER differs"
        else:
            a, b = [t.strip() for t in e[len("POWER("):-1].split(",")]
            xv, yv = [mid[int(t[1:])].data.astype(np.float64) if t.startswith("$") else
                      np.full(len(r.data), float(t)) for t in (a, b)]
            _check_float(gd, r.data, f"{e} kernel={kernel} {mode}", str(gd.dtype), op="POWER", x=xv, y=yv)
        rv = r.valid if r.valid is not None else np.ones(len(r.data), bool)
        assert np.array_equal(gv, rv) and nc == r.null_count, f"{e}: validity differs"


PARTS = ["YEAR", "MONTH", "DAY", "HOUR", "MINUTE", "SECOND"]
MS_DAY, MS_HOUR, MS_MIN = 86_400_000, 3_600_000, 60_000


def _date_frame(rng, n):
    days = rng.integers(-800_000, 800_000, n).astype(np.int64)
    days[:10] = [0, -1, 59, 11016, -719468, -719469, 2 ** 31 - 1, -(2 ** 31), 2 ** 31 - 719469, 2932896]
    ms = rng.integers(-10 ** 14, 10 ** 14, n).astype(np.int64)
    # exact negative multiples of a day / hour / minute / second (the reference's
    # HOUR / MINUTE / SECOND give 24 / 60 / 60 there), and their neighbours
    edge = []
    for unit in (MS_DAY, MS_HOUR, MS_MIN, 1000):
        for k in (1, 2, 7, 365, 10 ** 6):
            edge += [-k * unit, -k * unit - 1, -k * unit + 1, k * unit]
    edge += [0, -1, 1, 86_399_999, -86_399_999, 9 * 10 ** 18, -9 * 10 ** 18]
    ms[10:10 + len(edge)] = edge
    ts = ms.copy()
    ts[len(edge) + 10:] = rng.integers(-(2 ** 50), 2 ** 50, n - len(edge) - 10)
    return [O.HostColumn(days.astype(np.int32), P.GDF_DATE32, rng.random(n) > 0.05),
            O.HostColumn(ms, P.GDF_DATE64), O.HostColumn(ts, P.GDF_TIMESTAMP, rng.random(n) > 0.1),
            O.HostColumn(rng.integers(0, 1000, n).astype(np.int32), P.GDF_INT32)]


@pytest.mark.parametrize("mode", ["project", "filter"])
@pytest.mark.parametrize("kernel", [4, 3, 5])
@pytest.mark.parametrize("col", [0, 1, 2], ids=["date32", "date64", "timestamp"])
def test_date_parts(B, col, kernel, mode):
    """EXTRACT(FLAG(part), $k) for every part over DATE32 (days; HOUR / MINUTE /
    SECOND are 0), DATE64 and TIMESTAMP (milliseconds, negative ones included, with
    the exact negative multiples), bit-exact as INT16."""
    rng = np.random.default_rng(13_000 + col)
    n = 30_000
    host = _date_frame(rng, n)
    proj = "LogicalProject(" + ", ".join(f"{p}=[EXTRACT(FLAG({p}), ${col})]" for p in PARTS) + ")"
    filt = "LogicalFilter(condition=[<($3, 700)])" if mode == "filter" else None
    _print_case(test="date_parts", col=col, kernel=kernel, mode=mode, n=n)
    mid = O.process_filter(host, filt) if filt else host
    ref = O.execute_project(mid, proj)
    kept, outs = _run_program(B, host, filt, proj, kernel)
    assert kept == len(ref[0])
    for p, (gd, gv, nc), r in zip(PARTS, outs, ref):
        assert gd.dtype == np.int16 and r.data.dtype == np.int16
        assert gd.tobytes() == r.data.tobytes(), f"{p}: date part differs"
        rv = r.valid if r.valid is not None else np.ones(len(r.data), bool)
        assert np.array_equal(gv, rv) and nc == r.null_count


def test_date_parts_in_a_predicate(B):
    """Date parts compared inside a filter predicate (the kept rows must be the
    oracle's, in row order)."""
    rng = np.random.default_rng(13_100)
    n = 50_000
    host = _date_frame(rng, n)
    filt = "LogicalFilter(condition=[AND(=(EXTRACT(FLAG(MONTH), $0), 2), >=(EXTRACT(FLAG(HOUR), $1), 12))])"
    proj = "LogicalProject(D=[$0], Y=[EXTRACT(FLAG(YEAR), $2)], S=[EXTRACT(FLAG(SECOND), $1)])"
    _print_case(test="date_predicate", n=n)
    ref = O.execute_project(O.process_filter(host, filt), proj)
    for kernel in (0, 3, 5):
        kept, outs = _run_program(B, host, filt, proj, kernel if kernel else 4)
        assert kept == len(ref[0])
        for (gd, gv, nc), r in zip(outs, ref):
            assert gd.tobytes() == r.data.tobytes()


UNARY_PLAN_OPS = ["SIN", "COS", "TAN", "COTAN", "ASIN", "ACOS", "ATAN", "LN", "LOG"]


@pytest.mark.parametrize("dtype", [P.GDF_FLOAT64, P.GDF_FLOAT32], ids=["f64", "f32"])
def test_unary_register_plans_within_one_ulp(B, dtype):
    """Hand-built one-op register plans through perform_operation
    (interpreter_cpp.h:25-37), COTAN included (it has no Calcite token: cos / sin,
    interpreter_ops.cuh:765-767)."""
    rng = np.random.default_rng(14_000 + dtype)
    n = 16_384 + 17
    npd = O.NP_OF[dtype]
    x = _float_inputs(rng, n, npd)
    valid = rng.random(n) > 0.05
    host = [O.HostColumn(x, dtype, valid)]
    dev = [to_device(B, h) for h in host]
    dummy = B.make_scalar(B.GDF_INT32, 0)
    for name in UNARY_PLAN_OPS:
        u = getattr(P, "BLZ_" + name)
        plan = P.Plan(left=[0], right=[-1], outputs=[1], binops=[P.GDF_INVALID_BINARY], unops=[u],
                      lscalars=[P.DUMMY], rscalars=[P.DUMMY], final_positions=[1])
        _print_case(test="unary_plan", op=name, dtype=dtype, n=n)
        ref = O.perform_operation(host, plan, [dtype], [True])[0]
        out = B.DeviceColumn(dtype, n, with_valid=True)
        nulls = B.perform_operation([out], dev, [0], [-1], [1], [1], [B.GDF_INVALID_BINARY], [u],
                                    [dummy], [dummy])
        _check_float(out.data(), ref.data, f"{name} plan", str(np.dtype(npd)), ULP_TOL_OP.get(name, ULP_TOL),
                     op=name, x=x.astype(np.float64))
        assert np.array_equal(out.valid(), ref.valid) and nulls[0] == ref.null_count
        out.free()


@pytest.mark.parametrize("scalar_mode", [0, 1], ids=["intended", "shipped"])
def test_register_plan_float_scalars_and_slot_reuse(B, scalar_mode):
    """A hand-built plan with float and integer scalars on either side, mixed
    int32 / float32 / float64 inputs, temporaries reused across ops and two outputs
    narrowed to different dtypes, in both scalar modes (interpreter_ops.cuh:156-175:
    SHIPPED reads the int64 cast of a float literal as a double)."""
    rng = np.random.default_rng(14_500 + scalar_mode)
    n = 50_001
    a = rng.integers(-100, 100, n).astype(np.int32)
    b = rng.uniform(-3.0, 3.0, n).astype(np.float32)
    c = rng.uniform(-50.0, 50.0, n)
    host = [O.HostColumn(a, P.GDF_INT32), O.HostColumn(b, P.GDF_FLOAT32, rng.random(n) > 0.1),
            O.HostColumn(c, P.GDF_FLOAT64)]
    # registers: inputs 0..2, outputs 3..4, temporaries from 5; the transcendental is
    # the last op of its output (a 1-ulp difference before a cancelling subtraction
    # would grow without bound)
    # t5 = c * 2.5 ; t6 = t5 + b ; t5 = t6 - a ; out3 = SIN(t5) ; t6 = 7 - a ; out4 = t6 * 3
    f25 = P.Scalar(P.GDF_FLOAT64, 2.5, True)
    i7 = P.Scalar(P.GDF_INT64, 7, True)
    i3 = P.Scalar(P.GDF_INT64, 3, True)
    plan = P.Plan(left=[2, 5, 6, 5, -2, 6], right=[-2, 1, 0, -1, 0, -2], outputs=[5, 6, 5, 3, 6, 4],
                  binops=[P.GDF_MUL, P.GDF_ADD, P.GDF_SUB, P.GDF_INVALID_BINARY, P.GDF_SUB, P.GDF_MUL],
                  unops=[P.BLZ_INVALID_UNARY] * 3 + [P.BLZ_SIN] + [P.BLZ_INVALID_UNARY] * 2,
                  lscalars=[P.DUMMY, P.DUMMY, P.DUMMY, P.DUMMY, i7, P.DUMMY],
                  rscalars=[f25, P.DUMMY, P.DUMMY, P.DUMMY, P.DUMMY, i3], final_positions=[3, 4])
    out_dt = [P.GDF_FLOAT64, P.GDF_INT32]
    _print_case(test="register_plan_scalars", mode=scalar_mode, n=n)
    ref = O.perform_operation(host, plan, out_dt, [True, True], scalar_mode)
    dev = [to_device(B, h) for h in host]
    outs = [B.DeviceColumn(dt, n, with_valid=True) for dt in out_dt]

    def sc(s):
        return B.make_scalar(s.dtype, s.value, s.is_valid) if s is not P.DUMMY else B.make_scalar(B.GDF_INT32, 0)

    nulls = B.perform_operation(outs, dev, plan.left, plan.right, plan.outputs, plan.final_positions,
                                plan.binops, plan.unops, [sc(s) for s in plan.lscalars],
                                [sc(s) for s in plan.rscalars], scalar_mode=scalar_mode)
    _check_float(outs[0].data(), ref[0].data, f"sin plan mode={scalar_mode}", "float64")
    assert outs[1].data().tobytes() == ref[1].data.tobytes()
    for o, r, nc in zip(outs, ref, nulls):
        assert np.array_equal(o.valid(), r.valid) and nc == r.null_count

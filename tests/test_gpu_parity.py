"""GPU parity: the HIP path through the C ABI against the oracle, bit-exact.

Every test here runs on an MI355X (marker `gpu`). Inputs are seeded; sizes keep
the oracle within seconds.
"""
import numpy as np
import pytest

from oracle import planner as P
from oracle import pyoracle as O
from tests import exprgen
from tests.helpers import (C1_FILTER, C1_PROJECT, assert_column_equal, c1_frame, gen_column,
                           to_device)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def B():
    import blz_ral
    if blz_ral.device_count() < 1:
        pytest.fail("no HIP device visible")
    return blz_ral


def _print_case(**kw):
    print("CASE", " ".join(f"{k}={v}" for k, v in kw.items()), flush=True)


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 1000, 4097, 1_000_000])
def test_c1_filter_project(B, n):
    _print_case(test="c1", n=n)
    host = c1_frame(n, seed=11)
    dev = [to_device(B, h) for h in host]
    outs, kept = B.filter_project(dev, C1_FILTER, C1_PROJECT)
    ref = O.execute_project(O.process_filter(host, C1_FILTER), C1_PROJECT)
    assert kept == len(ref[0])
    assert len(outs) == 2
    for g, r in zip(outs, ref):
        assert_column_equal(g, r, what="c1")


def test_c1_shipped_scalar_mode(B):
    n = 200_000
    _print_case(test="c1_shipped", n=n)
    host = c1_frame(n, seed=12)
    dev = [to_device(B, h) for h in host]
    outs, kept = B.filter_project(dev, C1_FILTER, C1_PROJECT, scalar_mode=B.SCALAR_SHIPPED)
    ref = O.execute_project(O.process_filter(host, C1_FILTER, 1), C1_PROJECT, 1)
    assert kept == len(ref[0])
    for g, r in zip(outs, ref):
        assert_column_equal(g, r, what="c1 shipped")


def test_process_filter_then_project_equals_fused(B):
    n = 300_001
    _print_case(test="filter_then_project", n=n)
    host = c1_frame(n, seed=13)
    dev = [to_device(B, h) for h in host]
    filtered = B.process_filter(dev, C1_FILTER)
    ref_f = O.process_filter(host, C1_FILTER)
    for g, r in zip(filtered, ref_f):
        assert_column_equal(g, r, what="filter")
    proj = B.execute_project(filtered, C1_PROJECT)
    fused, _ = B.filter_project(dev, C1_FILTER, C1_PROJECT)
    for g, f in zip(proj, fused):
        assert g.data().tobytes() == f.data().tobytes()


def test_interpreter_golden_test00(B):
    """interpreter_tests.cu TEST_00 (:68-156): ($0+$1)*$2+$1 and $1+2 on int32."""
    _print_case(test="golden_test00")
    i = np.arange(10)
    ins = [B.DeviceColumn.from_numpy((i * 2).astype(np.int32)),
           B.DeviceColumn.from_numpy((i * 10).astype(np.int32)),
           B.DeviceColumn.from_numpy((i * 20).astype(np.int32))]
    outs = [B.DeviceColumn.from_numpy(np.zeros(10, np.int32)) for _ in range(2)]
    junk = B.make_scalar(B.GDF_INT32, 0)
    two = B.make_scalar(B.GDF_INT32, 2)
    B.perform_operation(outs, ins, [0, 5, 5, 1], [1, 2, 1, -2], [5, 5, 3, 4], [3, 4],
                        [B.GDF_ADD, B.GDF_MUL, B.GDF_ADD, B.GDF_ADD], [B.BLZ_INVALID_UNARY] * 4,
                        [junk] * 4, [junk, junk, junk, two])
    assert outs[0].data().tolist() == [0, 250, 980, 2190, 3880, 6050, 8700, 11830, 15440, 19530]
    assert outs[1].data().tolist() == [2, 12, 22, 32, 42, 52, 62, 72, 82, 92]


def _logical_filter_frame():
    n = 32
    i = np.arange(n)
    x = np.where(i % 2 == 0, 1, i).astype(np.int8)
    return [O.HostColumn(x, P.GDF_INT8), O.HostColumn(i.astype(np.int8), P.GDF_INT8),
            O.HostColumn((i % 2).astype(np.int8), P.GDF_INT8),
            O.HostColumn((i + 1).astype(np.int8), P.GDF_INT8)]


LOGICAL_FILTER_EXPRS = [
    ">($1, 5)", "+(*($1, $0), $2)", "=(=($1, $0), $2)", "*($0, $0)", "=(*($0, $0), 1))",
    "=($1, $2)", "AND($0, $1)", "AND(=(*($0, $0), 1), =($1, 2))",
    "AND(=(*($0, $0), 1), =($1, 2), >($2, 0))", "AND(=($0, 1), >($1, 5), >($2, 0))",
    "AND(=($0, 1), >($1, 5), >($3, 10))", "AND(=($0, 1), >($1, 5))",
    "AND(>($1, 5), >($2, 0)))", "AND(>($1, 50), >($2, 0)))", "AND(>($1, 5), >($2, 100)))",
    "AND(=($0, 1), >($1, 5), =($2, 0), >($3, 5))", "AND(=($0, 1), >($1, 5), =($2, 1), >($3, 5))",
]


@pytest.mark.parametrize("expr", LOGICAL_FILTER_EXPRS)
def test_logical_filter_expressions(B, expr):
    """logical_filter-test.cu:93-366: evaluate_expression into an INT8 column."""
    _print_case(test="logical_filter", expr=expr)
    host = _logical_filter_frame()
    dev = [to_device(B, h) for h in host]
    out = B.DeviceColumn(B.GDF_INT8, 32, with_valid=True)
    B.evaluate_expression(dev, expr, out)
    ref = O.evaluate_expression(host, expr, P.GDF_INT8, True)
    assert_column_equal(out, ref, what=expr)


@pytest.mark.parametrize("seed", range(40))
def test_random_expressions_with_nulls(B, seed):
    rng = np.random.default_rng(1000 + seed)
    k = int(rng.integers(2, 7))
    dts = exprgen.rand_frame_dtypes(rng, k)
    n = int(rng.choice([1, 17, 64, 129, 5000, 70_001]))
    null_rates = [float(rng.choice([0.0, 0.0, 0.05, 0.5])) for _ in range(k)]
    host = [gen_column(rng, dt, n, nr) for dt, nr in zip(dts, null_rates)]
    nullable = [i for i, nr in enumerate(null_rates) if nr > 0]
    expr = exprgen.rand_expr(rng, dts, int(rng.integers(1, 5)), allow_pow=False, nullable=nullable)
    out_dt = int(rng.choice([P.GDF_INT8, P.GDF_INT16, P.GDF_INT32, P.GDF_INT64, P.GDF_FLOAT32,
                             P.GDF_FLOAT64]))
    if expr.startswith("$"):  # a bare column is cloned, so the output takes its dtype
        out_dt = dts[P.get_index(expr)]
    _print_case(test="random_expr", seed=seed, n=n, expr=expr, out=out_dt)
    dev = [to_device(B, h) for h in host]
    out = B.DeviceColumn(out_dt, n, with_valid=True)
    try:
        ref = O.evaluate_expression(host, expr, out_dt, True)
    except (P.PlanError, ValueError) as e:
        with pytest.raises(B.BlzError):
            B.evaluate_expression(dev, expr, out)
        return
    nulls = B.evaluate_expression(dev, expr, out)
    assert_column_equal(out, ref, check_valid=False, what=expr)
    gv = out.valid()
    ref_valid = ref.valid if ref.valid is not None else np.ones(n, bool)
    assert np.array_equal(gv, ref_valid), f"validity differs: {expr}"
    assert nulls == ref.null_count


@pytest.mark.parametrize("seed", range(24))
def test_random_filter_project_with_nulls(B, seed):
    rng = np.random.default_rng(2000 + seed)
    k = int(rng.integers(2, 6))
    dts = exprgen.rand_frame_dtypes(rng, k)
    n = int(rng.choice([100, 4096, 33_333, 200_000]))
    null_rates = [float(rng.choice([0.0, 0.1])) for _ in range(k)]
    host = [gen_column(rng, dt, n, nr) for dt, nr in zip(dts, null_rates)]
    nullable = [i for i, nr in enumerate(null_rates) if nr > 0]
    cond = f"<({exprgen.rand_expr(rng, dts, 2, False, nullable)}, {int(rng.integers(-5, 6))})"
    exprs = [exprgen.rand_expr(rng, dts, 2, False, nullable) for _ in range(int(rng.integers(1, 4)))]
    items = [f"E{i}=[{e}]" for i, e in enumerate(exprs)] + [f"A=[${int(rng.integers(0, k))}]"]
    proj = "LogicalProject(" + ", ".join(items) + ")"
    filt = f"LogicalFilter(condition=[{cond}])"
    _print_case(test="random_filter_project", seed=seed, n=n, filt=filt, proj=proj)
    dev = [to_device(B, h) for h in host]
    try:
        ref = O.execute_project(O.process_filter(host, filt), proj)
    except (P.PlanError, ValueError):
        with pytest.raises(B.BlzError):
            B.filter_project(dev, filt, proj)
        return
    outs, kept = B.filter_project(dev, filt, proj)
    assert kept == len(ref[0])
    for g, r in zip(outs, ref):
        assert_column_equal(g, r, what=f"{filt} {proj}")


@pytest.mark.parametrize("dtype", [P.GDF_INT8, P.GDF_INT16, P.GDF_INT32, P.GDF_INT64,
                                   P.GDF_FLOAT32, P.GDF_FLOAT64])
def test_materialize_column(B, dtype):
    rng = np.random.default_rng(dtype)
    n, m = 10_000, 7_777
    _print_case(test="gather", dtype=dtype)
    col = gen_column(rng, dtype, n, 0.2)
    idx = rng.integers(-3, n, m).astype(np.int32)
    ref = O.gather(col, idx, True)
    dcol = to_device(B, col)
    didx = B.DeviceColumn.from_numpy(idx, B.GDF_INT32)
    out = B.DeviceColumn(dtype, m, with_valid=True)
    B.materialize_column(dcol, out, didx)
    # rows with a negative index are left untouched by both: compare the rest
    keep = idx >= 0
    assert np.array_equal(out.data()[keep], ref.data[keep])
    assert np.array_equal(out.valid(), ref.valid)
    assert out.null_count == ref.null_count


def test_sequence_and_boolean_mask(B):
    n = 100_000
    _print_case(test="sequence_mask")
    seq = B.DeviceColumn(B.GDF_INT32, n)
    B.gdf_sequence(seq, 0, 1)
    assert np.array_equal(seq.data(), np.arange(n, dtype=np.int32))
    rng = np.random.default_rng(5)
    st = O.HostColumn((rng.random(n) < 0.3).astype(np.int8) * rng.integers(1, 3, n).astype(np.int8),
                      P.GDF_INT8, rng.random(n) > 0.1)
    got = B.apply_boolean_mask(seq, to_device(B, st))
    assert np.array_equal(got.data(), O.stencil_select(st))


def test_program_matches_one_shot(B):
    n = 500_000
    _print_case(test="program", n=n)
    host = c1_frame(n, seed=14)
    dev = [to_device(B, h) for h in host]
    prog = B.Program(dev, C1_FILTER, C1_PROJECT)
    for _ in range(3):
        prog.launch()
        kept, nulls = prog.result()
    ref = O.execute_project(O.process_filter(host, C1_FILTER), C1_PROJECT)
    assert kept == len(ref[0])
    for g, r in zip(prog.outputs, ref):
        assert g.data().tobytes() == r.data.tobytes()
    prog.close()


@pytest.mark.parametrize("kernel", [3, 4, 5])
@pytest.mark.parametrize("seed", range(8))
def test_kernel_kinds_random_filter_project(B, kernel, seed):
    """Both kernel kinds (direct register file / tile VM) on random plans."""
    rng = np.random.default_rng(3000 + seed)
    k = int(rng.integers(2, 6))
    dts = exprgen.rand_frame_dtypes(rng, k)
    n = int(rng.choice([1000, 5000, 70_001]))
    null_rates = [float(rng.choice([0.0, 0.1])) for _ in range(k)]
    host = [gen_column(rng, dt, n, nr) for dt, nr in zip(dts, null_rates)]
    nullable = [i for i, nr in enumerate(null_rates) if nr > 0]
    cond = f"<({exprgen.rand_expr(rng, dts, 2, False, nullable)}, {int(rng.integers(-5, 6))})"
    exprs = [exprgen.rand_expr(rng, dts, 2, False, nullable) for _ in range(int(rng.integers(1, 4)))]
    items = [f"E{i}=[{e}]" for i, e in enumerate(exprs)] + [f"A=[${int(rng.integers(0, k))}]"]
    proj = "LogicalProject(" + ", ".join(items) + ")"
    filt = f"LogicalFilter(condition=[{cond}])"
    _print_case(test="kernel_kinds", kernel=kernel, seed=seed, n=n, filt=filt, proj=proj)
    try:
        ref = O.execute_project(O.process_filter(host, filt), proj)
    except (P.PlanError, ValueError):
        return
    dev = [to_device(B, h) for h in host]
    try:
        prog = B.Program(dev, filt, proj, kernel=kernel)
    except B.BlzError as e:
        pytest.skip(f"plan does not fit kernel {kernel}: {e}")
    assert prog.kernel == kernel
    prog.launch()
    kept, nulls = prog.result()
    assert kept == len(ref[0])
    for g, r, nc in zip(prog.outputs, ref, nulls):
        g.col.null_count = nc
        assert_column_equal(g, r, what=f"kernel {kernel}: {filt} {proj}")
    prog.close()


RANGE_FILTERS = [
    "AND(>=($0, -10), <($0, 10))",                     # int range, inclusive / exclusive
    "AND(<=($0, 10), >($0, -10), <($1, 500))",          # bounds in either order, a third term
    "AND(>=($0, -1000), <=($0, 1000))",                 # the whole domain
    "AND(>=($0, 5), >=($0, 7))",                        # two lower bounds: not a range
    "AND(>=($2, -2.5), <=($2, 2.5))",                   # f64 inclusive, with NaN rows
    "AND(>($2, -2.5), <($2, 2.5), >=($3, 0.0), <($3, 40.0))",  # f64 exclusive, f32 range
    "AND(<($3, 40.0), >($1, 100), >=($3, -40.0), <=($1, 900))",  # interleaved columns
    "AND(>=($1, -9223372036854775808), <=($1, 9223372036854775807))",  # int64 extremes
    "AND(>=($4, 0), <($4, 2))",                         # int8 column (general leaf)
    "AND(=($0, 3), <>($2, 1.0))",
]


@pytest.mark.parametrize("cond", RANGE_FILTERS)
def test_range_conjunctions(B, cond):
    """Conjunctions of comparisons, the shape the predicate compiler fuses into
    range ops (lower and upper bound on one operand), against the oracle; float
    columns hold NaNs (every comparison with NaN is false)."""
    rng = np.random.default_rng(7000 + RANGE_FILTERS.index(cond))
    n = 70_001
    host = [gen_column(rng, P.GDF_INT32, n, 0.05), gen_column(rng, P.GDF_INT64, n),
            gen_column(rng, P.GDF_FLOAT64, n, 0.05, -5.0, 5.0), gen_column(rng, P.GDF_FLOAT32, n, 0.0, -60.0, 60.0),
            gen_column(rng, P.GDF_INT8, n, 0.0, -3, 3)]
    host[2].data[rng.integers(0, n, 500)] = np.nan
    host[3].data[rng.integers(0, n, 500)] = np.nan
    filt = f"LogicalFilter(condition=[{cond}])"
    proj = "LogicalProject(A=[$0], B=[$1], C=[+($2, $3)], D=[*($1, 2)])"
    _print_case(test="range_conjunctions", cond=cond)
    ref = _oracle(host, filt, proj)
    dev = [to_device(B, h) for h in host]
    for kernel in (0, 3, 5):
        prog = B.Program(dev, filt, proj, kernel=kernel)
        prog.launch()
        kept, nulls = prog.result()
        assert kept == len(ref[0]), f"kernel {kernel}: {cond}"
        for g, r, nc in zip(prog.outputs, ref, nulls):
            g.col.null_count = nc
            assert_column_equal(g, r, what=f"kernel {kernel}: {cond}")
        prog.close()


@pytest.mark.parametrize("nc0", [None, 0])
def test_range_conjunction_bitmap_null_count(B, nc0):
    """A range conjunction over nullable columns (the register-load predicate
    applies the columns' bitmaps itself); with null_count 0 a bitmap that has
    cleared bits still counts as all valid (interpreter_ops.cuh:1312-1327)."""
    rng = np.random.default_rng(7100)
    n = 70_001
    host = [gen_column(rng, P.GDF_INT32, n, 0.05), gen_column(rng, P.GDF_INT64, n),
            gen_column(rng, P.GDF_FLOAT64, n, 0.05, -5.0, 5.0)]
    if nc0 is not None:
        host[0].null_count = nc0
    filt = "LogicalFilter(condition=[AND(>=($0, -500), <($0, 500), >($2, -4.0))])"
    proj = "LogicalProject(B=[$1], C=[*($2, 2.0)])"
    _print_case(test="range_conjunction_bitmap_null_count", nc0=nc0)
    ref = _oracle(host, filt, proj)
    dev = [to_device(B, h) for h in host]
    for kernel in (0, 3, 5):
        prog = B.Program(dev, filt, proj, kernel=kernel)
        prog.launch()
        kept, nulls = prog.result()
        assert kept == len(ref[0]), f"kernel {kernel}"
        for g, r, nc in zip(prog.outputs, ref, nulls):
            g.col.null_count = nc
            assert_column_equal(g, r, what=f"kernel {kernel} nc0={nc0}")
        prog.close()


WIDE_MIX = [P.GDF_INT32, P.GDF_INT32, P.GDF_INT64, P.GDF_INT64, P.GDF_FLOAT32, P.GDF_FLOAT32,
            P.GDF_FLOAT64, P.GDF_FLOAT64]


@pytest.mark.parametrize("kernel", [0, 3, 4, 5])
@pytest.mark.parametrize("seed", range(10))
def test_wide_workload_queries(B, kernel, seed):
    """Queries of the batched / served workloads (4-40 ops over eight nullable
    columns of the wide-expression mix, filtered or not): both kernel kinds and
    the default choice against the oracle. Every one of them fits the tile VM."""
    from blz_ral import workloads as W
    rng = np.random.default_rng(6000 + seed)
    n = int(rng.choice([1000, 70_001]))
    spec = {"rows": n, "dtypes": WIDE_MIX, "nulls": [0.05] * 8, "seeds": [0] * 8}
    host = [gen_column(rng, dt, n, 0.05) for dt in WIDE_MIX]
    filt, proj = W.rand_query(rng, spec, with_filter=bool(seed % 2))
    _print_case(test="wide_workload", kernel=kernel, seed=seed, n=n, filt=filt, proj=proj)
    ref = _oracle(host, filt, proj)
    dev = [to_device(B, h) for h in host]
    prog = B.Program(dev, filt, proj, kernel=kernel)
    assert prog.kernel == (5 if kernel == 0 else kernel)
    prog.launch()
    kept, nulls = prog.result()
    assert kept == len(ref[0])
    for g, r, nc in zip(prog.outputs, ref, nulls):
        g.col.null_count = nc
        assert_column_equal(g, r, what=f"kernel {kernel}: {filt} {proj}")
    prog.close()


@pytest.mark.parametrize("kernel", [3, 4, 5])
def test_kernel_kinds_c1(B, kernel):
    n = 1_000_003
    _print_case(test="kernel_kinds_c1", kernel=kernel, n=n)
    host = c1_frame(n, seed=15)
    dev = [to_device(B, h) for h in host]
    prog = B.Program(dev, C1_FILTER, C1_PROJECT, kernel=kernel)
    assert prog.kernel == kernel
    prog.launch()
    kept, _ = prog.result()
    ref = O.execute_project(O.process_filter(host, C1_FILTER), C1_PROJECT)
    assert kept == len(ref[0])
    for g, r in zip(prog.outputs, ref):
        assert g.data().tobytes() == r.data.tobytes()
    prog.close()


VM_SHAPES = [(8, 1), (8, 2), (4, 1), (4, 4), (2, 1), (2, 8)]


@pytest.mark.parametrize("shape", VM_SHAPES, ids=[f"r{r}p{p}" for r, p in VM_SHAPES])
@pytest.mark.parametrize("mode", ["filter", "project"])
def test_vm_shapes(B, shape, mode):
    """Every tile-VM shape, filter+project and pure projection, with nulls, on row
    counts below, at and just past whole tiles."""
    r, p = shape
    T = 256 * r * p
    for seed, n in enumerate([1, T - 1, T, 3 * T + 17, 40 * T + 5]):
        rng = np.random.default_rng(7100 + 17 * seed + r * 3 + p)
        k = int(rng.integers(2, 6))
        dts = exprgen.rand_frame_dtypes(rng, k)
        null_rates = [float(rng.choice([0.0, 0.1])) for _ in range(k)]
        host = [gen_column(rng, dt, n, nr) for dt, nr in zip(dts, null_rates)]
        nullable = [i for i, nr in enumerate(null_rates) if nr > 0]
        exprs = [exprgen.rand_expr(rng, dts, 2, False, nullable) for _ in range(int(rng.integers(1, 4)))]
        items = [f"E{i}=[{e}]" for i, e in enumerate(exprs)] + [f"A=[${int(rng.integers(0, k))}]"]
        proj = "LogicalProject(" + ", ".join(items) + ")"
        filt = None
        if mode == "filter":
            cond = f"<({exprgen.rand_expr(rng, dts, 2, False, nullable)}, {int(rng.integers(-5, 6))})"
            filt = f"LogicalFilter(condition=[{cond}])"
        _print_case(test="vm_shapes", r=r, p=p, n=n, filt=filt, proj=proj)
        try:
            mid = O.process_filter(host, filt) if filt else host
            ref = O.execute_project(mid, proj)
        except (P.PlanError, ValueError):
            continue
        dev = [to_device(B, h) for h in host]
        try:
            prog = B.Program(dev, filt, proj, kernel=4, r=r, t=T)
        except B.BlzError as e:
            if "fit" in str(e):
                continue
            raise
        assert prog.kernel == 4 and prog.r == r and prog.tile_rows == T
        for _ in range(2):  # a second launch must give the same result
            prog.launch()
            kept, nulls = prog.result()
            assert kept == len(ref[0])
            for g, rc, nc in zip(prog.outputs, ref, nulls):
                g.col.null_count = nc
                assert_column_equal(g, rc, what=f"vm r={r} T={T} n={n}: {filt} {proj}")
        prog.close()


def _rand_query(rng, k_min=2, k_max=6, filt_p=0.6):
    k = int(rng.integers(k_min, k_max))
    dts = exprgen.rand_frame_dtypes(rng, k)
    n = int(rng.choice([0, 1, 700, 5000, 70_001, 300_000]))
    null_rates = [float(rng.choice([0.0, 0.1])) for _ in range(k)]
    host = [gen_column(rng, dt, n, nr) for dt, nr in zip(dts, null_rates)]
    nullable = [i for i, nr in enumerate(null_rates) if nr > 0]
    exprs = [exprgen.rand_expr(rng, dts, 3, False, nullable) for _ in range(int(rng.integers(1, 4)))]
    items = [f"E{i}=[{e}]" for i, e in enumerate(exprs)] + [f"A=[${int(rng.integers(0, k))}]"]
    proj = "LogicalProject(" + ", ".join(items) + ")"
    filt = None
    if rng.random() < filt_p:
        cond = f"<({exprgen.rand_expr(rng, dts, 2, False, nullable)}, {int(rng.integers(-5, 6))})"
        filt = f"LogicalFilter(condition=[{cond}])"
    return host, filt, proj


def _oracle(host, filt, proj):
    mid = O.process_filter(host, filt) if filt else host
    return O.execute_project(mid, proj)


def _check_batch_query(batch, q, ref, what):
    kept, cols, nulls = batch.result(q)
    assert kept == len(ref[0]), what
    for g, r, nc in zip(cols, ref, nulls):
        g.col.null_count = nc
        assert_column_equal(g, r, what=what)


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("r", [4, 2])
def test_batch_matches_each_query_alone(B, r, graph):
    """Each query's result in a batch of 24 uneven queries equals its result alone
    (a Program of the same query) and the oracle's."""
    rng = np.random.default_rng(40_000 + r + 10 * graph)
    qs, refs, devs = [], [], []
    while len(qs) < 24:
        host, filt, proj = _rand_query(rng)
        try:
            ref = _oracle(host, filt, proj)
        except (P.PlanError, ValueError):
            continue
        dev = [to_device(B, h) for h in host]
        devs.append(dev)
        qs.append((dev, filt, proj))
        refs.append(ref)
    _print_case(test="batch", r=r, graph=graph, queries=len(qs))
    batch = B.Batch(qs, r=r, t=256 * r, use_graph=graph)
    assert batch.r == r and batch.graph == int(graph)
    for rep in range(3):  # replays, with eager work and synchronisation between them
        batch.launch()
        for q, ((dev, filt, proj), ref) in enumerate(zip(qs, refs)):
            _check_batch_query(batch, q, ref, f"batch r={r} q={q} rep={rep}: {filt} {proj}")
        if rep == 0:
            for q in (0, 5, 11):
                dev, filt, proj = qs[q]
                prog = B.Program(dev, filt, proj)
                prog.launch()
                kept, _ = prog.result()
                bk, bcols, _ = batch.result(q)
                assert kept == bk
                for a, b in zip(prog.outputs, bcols):
                    assert a.data().tobytes() == b.data().tobytes()
                prog.close()
    batch.close()


def test_batch_autotuner(B):
    rng = np.random.default_rng(41_000)
    qs, refs = [], []
    while len(qs) < 12:
        host, filt, proj = _rand_query(rng, filt_p=0.5)
        try:
            ref = _oracle(host, filt, proj)
        except (P.PlanError, ValueError):
            continue
        qs.append(([to_device(B, h) for h in host], filt, proj))
        refs.append(ref)
    _print_case(test="batch_autotuner", n=len(qs))
    batch = B.Batch(qs, r=0, use_graph=True)
    assert batch.tuned >= 1 and batch.r in (2, 4, 8) and batch.graph == 1
    batch.launch()
    for q, ref in enumerate(refs):
        _check_batch_query(batch, q, ref, f"autotuned batch q={q}")
    batch.close()


@pytest.mark.parametrize("budget_mb,max_batch,streams", [(4096, 8, 3), (1, 4, 2)])
def test_server_each_result_same_alone_and_in_the_mix(B, budget_mb, max_batch, streams):
    """Queries submitted back to back are admitted by budget, batched and overlapped
    on several streams; every result equals the query's result alone and the
    oracle's. A 1 MiB budget admits a large query only alone."""
    rng = np.random.default_rng(42_000 + budget_mb)
    items = []
    while len(items) < 20:
        host, filt, proj = _rand_query(rng)
        try:
            ref = _oracle(host, filt, proj)
        except (P.PlanError, ValueError):
            continue
        items.append(([to_device(B, h) for h in host], filt, proj, ref))
    _print_case(test="server", budget_mb=budget_mb, max_batch=max_batch, streams=streams)
    srv = B.Server(budget_mb << 20, max_batch=max_batch, n_streams=streams, r=4)
    tickets = [srv.submit(dev, filt, proj) for dev, filt, proj, _ in items]
    for q, (tk, (dev, filt, proj, ref)) in enumerate(zip(tickets, items)):
        kept, cols, nulls, lat = srv.wait(tk)
        assert lat >= 0
        assert kept == len(ref[0]), f"server q={q}: {filt} {proj}"
        for g, r, nc in zip(cols, ref, nulls):
            g.col.null_count = nc
            assert_column_equal(g, r, what=f"server q={q}: {filt} {proj}")
        if q % 7 == 0:  # and the same query alone, through a Program
            prog = B.Program(dev, filt, proj)
            prog.launch()
            pk, _ = prog.result()
            assert pk == kept
            for a, b in zip(prog.outputs, cols):
                assert a.data().tobytes() == b.data().tobytes()
            prog.close()
        srv.release(tk)
    st = srv.stats()
    assert st["batches"] >= 1
    if budget_mb == 1:  # only small queries share a batch under a 1 MiB budget
        assert st["batches"] >= len(items) // 2
    srv.close()


def test_server_replay_poisson(B):
    """The replay client: Poisson arrivals, every query completes, latencies > 0."""
    rng = np.random.default_rng(43_000)
    qs = []
    while len(qs) < 40:
        host, filt, proj = _rand_query(rng)
        try:
            _oracle(host, filt, proj)
        except (P.PlanError, ValueError):
            continue
        qs.append(([to_device(B, h) for h in host], filt, proj))
    arrivals = np.cumsum(rng.exponential(1 / 2000.0, len(qs)))
    _print_case(test="server_replay", n=len(qs))
    srv = B.Server(8 << 30, max_batch=16, n_streams=4, r=4)
    lat, kept, wall = srv.replay(qs, arrivals)
    assert len(lat) == len(qs) and (lat > 0).all() and wall > 0
    srv.close()


def test_synth_column_deterministic(B):
    n = 100_003
    _print_case(test="synth", n=n)
    a = B.DeviceColumn(B.GDF_FLOAT64, n, with_valid=True)
    b = B.DeviceColumn(B.GDF_FLOAT64, n, with_valid=True)
    B.synth_column(a, 7, 1, -1.0, 1.0, 0.05)
    B.synth_column(b, 7, 1, -1.0, 1.0, 0.05)
    x = a.data()
    assert x.tobytes() == b.data().tobytes()
    assert (x >= -1.0).all() and (x < 1.0).all()
    va = a.valid()
    assert 0.03 < 1 - va.mean() < 0.07
    c = B.DeviceColumn(B.GDF_INT32, n)
    B.synth_column(c, 9, 5, 10, 0)
    assert np.array_equal(c.data(), np.arange(10, 10 + n, dtype=np.int32))


def test_cpp_operator_api_filter_project(B):
    """The reference's call sequence (evaluate_expression, process_filter,
    execute_project_plan, process_aggregate) from C++ through include/blz/ral.hpp,
    against the oracle on the same generated frame, bit-exact."""
    import os
    import subprocess
    shim = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "ral_shim_test")
    n = 20_000
    _print_case(test="cpp_operator_api", n=n)
    r = subprocess.run([shim, str(n)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    lines = r.stdout.splitlines()
    i = np.arange(n, dtype=np.int64)
    a = (1 + (i * 7919) % 50).astype(np.int32)
    b = 900.0 + ((i * 104729) % 104100).astype(np.float64)
    c = ((i * 31) % 11).astype(np.float64) / 100.0
    d = ((i * 40503) % 2556).astype(np.int32)
    host = [O.HostColumn(a, P.GDF_INT32), O.HostColumn(b, P.GDF_FLOAT64), O.HostColumn(c, P.GDF_FLOAT64),
            O.HostColumn(d, P.GDF_INT32)]
    st = O.evaluate_expression(host, ">=($3, 730)", P.GDF_INT8, True)
    assert lines[0] == f"stencil {int(np.count_nonzero(st.data))}"
    filt = "LogicalFilter(condition=[AND(>=($3, 730), <($3, 1095), >=($2, 0.03), <=($2, 0.05), <($0, 30))])"
    proj = "LogicalProject(EXPR$0=[*($1, -(1, $2))], EXPR$1=[+($0, 1)], D=[$3])"
    ref = O.execute_project(O.process_filter(host, filt), proj)
    assert lines[1] == f"kept {len(ref[0])}"
    rows = [ln.split() for ln in lines[2:] if ln.startswith("row ")]
    assert len(rows) == len(ref[0])
    agg = O.process_aggregate(ref, "LogicalAggregate(group=[{}], n=[COUNT()], m=[MAX($1)], s=[SUM($1)])")
    assert lines[-1] == f"agg {agg[0].data[0]} {agg[1].data[0]} {agg[2].data[0]}"
    got0 = np.array([float(x[1]) for x in rows])
    got1 = np.array([int(x[2]) for x in rows], dtype=np.int32)
    got2 = np.array([int(x[3]) for x in rows], dtype=np.int32)
    assert got0.tobytes() == ref[0].data.tobytes()
    assert got1.tobytes() == ref[1].data.tobytes()
    assert got2.tobytes() == ref[2].data.tobytes()


DIVISORS = [2, 3, 7, 10, 87, 1000, 65536, -2, -7, -1000, 6700417, 1099511627781, 9223372036854775807,
            -9223372036854775807, 0, 1, -1]


@pytest.mark.parametrize("kernel", [0, 3, 5])
def test_integer_division_by_constants(B, kernel):
    """Integer / and MOD by literal divisors (the tile VM divides by a magic
    multiplier; 0, 1 and -1 take the general path) against the oracle, with
    int64 extremes among the dividends."""
    rng = np.random.default_rng(8100 + kernel)
    n = 20_000
    x = rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64)
    x[:6] = [0, 1, -1, 2**63 - 1, -2**63, -2**63 + 1]
    x[6:2000] >>= rng.integers(0, 63, 1994)
    y = rng.integers(-2**31, 2**31 - 1, n).astype(np.int32)
    host = [O.HostColumn(x, P.GDF_INT64, rng.random(n) > 0.05), O.HostColumn(y, P.GDF_INT32)]
    dev = [to_device(B, h) for h in host]
    items = []
    for i, d in enumerate(DIVISORS):
        items.append(f"Q{i}=[/($0, {d})]")
        items.append(f"M{i}=[MOD($0, {d})]")
        items.append(f"R{i}=[MOD($1, {d})]")
    items.append("X=[+(MOD($0, 7), /($1, -3))]")
    for k in range(0, len(items), 16):
        proj = "LogicalProject(" + ", ".join(items[k:k + 16]) + ")"
        _print_case(test="division_by_constants", kernel=kernel, part=k)
        ref = O.execute_project(host, proj)
        prog = B.Program(dev, None, proj, kernel=kernel)
        prog.launch()
        kept, nulls = prog.result()
        for g, r, nc in zip(prog.outputs, ref, nulls):
            g.col.null_count = nc
            assert_column_equal(g, r, what=f"kernel {kernel}: {proj}")
        prog.close()


BARE_DTYPES = [P.GDF_INT8, P.GDF_BOOL8, P.GDF_INT16, P.GDF_INT32, P.GDF_INT64, P.GDF_FLOAT32,
               P.GDF_FLOAT64, P.GDF_DATE32, P.GDF_DATE64, P.GDF_TIMESTAMP]


@pytest.mark.parametrize("kernel", [0, 3, 5])
@pytest.mark.parametrize("mode", ["filter", "project"])
def test_bare_column_outputs_keep_their_bytes(B, kernel, mode):
    """`$k` outputs are the (kept) rows' bytes, unchanged (the reference gathers them
    by width, ColumnManipulation.cu:83-118, or aliases them): every dtype, with
    nulls, float NaNs with payloads (quiet and signalling) and signed zeros, next to
    an evaluated output; row counts around whole tiles and chunks."""
    for n in (1, 255, 4096, 70_001):
        rng = np.random.default_rng(9100 + n)
        host = [gen_column(rng, P.GDF_INT32, n, 0.1, -50, 50)]
        for i, dt in enumerate(BARE_DTYPES):
            h = gen_column(rng, dt if dt not in (P.GDF_BOOL8,) else P.GDF_INT8, n, 0.1 * (i % 2))
            if dt == P.GDF_BOOL8:
                h = O.HostColumn((h.data != 0).astype(np.int8), dt, h.valid)
            host.append(h)
        f32 = host[1 + BARE_DTYPES.index(P.GDF_FLOAT32)].data.view(np.uint32)
        f64 = host[1 + BARE_DTYPES.index(P.GDF_FLOAT64)].data.view(np.uint64)
        k = min(n, 40)
        f32[:k:4] = np.array([0x7FC00001, 0x7F800001, 0xFFA5A5A5, 0x80000000] * 10, np.uint32)[:len(f32[:k:4])]
        f64[:k:4] = np.array([0x7FF8000000000123, 0x7FF0000000000001, 0xFFF4000000000000,
                              0x8000000000000000] * 10, np.uint64)[:len(f64[:k:4])]
        filt = "LogicalFilter(condition=[<($0, 20)])" if mode == "filter" else None
        items = [f"C{i}=[${i + 1}]" for i in range(len(BARE_DTYPES))] + ["E=[*($0, 3)]"]
        proj = "LogicalProject(" + ", ".join(items) + ")"
        _print_case(test="bare_outputs", kernel=kernel, mode=mode, n=n)
        ref = _oracle(host, filt, proj)
        dev = [to_device(B, h) for h in host]
        prog = B.Program(dev, filt, proj, kernel=kernel)
        prog.launch()
        kept, nulls = prog.result()
        assert kept == len(ref[0])
        for g, r, nc in zip(prog.outputs, ref, nulls):
            g.col.null_count = nc
            assert_column_equal(g, r, what=f"kernel {kernel} n={n}: bare-column output {r.dtype}")
        prog.close()
        for d in dev:
            d.free()


DIRECT_CASES = [
    # (predicate, projection): 4- and 8-byte terms, nullable and null_count-0 columns
    ("AND(>=($3, 730), <($3, 1095), >=($2, 0.03), <=($2, 0.05), <($0, 30))",
     "LogicalProject(E0=[*($1, -(1, $2))], E1=[+($0, 1)])"),
    ("AND(>=($3, 100), <($3, 2000), >($2, 0.02))", "LogicalProject(D=[$3], C=[$2], B=[$1], E=[*($2, 2.0)])"),
    ("<($4, 5000)", "LogicalProject(K=[$4], A=[$0])"),                            # nullable int64 term
    ("AND(>($5, -50.0), <($0, 40))", "LogicalProject(F=[+($5, 1.0)], B=[$1])"),   # float32 + nullable
    ("AND(<>($5, 0.0), =($0, 7), >=($4, 100))", "LogicalProject(A=[$0], K=[+($4, 1)])"),
]


@pytest.mark.parametrize("t", [0, 512, 2048, 4096])
@pytest.mark.parametrize("case", range(len(DIRECT_CASES)))
def test_register_load_predicate(B, case, t):
    """Filters the register-load predicate runs (asserted through the program's
    description): 4- and 8-byte terms, nullable columns and a bitmap whose
    null_count is 0 (forced valid in the predicate; the projection of the filtered
    rows reads its bits, as after the reference's materialize_column), tile rows from
    512 to 4096 (a step spanning several slots, one slot, or a slot spanning several
    steps) and row counts that are not a multiple of the tile, against the oracle."""
    cond, proj = DIRECT_CASES[case]
    filt = f"LogicalFilter(condition=[{cond}])"
    for n in (1, 777, 70_001, 300_013):
        rng = np.random.default_rng(8200 + case * 10 + n % 97)
        host = c1_frame(n, seed=30 + case)
        host.append(gen_column(rng, P.GDF_INT64, n, 0.1, 0, 10_000))     # $4: nullable int64
        host.append(gen_column(rng, P.GDF_FLOAT32, n, 0.05, -60.0, 60.0))  # $5: nullable float32
        host[0] = O.HostColumn(host[0].data, P.GDF_INT32, rng.random(n) > 0.02, 0)  # bitmap, null_count 0
        _print_case(test="register_load_predicate", case=case, t=t, n=n)
        ref = _oracle(host, filt, proj)
        dev = [to_device(B, h) for h in host]
        try:
            prog = B.Program(dev, filt, proj, t=t, kernel=4)
        except B.BlzError as e:
            pytest.skip(f"{t}-row tiles do not fit this plan's staging: {e}")
        assert prog.kernel == 4 and prog.direct_pred is not None, "not the register-load predicate"
        if t:
            assert prog.tile_rows == t
        for rep in range(2):
            prog.launch()
            kept, nulls = prog.result()
            assert kept == len(ref[0])
            for g, r, nc in zip(prog.outputs, ref, nulls):
                g.col.null_count = nc
                assert_column_equal(g, r, what=f"case {case} t={t} n={n} rep={rep}")
        prog.close()
        outs, kept = B.filter_project(dev, filt, proj)  # the drop-in call: the same bytes
        assert kept == len(ref[0])
        for g, r in zip(outs, ref):
            assert g.data().tobytes() == r.data.tobytes()
        for c in outs + dev:
            c.free()


def test_one_shot_calls_reuse_and_release_the_streams_launch_buffers(B):
    """One-shot filters keep their launch buffers per stream, growing them for a
    larger input and reusing them for smaller ones: calls of alternating sizes on the
    default stream and on a stream of the library's, with every buffer released
    midway and the stream's going with it when it is destroyed, all give the
    oracle's result."""
    import ctypes as C
    st = C.c_void_p(0)
    assert B.lib().blz_stream_create(C.byref(st)) == B.BLZ_OK
    try:
        for i, n in enumerate([70_001, 1_000_003, 5_000, 1_000_003, 1]):
            host = c1_frame(n, seed=900 + i)
            ref = O.execute_project(O.process_filter(host, C1_FILTER), C1_PROJECT)
            dev = [to_device(B, h) for h in host]
            for stream in (None, st):
                print(f"CASE one_shot_lease n={n} stream={'lib' if stream else 'default'}", flush=True)
                outs, kept = B.filter_project(dev, C1_FILTER, C1_PROJECT, stream=stream)
                assert kept == len(ref[0])
                for g, r in zip(outs, ref):
                    assert_column_equal(g, r, what=f"n={n}")
                for c in outs:
                    c.free()
            for c in dev:
                c.free()
            if i == 2:   # every kept buffer released: the next calls allocate again
                assert B.lib().blz_release_launch_buffers() == B.BLZ_OK
    finally:
        assert B.lib().blz_stream_destroy(st) == B.BLZ_OK


def test_one_shot_calls_from_threads_on_shared_and_own_streams(B):
    """Four threads issue one-shot filters at once, two on streams of their own and
    two on the default stream (whose kept launch buffer they take in turns): every
    result equals the oracle's."""
    import ctypes as C
    import threading
    cases = []
    for i, n in enumerate([300_001, 40_000, 1_000_003, 77_777]):
        host = c1_frame(n, seed=950 + i)
        ref = O.execute_project(O.process_filter(host, C1_FILTER), C1_PROJECT)
        cases.append(([to_device(B, h) for h in host], ref))
    streams = [C.c_void_p(0), C.c_void_p(0)]
    for st in streams:
        assert B.lib().blz_stream_create(C.byref(st)) == B.BLZ_OK
    errors = []

    def work(i):
        dev, ref = cases[i]
        stream = streams[i] if i < 2 else None
        try:
            for _ in range(5):
                outs, kept = B.filter_project(dev, C1_FILTER, C1_PROJECT, stream=stream)
                assert kept == len(ref[0])
                for g, r in zip(outs, ref):
                    assert_column_equal(g, r, what=f"thread {i}")
                for c in outs:
                    c.free()
        except Exception as e:  # reported by the main thread
            errors.append(f"thread {i}: {e!r}")

    print("CASE one_shot_threads", flush=True)
    threads = [threading.Thread(target=work, args=(i,)) for i in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for st in streams:
        assert B.lib().blz_stream_destroy(st) == B.BLZ_OK
    for dev, _ in cases:
        for c in dev:
            c.free()
    assert not errors, errors

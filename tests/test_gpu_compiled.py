"""The compiled tier (kernel 5: the plan written out as a kernel and compiled for
the device at run time) against the oracle, bit-exact: the cases where its
kernels differ in shape from the interpreters' — row counts around its step and
tile sizes, every rows-per-lane value, filters that keep nothing / everything,
nullable outputs of a filter (validity compacted by the gather), bare-column
stencils, relaunches of one prepared program, and frames that differ only in
their data (one compiled kernel serves both)."""
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


def _check(prog, ref, what, launches=3):
    """Launches `prog` several times; every launch must give the oracle's result,
    with the validity bits past the last output row zero."""
    for i in range(launches):
        prog.launch()
        kept, nulls = prog.result()
        assert kept == len(ref[0]), f"{what}: launch {i}: kept {kept} != {len(ref[0])}"
        for g, r, nc in zip(prog.outputs, ref, nulls):
            g.col.null_count = nc
            assert_column_equal(g, r, what=f"{what} launch {i}")
            bits = g.valid_bits()
            if bits is not None and kept % 8:
                assert bits[-1] >> (kept % 8) == 0, f"{what}: bits past the last row"


@pytest.mark.parametrize("r", [1, 2, 4, 8, 16])
@pytest.mark.parametrize("n", [1, 63, 64, 65, 511, 512, 513, 4095, 4096, 4097, 70_001])
def test_c1_sizes_and_shapes(B, n, r):
    _print_case(test="compiled_c1", n=n, r=r)
    host = c1_frame(n, seed=100 + n)
    ref = O.execute_project(O.process_filter(host, C1_FILTER), C1_PROJECT)
    dev = [to_device(B, h) for h in host]
    prog = B.Program(dev, C1_FILTER, C1_PROJECT, kernel=5, r=r)
    assert prog.kernel == 5 and prog.r == r
    _check(prog, ref, f"c1 n={n} r={r}")
    prog.close()


@pytest.mark.parametrize("cond", ["<($0, -100000)", ">($0, -100000)"], ids=["none", "all"])
def test_keep_none_and_all(B, cond):
    rng = np.random.default_rng(5)
    n = 100_003
    host = [gen_column(rng, P.GDF_INT32, n, 0.0), gen_column(rng, P.GDF_FLOAT64, n, 0.1),
            gen_column(rng, P.GDF_INT64, n, 0.2)]
    filt = f"LogicalFilter(condition=[{cond}])"
    proj = "LogicalProject(A=[$0], B=[*($1, 2)], C=[+($2, $0)], D=[$2])"
    _print_case(test="compiled_keep", cond=cond)
    ref = O.execute_project(O.process_filter(host, filt), proj)
    dev = [to_device(B, h) for h in host]
    prog = B.Program(dev, filt, proj, kernel=5)
    _check(prog, ref, cond)
    prog.close()


@pytest.mark.parametrize("seed", range(6))
def test_nullable_filter_outputs(B, seed):
    """Filters over nullable columns whose outputs carry computed validity (the
    gather's ballot words), COALESCE and IS [NOT] NULL in both programs."""
    rng = np.random.default_rng(6100 + seed)
    n = int(rng.choice([777, 4096, 65_537, 300_000]))
    host = [gen_column(rng, P.GDF_INT32, n, 0.1), gen_column(rng, P.GDF_FLOAT64, n, 0.2),
            gen_column(rng, P.GDF_INT64, n, 0.05), gen_column(rng, P.GDF_FLOAT32, n, 0.0)]
    filt = ("LogicalFilter(condition=[OR(IS NULL($1), AND(>($0, -300), <(CASE(IS NOT NULL($2), $2, 7), "
            "500)))])")
    proj = ("LogicalProject(A=[+($0, $2)], B=[CASE(IS NOT NULL($1), $1, -1.5)], C=[IS NULL($2)], "
            "D=[$1], E=[*($3, $0)])")
    _print_case(test="compiled_nullable", seed=seed, n=n)
    ref = O.execute_project(O.process_filter(host, filt), proj)
    dev = [to_device(B, h) for h in host]
    prog = B.Program(dev, filt, proj, kernel=5)
    _check(prog, ref, f"nullable seed={seed}")
    prog.close()


def test_bare_column_stencil(B):
    """A filter whose condition is a bare column (the stencil is the column's value)."""
    rng = np.random.default_rng(77)
    n = 50_001
    flag = O.HostColumn(rng.integers(0, 3, n).astype(np.int8), P.GDF_INT8, rng.random(n) > 0.1)
    val = gen_column(rng, P.GDF_FLOAT64, n, 0.0)
    host = [flag, val]
    filt = "LogicalFilter(condition=[$0])"
    proj = "LogicalProject(V=[$1], W=[-($1, 1)])"
    _print_case(test="compiled_bare_stencil", n=n)
    ref = O.execute_project(O.process_filter(host, filt), proj)
    dev = [to_device(B, h) for h in host]
    prog = B.Program(dev, filt, proj, kernel=5)
    _check(prog, ref, "bare stencil")
    prog.close()


def test_same_kernel_for_another_frame(B):
    """Two frames of the same shape share one compiled kernel (the source does not
    depend on the data); each result is its own frame's."""
    for seed in (1, 2):
        n = 90_000 + seed
        host = c1_frame(n, seed=seed)
        ref = O.execute_project(O.process_filter(host, C1_FILTER), C1_PROJECT)
        dev = [to_device(B, h) for h in host]
        prog = B.Program(dev, C1_FILTER, C1_PROJECT, kernel=5)
        _print_case(test="compiled_shared", seed=seed, n=n)
        _check(prog, ref, f"frame {seed}", launches=1)
        prog.close()


@pytest.mark.parametrize("seed", range(10))
def test_random_projections(B, seed):
    """Random projections over nullable frames of every integer and float dtype,
    POWER and the transcendentals' operand types included (floats within the
    1-ulp contract are checked by the operator-arm tests; here the ops are exact)."""
    rng = np.random.default_rng(6300 + seed)
    k = int(rng.integers(2, 7))
    dts = exprgen.rand_frame_dtypes(rng, k)
    n = int(rng.choice([1, 100, 513, 20_000, 131_073]))
    null_rates = [float(rng.choice([0.0, 0.1])) for _ in range(k)]
    host = [gen_column(rng, dt, n, nr) for dt, nr in zip(dts, null_rates)]
    nullable = [i for i, nr in enumerate(null_rates) if nr > 0]
    exprs = [exprgen.rand_expr(rng, dts, 3, False, nullable) for _ in range(int(rng.integers(1, 5)))]
    proj = "LogicalProject(" + ", ".join(f"E{i}=[{e}]" for i, e in enumerate(exprs)) + \
           f", A=[${int(rng.integers(0, k))}])"
    _print_case(test="compiled_projection", seed=seed, n=n, proj=proj)
    try:
        ref = O.execute_project(host, proj)
    except (P.PlanError, ValueError):
        return
    dev = [to_device(B, h) for h in host]
    prog = B.Program(dev, None, proj, kernel=5)
    assert prog.kernel == 5
    _check(prog, ref, proj, launches=2)
    prog.close()


def test_programs_over_sources_compiled_together(B):
    """Plans whose sources were compiled as one program (one compilation for all of
    them): each prepared program finds its kernels there and gives the oracle's
    result, filters and projections alike."""
    rng = np.random.default_rng(6400)
    dts = [P.GDF_INT32, P.GDF_INT64, P.GDF_FLOAT32, P.GDF_FLOAT64]
    n = 77_777
    host = [gen_column(rng, dt, n, nr) for dt, nr in zip(dts, [0.1, 0.0, 0.05, 0.0])]
    plans = []
    for k in range(5):
        e = exprgen.rand_expr(rng, dts, 3, False, [])
        proj = f"LogicalProject(E=[+({e}, {6400 + k})], F=[$1], G=[*($3, {k + 2})])"
        filt = f"LogicalFilter(condition=[<($0, {100 * k - 150})])" if k % 2 else None
        plans.append((filt, proj))
    srcs = [B.program_source(dts, n, f, p, has_valid=[h.valid is not None for h in host],
                             null_counts=[h.null_count for h in host]) for f, p in plans]
    c0 = B.jit_counters()[0]
    B.jit_compile_many(srcs)
    assert B.jit_counters()[0] == c0 + 1
    dev = [to_device(B, h) for h in host]
    for filt, proj in plans:
        _print_case(test="compiled_together", filt=filt, proj=proj)
        mid = O.process_filter(host, filt) if filt else host
        ref = O.execute_project(mid, proj)
        prog = B.Program(dev, filt, proj, kernel=5)
        assert prog.kernel == 5
        _check(prog, ref, f"{filt} {proj}", launches=2)
        prog.close()
    assert B.jit_counters()[0] == c0 + 1   # no compilation when the programs were made
    for c in dev:
        c.free()

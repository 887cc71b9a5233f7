"""The compiled tier's kernel sources (no GPU): generated for random plans and the
benchmark workloads, and compiled for gfx950 with the library's run-time
compiler, so a plan shape the generator writes wrongly fails here, not on a
device."""
import numpy as np
import pytest

from tests import exprgen
from oracle import planner as P

C1_FILTER = ("LogicalFilter(condition=[AND(>=($3, 730), <($3, 1095), >=($2, 0.03), "
             "<=($2, 0.05), <($0, 30))])")
C1_PROJECT = "LogicalProject(EXPR$0=[*($1, -(1, $2))], EXPR$1=[+($0, 1)])"


@pytest.fixture(scope="module")
def B():
    import blz_ral
    blz_ral.lib()
    return blz_ral


def test_c1_source_shape(B):
    src = B.program_source([P.GDF_INT32, P.GDF_FLOAT64, P.GDF_FLOAT64, P.GDF_INT32], 1000, C1_FILTER,
                           C1_PROJECT)
    assert "blz_cg_pred" in src and "blz_cg_gather" in src
    # the predicate kernel stages a and c (the projection reads them) for the kept
    # rows; the gather reads b at the kept rows' source rows
    import re
    assert len(re.findall(r"s\d\[p\] = r\d\[u\];", src)) == 2
    assert len(re.findall(r"__builtin_nontemporal_load\(c\d \+ row\[u\]\)", src)) == 1
    B.jit_compile(src)


def test_projection_source_compiles(B):
    from blz_ral import workloads as W
    dts = [c[0] for c in W.WIDE_COLUMNS]
    src = B.program_source(dts, 1000, None, W.WIDE_PROJECT, has_valid=[True] * len(dts),
                           null_counts=[5] * len(dts))
    assert "blz_cg_project" in src and "store_vword" in src
    B.jit_compile(src)


@pytest.mark.parametrize("seed", range(6))
def test_random_plans_compile(B, seed):
    """Random filter + projection plans over nullable frames of every integer and
    float dtype (COALESCE, IS [NOT] NULL, POWER, the transcendentals' dtypes)."""
    rng = np.random.default_rng(900 + seed)
    k = int(rng.integers(2, 6))
    dts = exprgen.rand_frame_dtypes(rng, k)
    has_valid = [bool(rng.random() < 0.5) for _ in range(k)]
    nullable = [i for i in range(k) if has_valid[i]]
    cond = f"<({exprgen.rand_expr(rng, dts, 3, True, nullable)}, {int(rng.integers(-5, 6))})"
    exprs = [exprgen.rand_expr(rng, dts, 3, True, nullable) for _ in range(int(rng.integers(1, 4)))]
    items = [f"E{i}=[{e}]" for i, e in enumerate(exprs)] + [f"A=[${int(rng.integers(0, k))}]"]
    proj = "LogicalProject(" + ", ".join(items) + ")"
    filt = f"LogicalFilter(condition=[{cond}])" if seed % 2 == 0 else None
    try:
        src = B.program_source(dts, 5000, filt, proj, has_valid=has_valid, null_counts=[3] * k)
    except B.BlzError as e:
        pytest.skip(f"plan rejected by the planner: {e}")
    B.jit_compile(src)


def test_transcendental_and_date_sources_compile(B):
    dts = [P.GDF_FLOAT64, P.GDF_DATE32, P.GDF_DATE64, P.GDF_INT32]
    proj = ("LogicalProject(a=[SIN($0)], b=[LN(ABS($0))], c=[EXTRACT(FLAG(YEAR), $1)], "
            "d=[EXTRACT(FLAG(HOUR), $2)], e=[POWER($3, 3)], f=[POWER($0, 0.5)], g=[LOG10($0)], "
            "h=[EXTRACT(FLAG(MINUTE), $2)])")
    src = B.program_source(dts, 777, "LogicalFilter(condition=[>($3, 0)])", proj)
    B.jit_compile(src)


def test_bad_source_reports_the_compiler_log(B):
    with pytest.raises(B.BlzError) as e:
        B.jit_compile("this is not HIP")
    assert "hiprtc" in str(e.value)


def test_many_sources_compile_as_one_program(B):
    """Sources compiled together are compiled once (one run of the compiler for
    all of them, filters and projections alike), and each is found compiled
    afterwards; a source that does not compile leaves the others compiled."""
    rng = np.random.default_rng(4242)
    dts = [P.GDF_INT32, P.GDF_INT64, P.GDF_FLOAT32, P.GDF_FLOAT64]
    exprs = [exprgen.rand_expr(rng, dts, 3, True, []) for _ in range(5)]

    def sources(base, ks):
        out = []
        for k in ks:
            proj = f"LogicalProject(E=[+({exprs[k]}, {base + k})], F=[$1])"  # distinct literals: not cached yet
            filt = f"LogicalFilter(condition=[<($0, {base + 100 + k})])" if k % 2 else None
            out.append(B.program_source(dts, 4096, filt, proj, has_valid=[True, False, True, False],
                                        null_counts=[2, 0, 2, 0]))
        return out
    srcs = sources(7000, range(5))
    broken = srcs[0].replace("extern \"C\"", "extern \"C\" this_does_not_parse", 1) + "\n// x"
    c0, _, _ = B.jit_counters()
    B.jit_compile_many(srcs)
    c1, h1, _ = B.jit_counters()
    assert c1 - c0 == 1
    for s in srcs:
        B.jit_compile(s)                      # found compiled: no new compilation
    assert B.jit_counters()[0] == c1
    more = sources(8000, range(3))
    B.jit_compile_many(more + [broken])     # the combined compile fails: each compiled on its own
    assert B.jit_counters()[0] == c1 + 3      # (a failed compilation is not counted)
    for s in more:
        B.jit_compile(s)
    assert B.jit_counters()[0] == c1 + 3
    with pytest.raises(B.BlzError):
        B.jit_compile(broken)


def test_generated_loads_are_unconditional_and_null_counts_per_workgroup(B):
    """The generated gather loads every lane's index, staged value, source row and
    bitmap byte without a condition of its own (a conditional load compiles to a
    branch that waits for it alone), and the kernels add their null counts once per
    workgroup (one atomic per wave serializes on the count words)."""
    import re
    dts = [P.GDF_INT32, P.GDF_INT32, P.GDF_INT64, P.GDF_INT64, P.GDF_FLOAT32, P.GDF_FLOAT32, P.GDF_FLOAT64,
           P.GDF_FLOAT64]
    filt = "LogicalFilter(condition=[<($4, 39.26)])"
    proj = "LogicalProject(E0=[*(CASE(IS NOT NULL($6), $6, $2), 48)], E1=[AND(NOT($1), $5)], E2=[+($4, 1)])"
    src = B.program_source(dts, 1 << 20, filt, proj, has_valid=[True] * 8, null_counts=[100] * 8)
    gather = src[src.index("blz_cg_gather"):]
    assert not re.search(r"\?\s*\(?(int64_t\)\s*)?__builtin_nontemporal_load", gather)
    assert "valid_bit(" not in gather
    assert gather.count("atomicAdd(ws.null_counts") == 3            # one per nullable output ...
    assert "s_nn[threadIdx.x >> 6]" in gather and "__syncthreads()" in gather   # ... per workgroup
    proj_src = B.program_source(dts, 1 << 20, None, proj, has_valid=[True] * 8, null_counts=[100] * 8)
    assert "s_nn[threadIdx.x >> 6]" in proj_src
    B.jit_compile(src)
    B.jit_compile(proj_src)


def test_tail_step_loads_stay_inside_the_column(B):
    """A filter's last tile holds slots that start past the last row (four slots per
    tile): the generated predicate's tail step loads the column's last row in the
    lanes past it, never row g0 * 64 of such a slot (past the column's end)."""
    src = B.program_source([P.GDF_FLOAT64, P.GDF_INT16, P.GDF_FLOAT64], 70_001,
                           "LogicalFilter(condition=[<(MOD(*($2, $2), CEIL($0)), 0)])",
                           "LogicalProject(E0=[/(5, $1)], A=[$1])", has_valid=[False, False, True],
                           null_counts=[0, 0, 7000])
    pred = src[:src.index("blz_cg_gather")]
    loads = [l for l in pred.splitlines() if "__builtin_nontemporal_load" in l]
    assert loads
    for l in loads:
        assert "in_range ? b" in l and "+ (n - 1))" in l, l

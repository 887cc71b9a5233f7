"""Host-side checks that need no GPU.

- The C-ABI library loads and exports every entry point include/blz/*.h declares.
- The C++ planner (the product's, behind blz_plan_expression) produces the same
  plan vectors as the oracle's restatement of add_expression_to_plan
  (LogicalFilter.cpp:44-267) on seeded random expressions, and both reject the
  plans the reference cannot run.
- The C1 predicate's plan matches its hand derivation from the reference's
  register allocation (inputs in first-use order, final output at n_in, temps
  lowest-free from n_in + 1).
"""
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import planner as P

from . import exprgen
from .conftest import ROOT

HEADERS = [os.path.join(ROOT, "include", "blz", h) for h in ("blz_ral.h", "gdf_types.h")]


@pytest.fixture(scope="module")
def B():
    import blz_ral
    blz_ral.lib()  # loads on a host without a GPU; no compute is called here
    return blz_ral


def declared_entry_points():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:[\w\s\*]+?)\b(blz_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_entry_point(B):
    names = declared_entry_points()
    assert len(names) >= 25, names
    out = subprocess.run(["nm", "-D", "--defined-only", B.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    missing = sorted(n for n in names if n not in exported and not n.startswith("blz_dtype"))
    assert not missing, missing
    for n in names:
        if n in exported:
            getattr(B.lib(), n)  # resolvable through ctypes


def test_version_and_no_device(B):
    assert B.lib().blz_version() is not None


def _scalar_value(s):
    field = {P.GDF_FLOAT32: "fp32", P.GDF_FLOAT64: "fp64", P.GDF_INT8: "si08", P.GDF_INT16: "si16",
             P.GDF_INT32: "si32"}.get(s.dtype, "si64")
    v = getattr(s.data, field)
    return float(v) if field.startswith("fp") else int(v)


def _py_scalar_value(s):
    if s.dtype == P.GDF_FLOAT32:
        return float(np.float32(s.value))
    if s.dtype == P.GDF_FLOAT64:
        return float(s.value)
    return int(s.value)


def compare_plans(B, dtypes, expr):
    try:
        want = P.plan_expression(dtypes, expr)
    except P.PlanError:
        with pytest.raises(B.BlzError):
            B.plan_expression(dtypes, expr)
        return "rejected"
    if want is None:
        return "bare"
    got = B.plan_expression(dtypes, expr)
    assert got["left"] == want.left, expr
    assert got["right"] == want.right, expr
    assert got["outputs"] == want.outputs, expr
    assert got["binops"] == want.binops, expr
    assert got["unops"] == want.unops, expr
    assert got["input_columns"] == want.input_columns, expr
    assert got["final_position"] == want.final_positions[0], expr
    for side, wl in (("lscalars", want.lscalars), ("rscalars", want.rscalars)):
        for k, (g, w) in enumerate(zip(got[side], wl)):
            uses = (want.left[k] if side == "lscalars" else want.right[k]) in (-2, -3)
            if not uses:
                continue
            assert g.dtype == w.dtype, (expr, side, k)
            assert bool(g.is_valid) == bool(w.is_valid), (expr, side, k)
            if w.is_valid:
                assert _scalar_value(g) == _py_scalar_value(w), (expr, side, k)
    return "ok"


@pytest.mark.parametrize("seed", range(40))
def test_cpp_planner_matches_oracle_planner(B, seed):
    rng = np.random.default_rng(900 + seed)
    k = int(rng.integers(1, 7))
    dtypes = exprgen.rand_frame_dtypes(rng, k)
    seen = set()
    for _ in range(25):
        expr = exprgen.rand_expr(rng, dtypes, int(rng.integers(1, 5)), True, list(range(k)))
        seen.add(compare_plans(B, dtypes, expr))
    assert "ok" in seen


@pytest.mark.parametrize("expr", [
    "+(1, 2)",               # both operands literal (LogicalFilter.cpp:121-129)
    "SIN(1.5)",              # unary on a literal (:234-235)
    "=($0, 'abc')",          # string literal (NVCategory path)
    "BITWISE_AND($0, $1)",   # no kernel branch
])
def test_planner_rejections_agree(B, expr):
    assert compare_plans(B, [P.GDF_INT32, P.GDF_INT32], expr) == "rejected"


def test_c1_predicate_plan_by_hand(B):
    """AND(>=($3, 730), <($3, 1095), >=($2, 0.03), <=($2, 0.05), <($0, 30)) over
    (a i32, b f64, c f64, d i32): expand_if_logical_op makes it left-deep
    AND(AND(AND(AND(p1, p2), p3), p4), p5); tokens are processed right to left, so
    p5 (<($0, 30)) is planned first. Inputs are the used columns in frame order
    (a, c, d -> slots 0, 1, 2), the result slot is 3, temporaries start at 4."""
    dtypes = [P.GDF_INT32, P.GDF_FLOAT64, P.GDF_FLOAT64, P.GDF_INT32]
    expr = "AND(>=($3, 730), <($3, 1095), >=($2, 0.03), <=($2, 0.05), <($0, 30))"
    got = B.plan_expression(dtypes, expr)
    assert got["input_columns"] == [0, 2, 3]
    assert got["final_position"] == 3
    assert len(got["binops"]) == 9
    assert sorted(got["binops"]).count(P.GDF_MUL) == 4  # AND -> GDF_MUL
    assert got["outputs"][-1] == 3
    assert all(o == 3 or o >= 4 for o in got["outputs"])
    assert compare_plans(B, dtypes, expr) == "ok"


SHIM = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "ral_shim_test")


def _shim():
    if not os.path.exists(SHIM):
        subprocess.run(["make", "-C", os.path.dirname(SHIM)], check=True, capture_output=True)
    return SHIM


def test_cpp_operator_api_compiles_and_reports_errors():
    """include/blz/ral.hpp (the reference's C++ operator API over the C ABI) builds
    into a program, and a failing call surfaces as std::runtime_error with the
    library's message (the reference's error behaviour, Utils.cuh:25-55)."""
    r = subprocess.run([_shim(), "errors"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("runtime_error: evaluate_expression:"), r.stdout


INTEROP_PLANS = [  # benchmarks/interops/interops_benchmark.cpp:32-37
    "LogicalProject(EXPR$0=[+($0, $2)])",
    "LogicalProject(EXPR$0=[*(SIN(/(+($0, $2), $1)), POWER(/(-($0, $2), $1), 2))])",
    "LogicalProject(EXPR$1=[-(FLOOR(+(POWER(SIN($1), 2), POWER(COS($1), 2))), CEIL(+(POWER(SIN(+($1, 53.42)), 2), "
    "POWER(COS(+($1, 53.42)), 2))))])",
    "LogicalProject(EXPR$0=[*(*(10, SIN(/(+($0, $2), $1))), POWER(/(-($0, $2), $1), 2))], EXPR$1=[-(FLOOR(+(POWER("
    "SIN($1), 2), POWER(COS($1), 2))), CEIL(+(POWER(SIN(+($1, 53.42)), 2), POWER(COS(+($1, 53.42)), 2))))], "
    "EXPR$2=[+(-(/(*(FLOOR(+($1, 0.1)), POWER(MOD($0, 13), 2.0)), 5), CEIL(MOD(*(2, $2), 57))), 0.001)])",
    "LogicalProject(a=[$2], b=[*($0, 2)], c=[$0], d=[CASE(IS NOT NULL($1), $1, 7.5)], e=[>($2, $0)])",
]


@pytest.mark.parametrize("qp", INTEROP_PLANS)
def test_parse_project_plan_matches_oracle_planner(B, qp):
    """blz_parse_project_plan (parse_project_plan, CalciteInterpreter.cpp:123-309)
    against the oracle's restatement: the same register plan, input columns in
    frame order, final registers and per-output kind / index / dtype."""
    dts = [P.GDF_INT32, P.GDF_FLOAT64, P.GDF_INT32]
    want, outs, _ = P.parse_project(dts, qp)
    got = B.parse_project_plan(dts, qp)
    for key in ("left", "right", "outputs", "binops", "unops"):
        assert got[key] == getattr(want, key), key
    assert got["input_columns"] == want.input_columns
    assert got["final_positions"] == want.final_positions
    assert got["outputs_kind"] == [0 if k == "expr" else 1 for k, _ in outs]
    assert got["outputs_index"] == [i for _, i in outs]
    ev = iter(want.output_dtypes)
    assert got["outputs_dtype"] == [next(ev) if k == "expr" else dts[i] for k, i in outs]


@pytest.mark.parametrize("seed", range(10))
def test_add_expression_to_plan_matches_oracle_planner(B, seed):
    """blz_add_expression_to_plan (LogicalFilter.cpp:44-267) appends exactly the ops
    the oracle's add_expression_to_plan appends, for an expression placed at an
    arbitrary output position of a multi-output plan."""
    rng = np.random.default_rng(950 + seed)
    k = int(rng.integers(2, 7))
    dtypes = exprgen.rand_frame_dtypes(rng, k)
    n_done = 0
    for _ in range(10):
        expr = exprgen.rand_expr(rng, dtypes, int(rng.integers(1, 5)), True, list(range(k)))
        if expr.startswith("$"):
            continue
        used = sorted({int(m) for m in re.findall(r"\$(\d+)", expr)})
        idx = [-1] * k
        for slot, c in enumerate(used):
            idx[c] = slot
        num_outputs = int(rng.integers(1, 4))
        pos = int(rng.integers(0, num_outputs))
        want = P.Plan()
        try:
            P.add_expression_to_plan(want, expr, pos, num_outputs, len(used), idx)
        except P.PlanError:
            with pytest.raises(B.BlzError):
                B.add_expression_to_plan(expr, pos, num_outputs, len(used), idx)
            continue
        got = B.add_expression_to_plan(expr, pos, num_outputs, len(used), idx)
        for key in ("left", "right", "outputs", "binops", "unops"):
            assert got[key] == getattr(want, key), (expr, key)
        n_done += 1
    assert n_done > 0

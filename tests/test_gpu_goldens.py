"""The reference's own golden tables through the HIP path (the same fixtures the
oracle is checked against on the CPU, tests/golden/reference_goldens.json):
evaluate_query.cu (TEST_IS_NULL, TEST_IS_NOT_NULL, TEST_01..09, TEST_UNARY),
unary-tests.cu (trig, LN, CEIL / FLOOR), where-test.cpp (six filters) through
process_filter / execute_project_plan, and interpreter_tests.cu's register plans
through perform_operation. Each result is checked as the reference's test checks
it and against the oracle (bit-exact; the transcendentals within 1 ulp).
"""
import numpy as np
import pytest

from oracle import planner as P
from oracle import pyoracle as O
from tests.helpers import ulp_diff
from tests.test_oracle_golden import CASES, DT, GOLD, _scalar, check_columns, host_table

pytestmark = pytest.mark.gpu

TRANSCENDENTAL = ("SIN(", "COS(", "TAN(", "LN(", "LOG10(", "POWER(")


@pytest.fixture(scope="module")
def B():
    import blz_ral
    if blz_ral.device_count() < 1:
        pytest.fail("no HIP device visible")
    return blz_ral


def _to_host(c):
    return O.HostColumn(c.data(), c.dtype if c.dtype else c.col.dtype, c.valid(), c.null_count)


def _device_table(B, tables):
    return [B.DeviceColumn.from_numpy(h.data, h.dtype, valid=h.valid, null_count=h.null_count)
            for h in tables]


def run_logical_plan_gpu(B, tables, plan_text):
    """The bottom-up walk of a scan / filter / project plan
    (CalciteInterpreter.cpp:731-870) over device columns."""
    lines = [l.strip() for l in plan_text.strip().split("\n") if l.strip()]
    frame = _device_table(B, tables)
    for line in reversed(lines):
        if "TableScan" in line:
            continue
        if line.startswith("LogicalFilter"):
            frame = B.process_filter(frame, line)
        elif line.startswith("LogicalProject"):
            frame = B.execute_project(frame, line)
        else:
            raise ValueError(f"unsupported plan line: {line}")
    return [_to_host(c) for c in frame]


@pytest.mark.parametrize("kind,case", CASES, ids=[f"{k}-{c['name']}" for k, c in CASES])
def test_logical_plan_goldens_on_gpu(B, kind, case):
    print(f"CASE golden {kind} {case['name']}", flush=True)
    tables = host_table(case["tables"])
    got = run_logical_plan_gpu(B, tables, case["plan"])
    check_columns(got, case["expected"])
    ref = O.run_logical_plan(tables, case["plan"])
    approx = any(t in case["plan"] for t in TRANSCENDENTAL)
    for g, r in zip(got, ref):
        assert g.data.dtype == r.data.dtype
        if approx and r.data.dtype.kind == "f":
            d = ulp_diff(g.data, r.data)
            print(f"ULP golden {case['name']} max={d.max() if len(d) else 0:g}", flush=True)
            assert (d <= 1).all()
        else:
            assert g.data.tobytes() == r.data.tobytes()
        if r.valid is not None and len(r.data):
            assert g.valid is not None and np.array_equal(g.valid, r.valid)
            assert g.null_count == r.null_count


@pytest.mark.parametrize("case", GOLD["interpreter"], ids=[c["name"] for c in GOLD["interpreter"]])
def test_interpreter_plan_goldens_on_gpu(B, case):
    """interpreter_tests.cu register plans through perform_operation, in the
    INTENDED scalar mode (TEST_02's expected values need it) and, against the
    oracle, in the SHIPPED mode."""
    print(f"CASE golden interpreter {case['name']}", flush=True)
    inputs = [O.HostColumn(np.array(c["values"], dtype=O.NP_OF[DT[c["dtype"]]]), DT[c["dtype"]])
              for c in case["inputs"]]
    binops = [getattr(P, "GDF_" + b) if b != "INVALID" else P.GDF_INVALID_BINARY for b in case["binops"]]
    unops = [getattr(P, "BLZ_" + u) if u != "INVALID" else P.BLZ_INVALID_UNARY for u in case["unops"]]
    ls = [_scalar(s) for s in case["lscalars"]]
    rs = [_scalar(s) for s in case["rscalars"]]
    plan = P.Plan(left=case["left"], right=case["right"], outputs=case["outputs"], binops=binops, unops=unops,
                  lscalars=ls, rscalars=rs, final_positions=case["final_positions"])
    out_dt = [DT[d] for d in case["out_dtypes"]]
    dev = [B.DeviceColumn.from_numpy(h.data, h.dtype) for h in inputs]
    n = len(inputs[0].data)
    for mode in (0, 1):
        outs = [B.DeviceColumn(dt, n) for dt in out_dt]
        B.perform_operation(outs, dev, plan.left, plan.right, plan.outputs, plan.final_positions, binops, unops,
                            [B.make_scalar(s.dtype, s.value, s.is_valid) for s in ls],
                            [B.make_scalar(s.dtype, s.value, s.is_valid) for s in rs], scalar_mode=mode)
        ref = O.perform_operation(inputs, plan, out_dt, [False] * len(out_dt), scalar_mode=mode)
        for o, r in zip(outs, ref):
            g = o.data()
            if r.data.dtype.kind == "f":
                d = ulp_diff(g, r.data)
                print(f"ULP golden {case['name']} mode={mode} max={d.max():g}", flush=True)
                assert (d <= 1).all()
            else:
                assert g.tobytes() == r.data.tobytes()
        if mode == 0:  # the reference test's own expectation
            for o, e in zip(outs, case["expected"]):
                g = o.data()
                e = np.array(e, dtype=np.float64)
                if case["compare"] == "exact":
                    assert np.array_equal(g.astype(np.float64), e)
                elif case["compare"] == "float32":
                    assert np.array_equal(g, e.astype(np.float32))
                else:
                    assert np.allclose(g.astype(np.float64), e, rtol=1e-5, atol=1e-6)

"""The oracle against the reference's own golden vectors (CPU only).

tests/golden/reference_goldens.json holds the literal input tables, plans and
expected tables of the reference's tests (extracted as data by
tests/golden/make_fixtures.py):
  evaluate_query.cu   TEST_IS_NULL :44, TEST_IS_NOT_NULL :77, TEST_01..09, TEST_UNARY
  unary-tests.cu      TEST_00 :43 (trig, 10 printed digits), TEST_01 LN, TEST_02 CEIL/FLOOR
  where-test.cpp      six FLOAT64 comparison filters (:66-118)
  interpreter_tests.cu TEST_00 :68, TEST_FLOATS :161, TEST_SCALAR_ADD :248,
                      TEST_01_Unary_Operators :334, TEST_02 :434 (register plans)
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import planner as P
from oracle import pyoracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_goldens.json")))

DT = {name: getattr(P, name) for name in dir(P) if name.startswith("GDF_")}


def host_table(cols):
    """Fixture columns -> HostColumns. A column shorter than the table (the
    reference's TEST_09 lists 9 salaries for 10 rows; its fixture reads past the
    literal) is padded with zeros, and those rows are never compared."""
    n = max(len(c["values"]) for c in cols)
    out = []
    for c in cols:
        dt = DT[c["dtype"]]
        v = list(c["values"]) + [0] * (n - len(c["values"]))
        valid = c["valid"]
        if valid is not None:
            valid = list(valid) + [True] * (n - len(valid))
        out.append(O.HostColumn(np.array(v, dtype=O.NP_OF[dt]), dt,
                                None if valid is None else np.array(valid, dtype=bool)))
    return out


def as_reference_strings(values, valid, is_float):
    """The reference's comparison (tests/utils/gdf/library/column.h:166-179): each
    value streamed with the default ostream format (6 significant digits for
    floating point), a null row as '@'."""
    out = []
    for i, v in enumerate(values):
        if valid is not None and not valid[i]:
            out.append("@")
        elif is_float:
            out.append("%g" % float(v))
        else:
            out.append(str(int(v)))
    return out


def check_columns(got, expected):
    assert len(got) == len(expected)
    for g, e in zip(got, expected):
        is_float = DT[e["dtype"]] in (P.GDF_FLOAT32, P.GDF_FLOAT64)
        assert len(g.data) == len(e["values"]), (len(g.data), len(e["values"]))
        want = as_reference_strings(e["values"], e.get("valid"), is_float)
        have = as_reference_strings(g.data.tolist(), None if g.valid is None else g.valid.tolist(), is_float)
        if e.get("valid") is None:  # the expected table has no validity: compare values only
            have = as_reference_strings(g.data.tolist(), None, is_float)
        assert have == want
        # and the values themselves, where the fixture holds them exactly
        if not is_float:
            assert np.array_equal(g.data.astype(np.int64), np.array(e["values"], dtype=np.int64))


CASES = [(k, c) for k in ("evaluate_query", "where", "unary") for c in GOLD[k]]


@pytest.mark.parametrize("kind,case", CASES, ids=[f"{k}-{c['name']}" for k, c in CASES])
def test_logical_plan_goldens(kind, case):
    tables = host_table(case["tables"])
    got = O.run_logical_plan(tables, case["plan"])
    check_columns(got, case["expected"])


def _scalar(s):
    if s is None:
        return P.DUMMY
    dt = DT[s[0]]
    return P.Scalar(dt, s[1], True)


@pytest.mark.parametrize("case", GOLD["interpreter"], ids=[c["name"] for c in GOLD["interpreter"]])
def test_interpreter_plan_goldens(case):
    """Register plans of interpreter_tests.cu through oracle_perform_operation."""
    inputs = [O.HostColumn(np.array(c["values"], dtype=O.NP_OF[DT[c["dtype"]]]), DT[c["dtype"]])
              for c in case["inputs"]]
    plan = P.Plan(left=case["left"], right=case["right"], outputs=case["outputs"],
                  binops=[getattr(P, "GDF_" + b) if b != "INVALID" else P.GDF_INVALID_BINARY
                          for b in case["binops"]],
                  unops=[getattr(P, "BLZ_" + u) if u != "INVALID" else P.BLZ_INVALID_UNARY
                         for u in case["unops"]],
                  lscalars=[_scalar(s) for s in case["lscalars"]],
                  rscalars=[_scalar(s) for s in case["rscalars"]],
                  final_positions=case["final_positions"])
    out_dt = [DT[d] for d in case["out_dtypes"]]
    got = O.perform_operation(inputs, plan, out_dt, [False] * len(out_dt))
    for g, e in zip(got, case["expected"]):
        e = np.array(e, dtype=np.float64)
        gv = g.data.astype(np.float64)
        if case["compare"] == "exact":
            assert np.array_equal(gv, e)
        elif case["compare"] == "float32":
            assert np.array_equal(g.data, e.astype(np.float32))
        else:  # printed with 6 significant digits in the reference test
            assert np.allclose(gv, e, rtol=1e-5, atol=1e-6)


def test_interpreter_test02_needs_intended_literals():
    """interpreter_tests.cu:434-514 expects sin(x)+2.33: only the bit-preserving
    scalar (INTENDED) reproduces it; the shipped int64 cast turns 2.33 into a
    denormal and the sum into sin(x)."""
    case = next(c for c in GOLD["interpreter"] if c["name"] == "TEST_02")
    inputs = [O.HostColumn(np.array(c["values"], dtype=O.NP_OF[DT[c["dtype"]]]), DT[c["dtype"]])
              for c in case["inputs"]]
    plan = P.Plan(left=case["left"], right=case["right"], outputs=case["outputs"],
                  binops=[getattr(P, "GDF_" + b) if b != "INVALID" else P.GDF_INVALID_BINARY
                          for b in case["binops"]],
                  unops=[getattr(P, "BLZ_" + u) if u != "INVALID" else P.BLZ_INVALID_UNARY
                         for u in case["unops"]],
                  lscalars=[_scalar(s) for s in case["lscalars"]],
                  rscalars=[_scalar(s) for s in case["rscalars"]],
                  final_positions=case["final_positions"])
    out_dt = [DT[d] for d in case["out_dtypes"]]
    shipped = O.perform_operation(inputs, plan, out_dt, [False, False], scalar_mode=1)
    x = inputs[1].data.astype(np.float64)
    assert np.array_equal(shipped[1].data, np.sin(x).astype(np.float32))

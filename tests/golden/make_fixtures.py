"""Writes tests/golden/reference_goldens.json from the reference's own test sources.

Run in the development container (it reads the reference tree as text; nothing of
the reference is imported or executed):

    python tests/golden/make_fixtures.py /path/to/reference

It extracts the literal input tables, logical plans and expected result tables of
  tests/evaluate_query/evaluate_query.cu   (TEST_IS_NULL, TEST_IS_NOT_NULL, TEST_01-09, TEST_UNARY)
  tests/unary/unary-tests.cu               (TEST_00 trig, TEST_01 LN, TEST_02 CEIL/FLOOR)
  tests/where/where-test.cpp               (6 FLOAT64 comparison filters)
and adds the hand-built register plans of tests/interpreter/interpreter_tests.cu
(TEST_00 :68-156, TEST_FLOATS :161-244, TEST_SCALAR_ADD :248-330,
TEST_01_Unary_Operators :334-417, TEST_02 :434-514), transcribed with their
expected tables. The JSON holds data only (inputs, plans, expected outputs).
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def balanced(s, i):
    """s[i] == '{' -> index of the matching '}'."""
    depth = 0
    for j in range(i, len(s)):
        if s[j] == "{":
            depth += 1
        elif s[j] == "}":
            depth -= 1
            if depth == 0:
                return j
    raise ValueError("unbalanced")


def numbers(body):
    return [float(x) if re.search(r"[.eE]", x) else int(x)
            for x in re.findall(r"-?\d+(?:\.\d+)?(?:[eE][-+]?\d+)?", body)]


def parse_literals(span):
    """All {"name", Literals<GDF_T>{...}} columns in span, in order."""
    cols = []
    for m in re.finditer(r'\{\s*"([^"]*)"\s*,\s*Literals<(GDF_\w+)>\s*\{', span):
        start = m.end() - 1
        end = balanced(span, start)
        body = span[start + 1:end]
        vm = re.search(r"::vector\s*\{([^}]*)\}", body)
        bm = re.search(r"::bool_vector\s*\{([^}]*)\}", body)
        if vm:
            values = numbers(vm.group(1))
            valid = [bool(x) for x in numbers(bm.group(1))] if bm else None
        else:
            values = numbers(body)
            valid = None
        cols.append({"name": m.group(1), "dtype": m.group(2), "values": values, "valid": valid})
    return cols


def c_strings(span):
    return "".join(bytes(x, "utf-8").decode("unicode_escape") for x in re.findall(r'"((?:[^"\\]|\\.)*)"', span))


def parse_query_tests(path, rel):
    src = open(path).read()
    out = []
    heads = list(re.finditer(r"TEST_F\(EvaluateQueryTest,\s*(\w+)\)", src))
    for k, h in enumerate(heads):
        block = src[h.end(): heads[k + 1].start() if k + 1 < len(heads) else len(src)]
        lp = block.find(".logicalPlan")
        tg = block.find(".tableGroup")
        rt = block.find(".resultTable")
        if min(lp, tg, rt) < 0:
            continue
        line = src[:h.start()].count("\n") + 1
        out.append({
            "source": f"{rel}:{line}",
            "name": h.group(1),
            "plan": c_strings(block[lp:tg]),
            "tables": parse_literals(block[tg:rt]),
            "expected": parse_literals(block[rt:]),
        })
    return out


def parse_where_test(path, rel):
    src = open(path).read()
    table = parse_literals(src[src.find("LiteralTableBuilder{"):src.find("auto group")])
    out = []
    for m in re.finditer(r'Item\{\s*"(\w+)",\s*((?:"(?:[^"\\]|\\.)*"\s*)+),\s*\{([^}]*)\}', src):
        line = src[:m.start()].count("\n") + 1
        out.append({"source": f"{rel}:{line}", "name": m.group(1), "plan": c_strings(m.group(2)),
                    "tables": table,
                    "expected": [{"name": "", "dtype": "GDF_FLOAT64", "values": numbers(m.group(3)),
                                  "valid": None}]})
    return out


def interpreter_plans():
    """interpreter_tests.cu register plans (perform_operation called directly)."""
    rel = "tests/interpreter/interpreter_tests.cu"
    i10 = list(range(10))
    return [
        {"source": f"{rel}:68", "name": "TEST_00",
         "inputs": [{"dtype": "GDF_INT32", "values": [i * 2 for i in i10]},
                    {"dtype": "GDF_INT32", "values": [i * 10 for i in i10]},
                    {"dtype": "GDF_INT32", "values": [i * 20 for i in i10]}],
         "left": [0, 5, 5, 1], "right": [1, 2, 1, -2], "outputs": [5, 5, 3, 4],
         "final_positions": [3, 4], "binops": ["ADD", "MUL", "ADD", "ADD"],
         "unops": ["INVALID"] * 4, "rscalars": [None, None, None, ["GDF_INT32", 2]],
         "lscalars": [None] * 4, "out_dtypes": ["GDF_INT32", "GDF_INT32"],
         "expected": [[0, 250, 980, 2190, 3880, 6050, 8700, 11830, 15440, 19530],
                      [2, 12, 22, 32, 42, 52, 62, 72, 82, 92]], "compare": "exact"},
        {"source": f"{rel}:161", "name": "TEST_FLOATS",
         "inputs": [{"dtype": "GDF_FLOAT32", "values": [1.1, 3.6, 5.3, 7.9, 9.35]},
                    {"dtype": "GDF_FLOAT32", "values": [1, 1, 1, 1, 1]},
                    {"dtype": "GDF_FLOAT32", "values": [0, 2, 4, 6, 8]}],
         "left": [2, 0], "right": [0, 1], "outputs": [3, 4], "final_positions": [3, 4],
         "binops": ["MUL", "ADD"], "unops": ["INVALID"] * 2, "lscalars": [None] * 2,
         "rscalars": [None] * 2, "out_dtypes": ["GDF_FLOAT32", "GDF_FLOAT32"],
         "expected": [[0, 7.2, 21.2, 47.4, 74.8], [2.1, 4.6, 6.3, 8.9, 10.35]],
         "compare": "float32"},
        {"source": f"{rel}:248", "name": "TEST_SCALAR_ADD",
         "inputs": [{"dtype": "GDF_INT32", "values": [1, 3, 5, 7, 9]},
                    {"dtype": "GDF_INT32", "values": [1, 1, 1, 1, 1]},
                    {"dtype": "GDF_INT32", "values": [0, 2, 4, 6, 8]}],
         "left": [0], "right": [-2], "outputs": [3], "final_positions": [3], "binops": ["ADD"],
         "unops": ["INVALID"], "lscalars": [None], "rscalars": [["GDF_INT32", 1]],
         "out_dtypes": ["GDF_INT32"], "expected": [[2, 4, 6, 8, 10]], "compare": "exact"},
        {"source": f"{rel}:334", "name": "TEST_01_Unary_Operators",
         "inputs": [{"dtype": "GDF_FLOAT32", "values": [1.1, 3.6, 5.3, 7.9, 9.35]},
                    {"dtype": "GDF_FLOAT32", "values": [0, 2, 4, 6, 8]},
                    {"dtype": "GDF_FLOAT32", "values": [0, 2, 4, 6, 8]}],
         "left": [0, 1], "right": [-1, -1], "outputs": [3, 4], "final_positions": [3, 4],
         "binops": ["INVALID", "INVALID"], "unops": ["FLOOR", "SIN"], "lscalars": [None] * 2,
         "rscalars": [None] * 2, "out_dtypes": ["GDF_FLOAT32", "GDF_FLOAT32"],
         "expected": [[1, 3, 5, 7, 9], [0, 0.9092974268, -0.7568024953, -0.2794154982, 0.9893582466]],
         "compare": "float32"},
        {"source": f"{rel}:434", "name": "TEST_02",
         "inputs": [{"dtype": "GDF_FLOAT32", "values": [(i + 1) * 2.0 for i in i10]},
                    {"dtype": "GDF_FLOAT32", "values": [(i + 1) * 10.0 for i in i10]},
                    {"dtype": "GDF_INT32", "values": [(i + 1) * 20 for i in i10]}],
         "left": [0, 5, 5, 1, 5], "right": [1, 2, 1, -1, -2], "outputs": [5, 5, 3, 5, 4],
         "final_positions": [3, 4], "binops": ["ADD", "MUL", "ADD", "INVALID", "ADD"],
         "unops": ["INVALID", "INVALID", "INVALID", "SIN", "INVALID"], "lscalars": [None] * 5,
         "rscalars": [None, None, None, None, ["GDF_FLOAT32", 2.33]],
         "out_dtypes": ["GDF_FLOAT32", "GDF_FLOAT32"],
         "expected": [[250.0, 980.0, 2190.0, 3880.0, 6050.0, 8700.0, 11830.0, 15440.0, 19530.0, 24100.0],
                      [1.78598, 3.24295, 1.34197, 3.07511, 2.06763, 2.02519, 3.10389, 1.33611, 3.224,
                       1.82363]],
         "compare": "6_significant"},
    ]


def main(ref):
    data = {
        "evaluate_query": parse_query_tests(os.path.join(ref, "tests/evaluate_query/evaluate_query.cu"),
                                            "tests/evaluate_query/evaluate_query.cu"),
        "unary": parse_query_tests(os.path.join(ref, "tests/unary/unary-tests.cu"), "tests/unary/unary-tests.cu"),
        "where": parse_where_test(os.path.join(ref, "tests/where/where-test.cpp"), "tests/where/where-test.cpp"),
        "interpreter": interpreter_plans(),
    }
    out = os.path.join(HERE, "reference_goldens.json")
    with open(out, "w") as f:
        json.dump(data, f, indent=1)
    print(out, {k: len(v) for k, v in data.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")

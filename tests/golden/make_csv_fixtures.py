"""Writes tests/golden/tpch_csv_goldens.json from the reference's own CSV test.

Run in the development container (it reads the reference tree as text; nothing of
the reference is imported or executed):

    python tests/golden/make_csv_fixtures.py /path/to/reference

tests/tpch/tpch_csv_tests.cu writes /tmp/customer.psv (150 TPC-H customer rows,
pipe-delimited) in its SetUp, loads it with the column types each test names
(cuDF's reader under csv_parser) and runs a filter + project plan. This script
keeps the file's text and, for the tests whose plan is a project over a filter of
that table (TEST_00, TEST_01), the plan, the column types and the expected result
table. The JSON holds data only.
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref):
    src = open(os.path.join(ref, "tests", "tpch", "tpch_csv_tests.cu")).read()
    m = re.search(r'outfile\("/tmp/customer\.psv".*?R"\((.*?)\)"', src, re.S)
    customer = m.group(1) + "\n"  # the test writes the text, then std::endl
    tests = []
    for tm in re.finditer(r"TEST_F\(EvaluateQueryTest, (TEST_\d+)\)", src):
        name = tm.group(1)
        body = src[tm.end():]
        end = body.find("TEST_F(")
        body = body if end < 0 else body[:end]
        if '"/tmp/customer.psv"' not in body:
            continue
        plan = "".join(re.findall(r'"((?:[^"\\]|\\.)*)"', re.search(r"\.logicalPlan\s*=(.*?),\s*\.filePaths",
                                                                     body, re.S).group(1)))
        plan = plan.replace("\\n", "\n")
        lines = [ln.strip() for ln in plan.split("\n") if ln.strip()]
        if not (len(lines) == 3 and lines[0].startswith("LogicalProject(") and lines[1].startswith("LogicalFilter(")
                and lines[2].startswith("EnumerableTableScan")):
            continue
        types = re.findall(r'"(\w+)"', re.search(r"\.columnTypes\s*=\s*\{\{(.*?)\}\}", body, re.S).group(1))
        expected = []
        for cm in re.finditer(r'\{\s*"(GDF_\w+)"\s*,\s*Literals<GDF_\w+>\s*\{([^}]*)\}', body):
            vals = [float(x) if re.search(r"[.eE]", x) else int(x)
                    for x in re.findall(r"-?\d+(?:\.\d+)?(?:[eE][-+]?\d+)?", cm.group(2))]
            # the float literals as written (the C++ test converts them to float)
            texts = re.findall(r"-?\d+(?:\.\d+)?(?:[eE][-+]?\d+)?", cm.group(2))
            expected.append({"dtype": cm.group(1), "values": vals, "text": texts})
        tests.append({"name": name, "project": lines[0], "filter": lines[1], "column_types": types,
                      "expected": expected, "source": "tests/tpch/tpch_csv_tests.cu"})
    out = {"customer_psv": customer, "customer_source": "tests/tpch/tpch_csv_tests.cu (SetUp)", "tests": tests}
    path = os.path.join(HERE, "tpch_csv_goldens.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {path}: {customer.count(chr(10))} rows, {len(tests)} tests")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")

"""The planner's number and date literal matchers against std::regex with the
reference's patterns (CalciteExpressionParsing.cpp:274, :324-325, :585): every
string up to length 6 over a small alphabet and random longer ones agree."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_literal_matchers_agree_with_std_regex():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "literal_matchers_test")
        subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "blazingdb-ral_amd", "csrc"),
                        "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "include", "blz"),
                        os.path.join(ROOT, "tests", "cpp", "literal_matchers_test.cpp"), "-o", exe],
                       check=True, capture_output=True, timeout=300)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout

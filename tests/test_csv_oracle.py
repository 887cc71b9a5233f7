"""Pins oracle/csv_oracle.py (the delimited-text reader's CPU restatement) without
the GPU: row and field splitting against Python's csv module, reals on the
correctly rounded path against float() and an exact binary32 rounding, integers
against int(), dates against datetime, and the whole reader, through the oracle's
filter and project, against the reference's expected tables for the customer
file of tests/tpch/tpch_csv_tests.cu."""
import csv
import datetime as dt
import io
import json
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import csv_oracle as CO
from oracle import planner as P
from oracle import pyoracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tpch_csv_goldens.json")
TYPES = {"int32": P.GDF_INT32, "int64": P.GDF_INT64, "float32": P.GDF_FLOAT32, "float64": P.GDF_FLOAT64}


def gen_text(rng, n_rows, n_cols, crlf=False, blank_p=0.05, quote_p=0.1, trailing=True):
    """Rows of words, numbers and quoted fields (quotes wrap whole fields and may
    hold the delimiter and line ends)."""
    words = ["alpha", "beta gamma", "x,y", "12", "-7.25", "", "3e5", "Customer#000000001"]
    eol = "\r\n" if crlf else "\n"
    lines = []
    for _ in range(n_rows):
        if rng.random() < blank_p:
            lines.append("")
        fields = []
        for _ in range(n_cols):
            w = words[int(rng.integers(0, len(words)))]
            if rng.random() < quote_p:
                inner = w + ("|" if rng.random() < 0.5 else "") + ("\n" if rng.random() < 0.3 else "")
                w = '"' + inner + '"'
            fields.append(w)
        lines.append("|".join(fields))
    text = eol.join(lines)
    return text + eol if trailing else text


def oracle_rows(data):
    out = []
    for st in CO.row_starts(data):
        out.append([data[a:e].decode() for a, e in CO.row_fields(data, st)])
    return out


def csv_rows(text):
    r = csv.reader(io.StringIO(text, newline=""), delimiter="|", quotechar='"', doublequote=False)
    return [row for row in r if row]


@pytest.mark.parametrize("seed", range(12))
def test_rows_and_fields_match_csv_module(seed):
    rng = np.random.default_rng(seed)
    text = gen_text(rng, int(rng.integers(1, 60)), int(rng.integers(1, 6)), crlf=seed % 3 == 0,
                    trailing=seed % 2 == 0)
    assert oracle_rows(text.encode()) == csv_rows(text)


def test_blank_lines_and_edges():
    assert CO.row_starts(b"") == []
    assert CO.row_starts(b"\n\n\r\n") == []
    assert CO.row_starts(b"a\n\nb") == [0, 3]
    assert CO.row_starts(b'"x\ny"|1\n2') == [0, 8]  # a line end inside quotes is data
    assert CO.row_fields(b"1|2", 0) == [(0, 1), (2, 3)]
    assert CO.row_fields(b"1|2|", 0) == [(0, 1), (2, 3), (4, 4)]


def _decimal_strings(rng, n, max_digits, emax):
    out = []
    for _ in range(n):
        nd = int(rng.integers(1, max_digits + 1))
        digits = "".join(str(int(d)) for d in rng.integers(0, 10, nd))
        if digits[0] == "0" and nd > 1:
            digits = str(int(rng.integers(1, 10))) + digits[1:]
        pos = int(rng.integers(0, nd + 1))
        s = digits[:pos] + ("." + digits[pos:] if pos < nd else "")
        if rng.random() < 0.3:
            s += "e" + str(int(rng.integers(-emax, emax + 1)))
        if rng.random() < 0.5:
            s = "-" + s
        out.append(s)
    return out


def test_binary64_correctly_rounded_path_equals_float():
    rng = np.random.default_rng(5)
    n_fast = 0
    for s in _decimal_strings(rng, 20000, 15, 6) + ["0.1", "711.56", "-272.6", "1e22", "9007199254740993",
                                                     "0.000001", "123456789012345e-22"]:
        neg, m, e = CO.scan_real(s.encode())
        if m <= (1 << 53) and -22 <= e <= 22:
            n_fast += 1
            assert CO.real_f64(neg, m, e) == float(s), s
    assert n_fast > 15000


def test_binary64_slow_path_within_a_few_ulp():
    rng = np.random.default_rng(6)
    for s in _decimal_strings(rng, 5000, 19, 40):
        neg, m, e = CO.scan_real(s.encode())
        got, want = CO.real_f64(neg, m, e), float(s)
        if want == 0 or not np.isfinite(want):
            continue
        assert abs(got - want) <= 4 * np.spacing(abs(want)), s


def _f32_exact(s):
    """The binary32 nearest to decimal s (ties to even), exactly."""
    x = Fraction(s)
    c = np.float32(float(s))
    best = None
    for cand in (np.nextafter(c, np.float32(-np.inf)), c, np.nextafter(c, np.float32(np.inf))):
        d = abs(Fraction(float(cand)) - x)
        key = (d, int(np.array(cand).view(np.uint32)) & 1)
        if best is None or key < best[0]:
            best = (key, cand)
    return best[1]


def test_binary32_correctly_rounded_path_is_exact():
    rng = np.random.default_rng(7)
    n_fast = 0
    for s in _decimal_strings(rng, 8000, 7, 4) + ["711.56", "7498.12", "-272.6", "16777216", "0.3"]:
        neg, m, e = CO.scan_real(s.encode())
        if m <= (1 << 24) and -10 <= e <= 10:
            n_fast += 1
            got = CO.real_f32(neg, m, e)
            want = _f32_exact(s)
            assert np.array(got).view(np.uint32) == np.array(want).view(np.uint32), s
    assert n_fast > 5000


def test_integers_match_int():
    rng = np.random.default_rng(8)
    for _ in range(5000):
        v = int(rng.integers(-(1 << 62), 1 << 62)) * int(rng.integers(1, 8))
        s = str(v)
        w = v % (1 << 64)
        w = w - (1 << 64) if w >= (1 << 63) else w
        assert CO.parse_int(s.encode()) == w
        for dtp, bits in ((P.GDF_INT8, 8), (P.GDF_INT16, 16), (P.GDF_INT32, 32), (P.GDF_INT64, 64)):
            t = v % (1 << bits)
            t = t - (1 << bits) if t >= (1 << (bits - 1)) else t
            assert CO.parse_value(s.encode(), dtp) == t
    assert CO.parse_int(b"25-989-741-2988") == 25 and CO.parse_int(b"Customer#1") == 0
    assert CO.parse_int(b"+17x") == 17 and CO.parse_int(b"") == 0


def test_dates_match_datetime():
    rng = np.random.default_rng(9)
    epoch = dt.datetime(1970, 1, 1)
    for _ in range(3000):
        d = epoch + dt.timedelta(days=int(rng.integers(-200000, 200000)),
                                 milliseconds=int(rng.integers(0, 86400000)))
        if d.year < 1:
            continue
        day = d.strftime("%Y-%m-%d")
        s = day + "T" + d.strftime("%H:%M:%S") + ".%03dZ" % (d.microsecond // 1000)
        days, ms = CO.parse_ms(s.encode())
        want_days = (dt.datetime(d.year, d.month, d.day) - epoch).days
        assert days == want_days and ms == (d - epoch) // dt.timedelta(milliseconds=1), s
        assert CO.parse_ms(day.encode())[1] == want_days * 86400000
        assert CO.parse_ms((day + " " + d.strftime("%H:%M:%S")).encode())[1] == \
            (d.replace(microsecond=0) - epoch) // dt.timedelta(milliseconds=1)


def _golden_frame(g, test):
    """The customer table read with the test's column types; string columns are
    not read (no plan references them) and stand as zero columns."""
    data = g["customer_psv"].encode()
    types = [TYPES.get(t, P.GDF_STRING) for t in test["column_types"]]
    want = [i for i, t in enumerate(types) if t != P.GDF_STRING]
    cols = CO.read_csv(data, types, want)
    n = len(cols[0].data)
    frame = []
    it = iter(cols)
    for t in types:
        frame.append(next(it) if t != P.GDF_STRING else O.HostColumn(np.zeros(n, np.int64), P.GDF_INT64))
    return frame


@pytest.mark.parametrize("name", ["TEST_00", "TEST_01"])
def test_reference_customer_file_through_filter_and_project(name):
    g = json.load(open(GOLD))
    test = next(t for t in g["tests"] if t["name"] == name)
    frame = _golden_frame(g, test)
    assert len(frame[0].data) == 150
    out = O.execute_project(O.process_filter(frame, test["filter"]), test["project"])
    assert len(out) == len(test["expected"])
    for col, exp in zip(out, test["expected"]):
        if exp["dtype"] == "GDF_FLOAT32":
            want = np.array([_f32_exact(t) for t in exp["text"]], dtype=np.float32)
            assert col.data.dtype == np.float32 and col.data.tobytes() == want.tobytes()
        else:
            assert col.data.tolist() == exp["values"]


def test_missing_fields_and_skip_rows():
    data = b"1|2.5|1996-01-02\n3\n\n4|x|1970-01-01T00:00:01.5Z\n"
    cols = CO.read_csv(data, [P.GDF_INT32, P.GDF_FLOAT64, P.GDF_DATE64])
    assert cols[0].data.tolist() == [1, 3, 4]
    assert cols[1].data.tolist() == [2.5, 0.0, 0.0]
    assert cols[2].data.tolist() == [(dt.date(1996, 1, 2) - dt.date(1970, 1, 1)).days * 86400000, 0, 1500]
    cols = CO.read_csv(data, [P.GDF_INT32, P.GDF_FLOAT64, P.GDF_DATE64], [2, 0], skip_rows=1)
    assert cols[1].data.tolist() == [3, 4] and cols[0].data.tolist() == [0, 1500]

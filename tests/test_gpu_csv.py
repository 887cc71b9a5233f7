"""Ingest on the GPU: delimited text to device columns (csv.hip) against the oracle
(oracle/csv_oracle.py) on the same bytes, bit for bit; the reference's customer
file (tests/tpch/tpch_csv_tests.cu) read and run through the GPU filter and
project against the reference's expected tables; a generated file at a few
million rows against the values it was written from."""
import datetime as dt
import json
import os

import numpy as np
import pytest

from oracle import csv_oracle as CO
from oracle import planner as P
from tests.test_csv_oracle import GOLD, TYPES, _f32_exact, gen_text

pytestmark = pytest.mark.gpu

ALL = [P.GDF_INT8, P.GDF_INT16, P.GDF_INT32, P.GDF_INT64, P.GDF_FLOAT32, P.GDF_FLOAT64, P.GDF_DATE32,
       P.GDF_DATE64, P.GDF_TIMESTAMP]


@pytest.fixture(scope="module")
def B():
    import blz_ral
    if blz_ral.device_count() < 1:
        pytest.fail("no HIP device visible")
    return blz_ral


def _check(B, data, dtypes, cols=None, skip=0, what=""):
    ref = CO.read_csv(data, dtypes, cols, skip_rows=skip)
    got, n = B.read_csv(data, dtypes, cols, skip_rows=skip)
    assert n == (len(ref[0].data) if ref else n), what
    for g, r in zip(got, ref):
        assert g.col.dtype == r.dtype and g.col.size == len(r.data), what
        gd = g.data()
        assert gd.tobytes() == r.data.tobytes(), (what, np.nonzero(gd != r.data)[0][:5])
    return got


def _value_text(rng, dtp):
    if dtp in (P.GDF_INT8, P.GDF_INT16, P.GDF_INT32, P.GDF_INT64):
        return str(int(rng.integers(-(1 << 40), 1 << 40)))
    if dtp in (P.GDF_FLOAT32, P.GDF_FLOAT64):
        k = rng.random()
        if k < 0.6:
            return f"{rng.uniform(-1e5, 1e5):.{int(rng.integers(0, 7))}f}"
        if k < 0.8:
            return repr(float(rng.standard_normal() * 10 ** int(rng.integers(-30, 30))))
        return str(int(rng.integers(-1000, 1000))) + "e" + str(int(rng.integers(-25, 25)))
    d = dt.datetime(1970, 1, 1) + dt.timedelta(seconds=int(rng.integers(-2 * 10 ** 9, 2 * 10 ** 9)))
    return d.strftime("%Y-%m-%dT%H:%M:%S") + ".%03dZ" % int(rng.integers(0, 1000))


def _typed_text(rng, dtypes, n_rows, crlf=False, quote_p=0.05, blank_p=0.02, drop_p=0.02):
    eol = "\r\n" if crlf else "\n"
    lines = []
    for _ in range(n_rows):
        if rng.random() < blank_p:
            lines.append("")
        f = []
        for dtp in dtypes:
            v = _value_text(rng, dtp) if dtp != P.GDF_STRING else "str|ing\nx" if rng.random() < 0.3 else "text"
            if dtp == P.GDF_STRING or rng.random() < quote_p:
                v = '"' + v + '"'
            f.append(v)
        if rng.random() < drop_p:
            f = f[:int(rng.integers(1, len(f) + 1))]
        lines.append("|".join(f))
    return (eol.join(lines) + eol).encode()


@pytest.mark.parametrize("seed", range(6))
def test_typed_columns_match_oracle(B, seed):
    rng = np.random.default_rng(60_000 + seed)
    dtypes = [ALL[int(i)] for i in rng.integers(0, len(ALL), int(rng.integers(2, 9)))]
    dtypes.insert(int(rng.integers(0, len(dtypes))), P.GDF_STRING)  # a string column, skipped
    want = [i for i, t in enumerate(dtypes) if t != P.GDF_STRING]
    rng.shuffle(want)
    n = int(rng.integers(1, 20_000))
    data = _typed_text(rng, dtypes, n, crlf=seed % 2 == 1)
    print(f"CASE csv typed seed={seed} rows={n} bytes={len(data)} dtypes={dtypes}", flush=True)
    _check(B, data, dtypes, want, what=f"seed {seed}")
    _check(B, data, dtypes, want[:1], skip=min(3, n), what=f"seed {seed} skip")


@pytest.mark.parametrize("seed", range(4))
def test_quoted_fields_and_blank_lines_match_oracle(B, seed):
    """Quoted fields holding delimiters and line ends, blank lines, CRLF, no
    trailing line end: the rows and the integer column they carry."""
    rng = np.random.default_rng(61_000 + seed)
    text = gen_text(rng, 3000, 4, crlf=seed % 2 == 0, blank_p=0.1, quote_p=0.3, trailing=seed < 2)
    data = text.encode()
    print(f"CASE csv quoted seed={seed} bytes={len(data)}", flush=True)
    _check(B, data, [P.GDF_INT64, P.GDF_FLOAT64, P.GDF_INT32, P.GDF_FLOAT32], what=f"quoted {seed}")


def test_rows_across_chunk_boundaries(B):
    """Rows and quoted fields longer than the 4 KiB chunks and 64-byte lanes the
    row pass works in, and rows that start on every lane offset."""
    rng = np.random.default_rng(62_000)
    lines = []
    for i in range(400):
        pad = "x" * int(rng.integers(0, 9000 if i % 7 == 0 else 130))
        q = '"' + pad[: len(pad) // 2] + "\n|\n" + pad[len(pad) // 2:] + '"' if i % 5 == 0 else pad
        lines.append(f"{i}|{q}|{i * 3}.5")
    data = ("\n".join(lines) + "\n").encode()
    print(f"CASE csv chunk_boundaries bytes={len(data)}", flush=True)
    got = _check(B, data, [P.GDF_INT32, P.GDF_STRING, P.GDF_FLOAT64], [0, 2], what="chunks")
    assert got[0].data().tolist() == list(range(400))


def test_edge_inputs(B):
    for data in (b"", b"\n\n\r\n", b"7", b"7|", b"|\n|\n", b"1\n2\n3"):
        print(f"CASE csv edge {data!r}", flush=True)
        _check(B, data, [P.GDF_INT64, P.GDF_FLOAT64], what=repr(data))
    with pytest.raises(B.BlzError, match="not supported"):
        B.read_csv(b"a|b\n", [P.GDF_STRING, P.GDF_INT32])
    with pytest.raises(B.BlzError, match="out of range"):
        B.read_csv(b"1|2\n", [P.GDF_INT32, P.GDF_INT32], [2])


@pytest.mark.parametrize("name", ["TEST_00", "TEST_01"])
def test_reference_customer_file_filter_project(B, name):
    """tpch_csv_tests.cu: the customer file read with the test's column types
    (string columns not read), then its filter and project on the GPU: the
    reference's expected tables, bit for bit."""
    g = json.load(open(GOLD))
    test = next(t for t in g["tests"] if t["name"] == name)
    types = [TYPES.get(t, P.GDF_STRING) for t in test["column_types"]]
    want = [i for i, t in enumerate(types) if t != P.GDF_STRING]
    print(f"CASE csv golden {name}", flush=True)
    cols, n = B.read_csv(g["customer_psv"].encode(), types, want)
    assert n == 150
    frame, it = [], iter(cols)
    for t in types:
        frame.append(next(it) if t != P.GDF_STRING else B.DeviceColumn.from_numpy(np.zeros(n, np.int64)))
    out, kept = B.filter_project(frame, test["filter"], test["project"])
    assert kept == len(test["expected"][0]["values"])
    for col, exp in zip(out, test["expected"]):
        d = col.data()
        if exp["dtype"] == "GDF_FLOAT32":
            want32 = np.array([_f32_exact(t) for t in exp["text"]], dtype=np.float32)
            assert d.dtype == np.float32 and d.tobytes() == want32.tobytes()
        else:
            assert d.tolist() == exp["values"]


def test_generated_file_round_trip(B, tmp_path):
    """Three million rows written from known values (int32, cents as binary64,
    dates), read back from device bytes and, split in two files, through the
    loader: every value equals the one it was written from."""
    n = 3_000_000
    rng = np.random.default_rng(63_000)
    a = rng.integers(-2 ** 31, 2 ** 31, n).astype(np.int32)
    cents = rng.integers(-10 ** 9, 10 ** 9, n)
    days = rng.integers(-25_000, 25_000, n)
    t_int = a.astype(str)
    t_real = np.char.add(np.char.add(np.where(cents < 0, "-", ""), (np.abs(cents) // 100).astype(str)),
                         np.char.add(".", np.char.zfill((np.abs(cents) % 100).astype(str), 2)))
    t_date = np.datetime_as_string(days.astype("datetime64[D]"))
    lines = np.char.add(np.char.add(np.char.add(t_int, "|"), np.char.add(t_real, "|")), t_date)
    data = ("\n".join(lines.tolist()) + "\n").encode()
    print(f"CASE csv round_trip rows={n} bytes={len(data)}", flush=True)
    dtypes = [P.GDF_INT32, P.GDF_FLOAT64, P.GDF_DATE32]
    buf = B.DeviceBuffer(len(data))
    buf.upload(np.frombuffer(data, np.uint8))
    cols, rows = B.read_csv((buf.value, len(data)), dtypes, on_device=True)
    assert rows == n
    assert cols[0].data().tobytes() == a.tobytes()
    assert cols[1].data().tobytes() == (cents / 100.0).tobytes()  # cents < 2^53, one correctly rounded divide
    assert cols[2].data().tobytes() == days.astype(np.int32).tobytes()
    half = data.index(b"\n", len(data) // 2) + 1
    p1, p2 = tmp_path / "part1.psv", tmp_path / "part2.psv"
    p1.write_bytes(data[:half])
    p2.write_bytes(data[half:])
    cols2, rows2 = B.load_csv_files([str(p1), str(p2)], dtypes, [2, 0])
    assert rows2 == n
    assert cols2[0].data().tobytes() == days.astype(np.int32).tobytes()
    assert cols2[1].data().tobytes() == a.tobytes()


def test_cpp_loader_filter_project_and_result_repository(B, tmp_path):
    """The C++ operator API (ral.hpp): data_loader over csv_parser reads the
    reference's customer file (TEST_01 types), process_filter and
    execute_project_plan run TEST_01, the frame is parked in the result repository
    and read back: the reference's expected table. The same file loaded twice by
    one data_loader gives twice the rows."""
    import subprocess
    g = json.load(open(GOLD))
    test = next(t for t in g["tests"] if t["name"] == "TEST_01")
    path = tmp_path / "customer.psv"
    path.write_text(g["customer_psv"])
    shim = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "ral_shim_test")
    print("CASE csv cpp loader", flush=True)
    r = subprocess.run([shim, "csv", str(path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    lines = r.stdout.splitlines()
    assert "rows 150" in lines
    assert "rows_twice 300 c_custkey c_acctbal" in lines
    assert lines[lines.index("rows 150") + 2] == "duration 1.5 error ''"
    cols = {ln.split()[1]: ln.split()[2:] for ln in lines if ln.startswith("col ")}
    assert [int(v) for v in cols["c_custkey"]] == test["expected"][0]["values"]
    assert [int(v) for v in cols["c_nationkey"]] == test["expected"][1]["values"]
    got = np.array([float.fromhex(v) for v in cols["c_acctbal"]], dtype=np.float32)
    want = np.array([_f32_exact(t) for t in test["expected"][2]["text"]], dtype=np.float32)
    assert got.tobytes() == want.tobytes()
    assert "freed 1" in lines
    assert any("Result set does not exist" in ln for ln in lines)

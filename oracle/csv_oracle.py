"""CPU restatement of the delimited-text reader (test infrastructure only).

Only tests and the benchmark's CPU baseline use this module; the product reads
text on the GPU (blazingdb-ral_amd/csrc/csv.hip) and never calls it.

The reference reads a table with csv_parser::parse (io/data_parser/CSVParser.cpp:
178-238) over cuDF's read_csv with the options of init_default_csv_args (:26-47):
one-byte delimiter and line terminator, quote '"' with minimal quoting and no
doubled quotes, blank lines skipped, no header, no NA detection. cuDF (version
not pinned by the reference, cmake/ImportCMakeModules.cmake) is absent here; this
module states the reader's rules, row by row:

  rows      a row starts at byte 0 or after a line terminator outside quotes
            (quotes toggle the state); a row that is empty or only "\\r" is
            blank and skipped; skip_rows rows are dropped first.
  fields    split at delimiters outside quotes, up to the row's terminator;
            a "\\r" before the terminator is dropped; a field that starts with a
            quote loses it and a closing quote; a missing field reads as 0.
  integers  sign, digits to the first other byte, wrapping 64-bit, truncated.
  binary64  up to 19 significant digits m, decimal exponent e; m <= 2^53 and
            |e| <= 22: one IEEE multiply or divide by 10^|e| (correctly rounded);
            else steps of 10^22 then 10^|e|.
  binary32  m <= 2^24 and |e| <= 10: one binary32 operation; else the binary64
            value rounded to binary32.
  dates     YYYY-MM-DD[(T| )hh:mm:ss[.fff]][Z]: days (DATE32) or milliseconds
            (DATE64, TIMESTAMP) since 1970-01-01, days_from_civil (Hinnant).

Pinned by tests/test_csv_oracle.py: Python's csv module (row and field
splitting), float() and an exact binary32 rounding (reals on the correctly
rounded path), int() and datetime (integers, dates), and the reference's own
expected tables for the customer file of tests/tpch/tpch_csv_tests.cu.
"""
import numpy as np

from oracle import planner as P
from oracle.pyoracle import HostColumn

NP_OF = {P.GDF_INT8: np.int8, P.GDF_INT16: np.int16, P.GDF_INT32: np.int32, P.GDF_INT64: np.int64,
         P.GDF_FLOAT32: np.float32, P.GDF_FLOAT64: np.float64, P.GDF_DATE32: np.int32,
         P.GDF_DATE64: np.int64, P.GDF_TIMESTAMP: np.int64}
P10 = [float(10 ** k) for k in range(23)]
P10F = [np.float32(10 ** k) for k in range(11)]


def row_starts(data: bytes, term=b"\n", quote=b'"'):
    """Byte offsets of the non-blank rows (csv.hip: csv_count / csv_rows)."""
    t, q = term[0], quote[0]
    n = len(data)
    out = []
    inq = 0
    prev = t
    for k in range(n):
        c = data[k]
        nx_end = k + 1 >= n
        blank = c == t or (c == 13 and (nx_end or data[k + 1] == t))
        if prev == t and not blank and inq == 0:
            out.append(k)
        if c == q:
            inq ^= 1
        prev = c
    return out


def row_fields(data: bytes, start, delim=b"|", term=b"\n", quote=b'"', last_col=None):
    """Field byte ranges [(a, e)] of the row at `start` (csv_fields_kernel), at most
    last_col + 1 of them; a row with fewer fields has fewer."""
    d, t, q = delim[0], term[0], quote[0]
    n = len(data)
    out = []
    a = start
    inq = False
    while True:
        e = a
        while e < n:
            c = data[e]
            if c == q:
                inq = not inq
            elif not inq and (c == d or c == t):
                break
            e += 1
        row_end = e >= n or data[e] == t
        fa, fe = a, e
        if fe > fa and data[fe - 1] == 13 and row_end:
            fe -= 1
        if fe > fa and data[fa] == q:
            fa += 1
            if fe > fa and data[fe - 1] == q:
                fe -= 1
        if a >= n:
            fe = fa
        out.append((fa, fe))
        if row_end or (last_col is not None and len(out) > last_col):
            return out
        a = e + 1


def parse_int(s: bytes) -> int:
    i, neg = 0, False
    if s[:1] in (b"-", b"+"):
        neg = s[:1] == b"-"
        i = 1
    v = 0
    while i < len(s) and 48 <= s[i] <= 57:
        v = (v * 10 + s[i] - 48) % (1 << 64)
        i += 1
    if neg:
        v = (-v) % (1 << 64)
    return v - (1 << 64) if v >= (1 << 63) else v


def scan_real(s: bytes):
    """(neg, m, e): sign, up to 19 significant digits, decimal exponent."""
    i, neg = 0, False
    if s[:1] in (b"-", b"+"):
        neg = s[:1] == b"-"
        i = 1
    m, e, nd, dot = 0, 0, 0, False
    while i < len(s):
        c = s[i]
        if c == 46 and not dot:
            dot = True
            i += 1
            continue
        if not 48 <= c <= 57:
            break
        d = c - 48
        if nd == 0 and d == 0:
            if dot:
                e -= 1
        elif nd < 19:
            m = m * 10 + d
            nd += 1
            if dot:
                e -= 1
        elif not dot:
            e += 1
        i += 1
    if i < len(s) and s[i] in (101, 69):
        i += 1
        eneg = False
        if i < len(s) and s[i] in (43, 45):
            eneg = s[i] == 45
            i += 1
        x = 0
        while i < len(s) and 48 <= s[i] <= 57:
            if x < 100000:
                x = x * 10 + s[i] - 48
            i += 1
        e += -x if eneg else x
    return neg, m, e


def real_f64(neg, m, e) -> float:
    v = float(m)
    if m != 0:
        if m <= (1 << 53) and -22 <= e <= 22:
            v = v / P10[-e] if e < 0 else v * P10[e]
        else:
            with np.errstate(all="ignore"):
                w = np.float64(v)
                while e > 22:
                    w = w * np.float64(1e22)
                    e -= 22
                while e < -22:
                    w = w / np.float64(1e22)
                    e += 22
                w = w / np.float64(P10[-e]) if e < 0 else w * np.float64(P10[e])
            v = float(w)
    return -v if neg else v


def real_f32(neg, m, e) -> np.float32:
    if m != 0 and m <= (1 << 24) and -10 <= e <= 10:
        v = np.float32(m)
        v = v / P10F[-e] if e < 0 else v * P10F[e]
        return -v if neg else v
    with np.errstate(all="ignore"):
        return np.float32(real_f64(neg, m, e))


def tdiv(a, b):
    """C++ integer division (truncation toward zero)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def days_from_civil(y, m, d):
    y -= 1 if m <= 2 else 0
    era = tdiv(y if y >= 0 else y - 399, 400)
    yoe = y - era * 400
    mp = (m + 9) % 12
    doy = tdiv(153 * mp + 2, 5) + d - 1
    doe = yoe * 365 + tdiv(yoe, 4) - tdiv(yoe, 100) + doy
    return era * 146097 + doe - 719468


def parse_ms(s: bytes):
    """(days, milliseconds) of YYYY-MM-DD[(T| )hh:mm:ss[.fff]][Z]."""
    i = 0

    def digits(w):
        nonlocal i
        v = 0
        for _ in range(w):
            if i < len(s) and 48 <= s[i] <= 57:
                v = v * 10 + s[i] - 48
                i += 1
            else:
                break
        return v

    def take(c):
        nonlocal i
        if i < len(s) and s[i] == c:
            i += 1
            return True
        return False

    y = digits(4)
    take(45)
    mo = digits(2)
    take(45)
    d = digits(2)
    days = days_from_civil(y, mo, d)
    ms = 0
    if take(84) or take(32):
        hh = digits(2)
        take(58)
        mi = digits(2)
        take(58)
        ss = digits(2)
        frac = 0
        if take(46):
            nd = 0
            while i < len(s) and nd < 3 and 48 <= s[i] <= 57:
                frac = frac * 10 + s[i] - 48
                i += 1
                nd += 1
            frac *= 10 ** (3 - nd)
        ms = ((hh * 60 + mi) * 60 + ss) * 1000 + frac
    return days, days * 86400000 + ms


def parse_value(s: bytes, dtype):
    if dtype in (P.GDF_INT8, P.GDF_INT16, P.GDF_INT32, P.GDF_INT64):
        v = parse_int(s)
        bits = np.dtype(NP_OF[dtype]).itemsize * 8
        v &= (1 << bits) - 1
        return v - (1 << bits) if v >= (1 << (bits - 1)) else v
    if dtype == P.GDF_FLOAT64:
        return real_f64(*scan_real(s))
    if dtype == P.GDF_FLOAT32:
        return real_f32(*scan_real(s))
    days, ms = parse_ms(s)
    if dtype == P.GDF_DATE32:
        v = days & 0xFFFFFFFF
        return v - (1 << 32) if v >= (1 << 31) else v
    return ms


def read_csv(data: bytes, dtypes, column_indices=None, delimiter="|", line_terminator="\n", skip_rows=0):
    """[HostColumn] of the requested columns (no validity), as blz_csv_read."""
    cols = list(range(len(dtypes))) if column_indices is None else list(column_indices)
    term, delim = line_terminator.encode(), delimiter.encode()
    starts = row_starts(data, term)[skip_rows:]
    last = max(cols) if cols else -1
    vals = {c: [] for c in cols}
    for st in starts:
        fields = row_fields(data, st, delim, term, last_col=last)
        for c in cols:
            if c < len(fields):
                a, e = fields[c]
                vals[c].append(parse_value(data[a:e], dtypes[c]))
            else:
                vals[c].append(0)
    return [HostColumn(np.array(vals[c], dtype=NP_OF[dtypes[c]]), dtypes[c]) for c in cols]

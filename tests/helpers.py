"""Shared test helpers: seeded frames mirrored on host (oracle) and device (HIP path)."""
import numpy as np

from oracle import planner as P
from oracle import pyoracle as O

DT_NP = O.NP_OF


def gen_column(rng, dtype, n, null_rate=0.0, lo=None, hi=None):
    if dtype in (P.GDF_FLOAT32, P.GDF_FLOAT64):
        lo = -100.0 if lo is None else lo
        hi = 100.0 if hi is None else hi
        data = rng.uniform(lo, hi, n).astype(DT_NP[dtype])
    else:
        info = np.iinfo(DT_NP[dtype])
        lo = max(info.min, -1000) if lo is None else lo
        hi = min(info.max, 1000) if hi is None else hi
        data = rng.integers(lo, hi + 1, n, dtype=np.int64).astype(DT_NP[dtype])
    valid = None
    if null_rate > 0:
        valid = rng.random(n) >= null_rate
    return O.HostColumn(data, dtype, valid)


def to_device(B, hc):
    return B.DeviceColumn.from_numpy(hc.data, hc.dtype, valid=hc.valid, null_count=hc.null_count)


def assert_column_equal(gpu_col, host_col, check_valid=True, what=""):
    """Bit-exact data for valid rows (null rows' data is compared too, it is defined),
    validity over bits < n, null count."""
    gd = gpu_col.data()
    hd = host_col.data
    assert len(gd) == len(hd), f"{what}: length {len(gd)} != {len(hd)}"
    if gd.tobytes() != hd.tobytes():
        diff = np.nonzero(gd.view(f"u{gd.itemsize}") != hd.view(f"u{hd.itemsize}"))[0]
        i = diff[0]
        raise AssertionError(f"{what}: data differs at {len(diff)} rows, first row {i}: "
                             f"gpu={gd[i]!r} oracle={hd[i]!r}")
    if check_valid and len(hd) > 0:  # 0-row columns carry no bitmap (GDFColumn.cu:323-329)
        gv = gpu_col.valid()
        if host_col.valid is None:
            assert gv is None or gv.all(), f"{what}: unexpected nulls"
        else:
            assert gv is not None, f"{what}: missing validity"
            assert np.array_equal(gv, host_col.valid), f"{what}: validity differs"
            assert gpu_col.null_count == host_col.null_count, \
                f"{what}: null_count {gpu_col.null_count} != {host_col.null_count}"


def c1_frame(n, seed=1):
    """Filter+project workload: a int32 U[1,50], b f64 U[900,105000], c f64 in
    {0.00,...,0.10} (correctly rounded decimals), d int32 U[0,2555]; no nulls."""
    rng = np.random.default_rng(seed)
    a = rng.integers(1, 51, n).astype(np.int32)
    b = rng.uniform(900, 105000, n)
    k = rng.integers(0, 11, n)
    c = np.array([float(f"0.{i:02d}") for i in range(11)])[k]
    d = rng.integers(0, 2556, n).astype(np.int32)
    return [O.HostColumn(a, P.GDF_INT32), O.HostColumn(b, P.GDF_FLOAT64),
            O.HostColumn(c, P.GDF_FLOAT64), O.HostColumn(d, P.GDF_INT32)]


C1_FILTER = ("LogicalFilter(condition=[AND(>=($3, 730), <($3, 1095), >=($2, 0.03), "
             "<=($2, 0.05), <($0, 30))])")
C1_PROJECT = "LogicalProject(EXPR$0=[*($1, -(1, $2))], EXPR$1=[+($0, 1)])"


def ulp_diff(got, want):
    """Per-row distance in units in the last place between two float arrays of one
    dtype (float32 or float64): 0 where the bits are equal or both rows are NaN;
    inf where exactly one is NaN or two zeros differ in sign; otherwise the number
    of representable values between them (the bit patterns mapped onto a monotone
    integer line, negative values reflected)."""
    g = np.ascontiguousarray(got)
    w = np.ascontiguousarray(want)
    assert g.dtype == w.dtype and g.dtype in (np.float32, np.float64), (g.dtype, w.dtype)
    it = np.int64 if g.dtype == np.float64 else np.int32
    gi = g.view(it).astype(np.int64)
    wi = w.view(it).astype(np.int64)
    out = np.zeros(len(g), dtype=np.float64)
    diff = np.nonzero(gi != wi)[0]
    if len(diff) == 0:
        return out
    mag = (1 << (63 if it == np.int64 else 31)) - 1
    for i in diff.tolist():
        gn, wn = bool(np.isnan(g[i])), bool(np.isnan(w[i]))
        if gn and wn:
            continue
        if gn or wn:
            out[i] = np.inf
            continue
        a, b = int(gi[i]), int(wi[i])
        oa = -(a & mag) if a < 0 else a
        ob = -(b & mag) if b < 0 else b
        out[i] = np.inf if oa == ob else abs(oa - ob)   # equal ordinals, different bits: +0 / -0
    return out


C2_KINDS = {
    # dtype -> host generator of the wide-expression column mix (int32 uniform
    # [-1e6, 1e6], int64 uniform [0, 2^40], float32 normal(0, 1), float64 lognormal(0, 2))
    P.GDF_INT32: lambda rng, n: rng.integers(-1_000_000, 1_000_001, n).astype(np.int32),
    P.GDF_INT64: lambda rng, n: rng.integers(0, 2 ** 40 + 1, n).astype(np.int64),
    P.GDF_FLOAT32: lambda rng, n: rng.normal(0.0, 1.0, n).astype(np.float32),
    P.GDF_FLOAT64: lambda rng, n: rng.lognormal(0.0, 2.0, n),
}


def c2_mix_table(rng, dtypes, null_rates, n):
    """Host columns of the wide-expression mix with the given null rates."""
    cols = []
    for dt, nr in zip(dtypes, null_rates):
        data = C2_KINDS[dt](rng, n)
        valid = (rng.random(n) >= nr) if nr > 0 else None
        cols.append(O.HostColumn(data, dt, valid))
    return cols

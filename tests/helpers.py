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

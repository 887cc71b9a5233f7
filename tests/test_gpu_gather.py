"""GPU parity of blz_materialize_columns (materialize_column over several columns
sharing the indices, one call) against the oracle's gather
(ColumnManipulation.cu:83-150) and the single-column call: every width, nulls,
float NaN payloads, repeated, negative and past-the-end indices."""
import numpy as np
import pytest

from oracle import planner as P
from oracle import pyoracle as O
from tests.helpers import gen_column, to_device

pytestmark = pytest.mark.gpu

DTYPES = [P.GDF_INT8, P.GDF_BOOL8, P.GDF_INT16, P.GDF_INT32, P.GDF_INT64, P.GDF_FLOAT32, P.GDF_FLOAT64,
          P.GDF_DATE32, P.GDF_DATE64]


@pytest.fixture(scope="module")
def B():
    import blz_ral
    if blz_ral.device_count() < 1:
        pytest.fail("no HIP device visible")
    return blz_ral


@pytest.mark.parametrize("n,m", [(1, 1), (1000, 255), (100_000, 257), (70_001, 300_000)])
def test_materialize_columns_matches_per_column_gather(B, n, m):
    rng = np.random.default_rng(4400 + n + m)
    cols = []
    for i, dt in enumerate(DTYPES):
        h = gen_column(rng, P.GDF_INT8 if dt == P.GDF_BOOL8 else dt, n, 0.2 if i % 2 else 0.0)
        if dt == P.GDF_BOOL8:
            h = O.HostColumn((h.data != 0).astype(np.int8), dt, h.valid)
        cols.append(h)
    f32 = cols[DTYPES.index(P.GDF_FLOAT32)].data.view(np.uint32)
    f32[: min(n, 3)] = np.array([0x7F800001, 0xFFC00123, 0x80000000], np.uint32)[: min(n, 3)]
    idx = rng.integers(-3, n + 3, m).astype(np.int32)   # negatives and past-the-end: not written
    print("CASE test=materialize_columns", "n", n, "m", m, flush=True)
    dev = [to_device(B, h) for h in cols]
    didx = B.DeviceColumn.from_numpy(idx, B.GDF_INT32)
    outs = [B.DeviceColumn(h.dtype, m, with_valid=True) for h in cols]
    singles = [B.DeviceColumn(h.dtype, m, with_valid=True) for h in cols]
    B.materialize_columns(dev, outs, didx)
    for d, s in zip(dev, singles):
        B.materialize_column(d, s, didx)
    keep = (idx >= 0) & (idx < n)
    for h, o, s in zip(cols, outs, singles):
        ref = O.gather(h, np.where(idx < n, idx, -1).astype(np.int32), True)
        od = o.data()
        assert od[keep].tobytes() == ref.data[keep].tobytes(), f"dtype {h.dtype}: data"
        assert od[keep].tobytes() == s.data()[keep].tobytes(), f"dtype {h.dtype}: vs materialize_column"
        assert np.array_equal(o.valid(), ref.valid), f"dtype {h.dtype}: validity"
        assert o.null_count == ref.null_count == s.null_count, f"dtype {h.dtype}: null count"
    for c in dev + outs + singles + [didx]:
        c.free()


def test_materialize_columns_permutation_is_a_permutation(B):
    """At a larger size, where the oracle is not run: gathering a column holding its
    row number by a random permutation gives back the permutation, and gathering
    twice by a permutation and its inverse gives the column back."""
    n = 20_000_000
    rng = np.random.default_rng(4500)
    perm = rng.permutation(n).astype(np.int32)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(n, dtype=np.int32)
    rows = B.DeviceColumn.from_numpy(np.arange(n, dtype=np.int64), B.GDF_INT64)
    vals = B.DeviceColumn.from_numpy(rng.random(n), B.GDF_FLOAT64)
    p, q = B.DeviceColumn.from_numpy(perm, B.GDF_INT32), B.DeviceColumn.from_numpy(inv, B.GDF_INT32)
    o1 = [B.DeviceColumn(B.GDF_INT64, n), B.DeviceColumn(B.GDF_FLOAT64, n)]
    o2 = [B.DeviceColumn(B.GDF_INT64, n), B.DeviceColumn(B.GDF_FLOAT64, n)]
    print("CASE test=materialize_columns_permutation n", n, flush=True)
    B.materialize_columns([rows, vals], o1, p)
    assert np.array_equal(o1[0].data(), perm.astype(np.int64))
    B.materialize_columns(o1, o2, q)
    assert np.array_equal(o2[0].data(), np.arange(n, dtype=np.int64))
    assert o2[1].data().tobytes() == vals.data().tobytes()
    for c in [rows, vals, p, q] + o1 + o2:
        c.free()

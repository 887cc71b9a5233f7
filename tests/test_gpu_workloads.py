"""The batched and served workloads at their own query distribution, checked query
by query against the oracle and against the same query alone.

Batched: plans drawn by blz_ral.workloads.rand_query (4-40 ops over the
wide-expression operator set) over their own 4-8 column tables of the mix
(int32 / int64 / float32 / float64, 0-10% nulls), rows log-uniform, at the row
counts the oracle finishes in seconds; built with the autotuner and a replayed HIP
graph (the benchmark's configuration) and with the shape the benchmark's autotuner
picks (8 rows per lane, 2048-row tiles).

Served: a Poisson replay of filter + project queries over tables of the mix
(eight columns, 5% nulls); every query's results are read back and compared.
"""
import math

import numpy as np
import pytest

from oracle import planner as P
from oracle import pyoracle as O
from tests.helpers import assert_column_equal, c2_mix_table, to_device

pytestmark = pytest.mark.gpu

MIX = [P.GDF_INT32, P.GDF_INT32, P.GDF_INT64, P.GDF_INT64, P.GDF_FLOAT32, P.GDF_FLOAT32, P.GDF_FLOAT64,
       P.GDF_FLOAT64]


@pytest.fixture(scope="module")
def B():
    import blz_ral
    if blz_ral.device_count() < 1:
        pytest.fail("no HIP device visible")
    return blz_ral


def _oracle(host, filt, proj):
    mid = O.process_filter(host, filt) if filt else host
    return O.execute_project(mid, proj)


def _check(kept, cols, nulls, ref, what):
    assert kept == len(ref[0]), what
    assert len(cols) == len(ref), what
    for g, r, nc in zip(cols, ref, nulls):
        g.col.null_count = nc
        assert_column_equal(g, r, what=what)


def _batch_queries(B, rng, n_plans, rows_lo, rows_hi, filt_p):
    from blz_ral import workloads as W
    qs, refs, hosts = [], [], []
    for _ in range(n_plans):
        rows = W.log_uniform_rows(rng, rows_lo, rows_hi)
        spec = W.table_spec(rng, rows)
        host = c2_mix_table(rng, spec["dtypes"], spec["nulls"], rows)
        filt, proj = W.rand_query(rng, spec, bool(rng.random() < filt_p))
        refs.append(_oracle(host, filt, proj))
        hosts.append(host)
        qs.append(([to_device(B, h) for h in host], filt, proj))
    return qs, refs


@pytest.mark.parametrize("shape", ["autotuned_graph", "r8_t2048_graph", "r8_t2048_direct"])
def test_batch_of_workload_plans(B, shape):
    """48 plans of the batched workload (the benchmark draws projections; a third of
    these carry a filter too) in one batch: every query's outputs, validity and
    null counts equal the oracle's, over three launches."""
    rng = np.random.default_rng(44_000 + ["autotuned_graph", "r8_t2048_graph", "r8_t2048_direct"].index(shape))
    qs, refs = _batch_queries(B, rng, 48, 1e4, 3e5, 0.33)
    print(f"CASE batch_workload shape={shape} plans={len(qs)} rows={sum(q[0][0].col.size for q in qs)}",
          flush=True)
    if shape == "autotuned_graph":
        batch = B.Batch(qs, r=0, use_graph=True)
        assert batch.tuned >= 1 and batch.graph == 1
    else:
        batch = B.Batch(qs, r=8, t=2048, use_graph=shape.endswith("graph"))
        assert batch.r == 8 and batch.t == 2048
    for rep in range(3):
        batch.launch()
        for q, ((dev, filt, proj), ref) in enumerate(zip(qs, refs)):
            kept, cols, nulls = batch.result(q)
            _check(kept, cols, nulls, ref, f"batch {shape} q={q} rep={rep}: {filt} {proj}")
    # and a few queries alone, through a prepared program: the same bytes
    for q in (0, 17, 47):
        dev, filt, proj = qs[q]
        prog = B.Program(dev, filt, proj)
        prog.launch()
        pk, _ = prog.result()
        bk, bcols, _ = batch.result(q)
        assert pk == bk
        for a, b in zip(prog.outputs, bcols):
            assert a.data().tobytes() == b.data().tobytes()
        prog.close()
    batch.close()


@pytest.mark.parametrize("mode", ["compile_all", "timed"])
def test_batch_with_large_plans_on_the_compiled_tier(B, mode):
    """Plans of at least 2^20 rows are compiled (in one compilation) and run as
    launches of their own beside the shared tile-VM launch: every one when asked
    (t < 0), else where the autotuner times them faster. Every query's result
    equals the oracle's over three launches of the graph."""
    rng = np.random.default_rng(46_000 + (mode == "timed"))
    big, ref_big = _batch_queries(B, rng, 3, 1.1e6, 1.6e6, 0.33)
    small, ref_small = _batch_queries(B, rng, 7, 1e4, 1e5, 0.33)
    qs, refs = big + small, ref_big + ref_small
    print(f"CASE batch_compiled mode={mode} rows={[q[0][0].col.size for q in qs]}", flush=True)
    batch = B.Batch(qs, r=0, t=-1 if mode == "compile_all" else 0, use_graph=True)
    print(f"kernels {batch.kernels}", flush=True)
    if mode == "compile_all":
        assert batch.n_compiled == 3
        assert all(k == (5, False) for k in batch.kernels[:3])
    assert all(k[0] != 5 for k in batch.kernels[3:])
    for rep in range(3):
        batch.launch()
        for q, ((dev, filt, proj), ref) in enumerate(zip(qs, refs)):
            kept, cols, nulls = batch.result(q)
            _check(kept, cols, nulls, ref, f"batch {mode} q={q} rep={rep}: {filt} {proj}")
    batch.close()
    for dev, _, _ in qs:
        for c in dev:
            c.free()


def test_served_poisson_replay_every_result(B):
    """A Poisson replay (2000 queries/s) of filter + project queries of the served
    workload over four resident tables of the mix: every query's result, read
    after the replay, equals the oracle's; three of them also equal the query run
    alone."""
    from blz_ral import workloads as W
    rng = np.random.default_rng(45_000)
    tables, specs, hosts = [], [], []
    for _ in range(4):
        rows = W.log_uniform_rows(rng, 2e4, 2.5e5)
        spec = W.table_spec(rng, rows, dtypes=list(MIX), null_max=0.0)
        spec["nulls"] = [0.05] * len(MIX)
        host = c2_mix_table(rng, MIX, spec["nulls"], rows)
        specs.append(spec)
        hosts.append(host)
        tables.append([to_device(B, h) for h in host])
    qs, refs = [], []
    for _ in range(64):
        i = int(rng.integers(0, len(tables)))
        filt, proj = W.rand_query(rng, specs[i], True)
        qs.append((tables[i], filt, proj))
        refs.append(_oracle(hosts[i], filt, proj))
    arrivals = np.cumsum(rng.exponential(1 / 2000.0, len(qs)))
    print(f"CASE served_replay queries={len(qs)}", flush=True)
    srv = B.Server(8 << 30, max_batch=16, n_streams=4, r=4)
    lat, kept, wall, tickets = srv.replay(qs, arrivals, keep_results=True)
    assert (lat > 0).all() and wall > 0 and all(t > 0 for t in tickets)
    for q, (tk, (dev, filt, proj), ref) in enumerate(zip(tickets, qs, refs)):
        k, cols, nulls, _ = srv.wait(tk)
        assert k == kept[q]
        _check(k, cols, nulls, ref, f"served q={q}: {filt} {proj}")
        if q in (0, 21, 63):
            prog = B.Program(dev, filt, proj)
            prog.launch()
            pk, _ = prog.result()
            assert pk == k
            for a, b in zip(prog.outputs, cols):
                assert a.data().tobytes() == b.data().tobytes()
            prog.close()
        srv.release(tk)
    st = srv.stats()
    assert st["used_bytes"] == 0, st
    srv.close()

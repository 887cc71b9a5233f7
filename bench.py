"""Benchmark: fused filter+project throughput on one MI355X.

Workload (the metric's "filter+project at 100M rows"): 100,000,000 rows of
a int32 U[1,50], b float64 U[900,105000], c float64 in {0.00,...,0.10},
d int32 U[0,2555] (no nulls), filter
  AND(>=($3, 730), <($3, 1095), >=($2, 0.03), <=($2, 0.05), <($0, 30))
and project  b*(1-c), a+1  over the rows kept. One step = one execution of the
compiled filter+project program over the device-resident table: the predicate
kernel (keep words + per-slot counts, plus the kept rows' a and c staged per slot),
the slot scan, and the gather kernel that evaluates the projection on the kept
rows. By default the program is the compiled tier (the plan written out as kernels
and compiled for the device when the program is prepared); --kernel 4 measures
the tile VM instead.

value = algorithmic bytes per step / wall time per step (GB/s), where algorithmic
bytes = 24 B per input row (a, b, c, d read once) + 12 B per kept row (f64 + i32
written). roofline is for the dominant kernel (the predicate kernel): its
algorithmic bytes (predicate columns a, c, d read once: 16 B/row, keep words
written: 1/8 B/row, slot counts: 4 B per slot, and for the compiled tier the kept
rows' a and c staged: 12 B per kept row) over its own duration (HIP events on the
launch stream); "pipeline" gives the same figure for all three kernels together. traffic is the dominant kernel's HBM bytes per launch from the
rocprofv3 FETCH_SIZE / WRITE_SIZE summary in profiles/. cpu_baseline times the
scalar oracle on a 10M-row sample of the same workload on this host.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--rows N] [--calibrate]
python bench.py --workload wide-expression   (100M rows x 8 nullable columns, the 24-op
    four-output projection; its own JSON line)
python bench.py --workload csv-ingest   (a 20M-row lineitem-shaped text file in HBM parsed
    into seven device columns; its own JSON line)
python bench.py --workload filter-gather [--selectivity S]   (500M-row filter + gather:
    int32 key U[0, 2^31), 3 int64 sequential and 3 float64 U[0, 1) payload columns,
    key < S * 2^31, the payload compacted in row order; its own JSON line)
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "blazingdb-ral_amd"))

import numpy as np  # noqa: E402

METRIC = "filter+project GB/s (% HBM peak) at 100M rows; batched vs per-query p99 latency"
HBM_PEAK_GBS = 8000.0
FILTER = ("LogicalFilter(condition=[AND(>=($3, 730), <($3, 1095), >=($2, 0.03), "
          "<=($2, 0.05), <($0, 30))])")
PROJECT = "LogicalProject(EXPR$0=[*($1, -(1, $2))], EXPR$1=[+($0, 1)])"
# rocprofv3 summaries (kernel statistics and HBM counters) per workload and program kind
PROFILES = {
    ("c1", 5): "blz_cg_c1_filter_project_100M.txt", ("c1", 4): "vm_kernel_c1_filter_project_100M.txt",
    ("fg", 5): "blz_cg_filter_gather_500M_s0.9.txt", ("fg", 4): "vm_kernel_filter_gather_500M_s0.9.txt",
    ("wide", 5): "blz_cg_wide_expression_100M.txt", ("wide", 4): "vm_kernel_wide_expression_100M.txt",
}


def profile_path(workload, kernel):
    return os.path.join(ROOT, "profiles", PROFILES.get((workload, kernel), "none"))
STAGES = ["predicate", "slot_scan", "gather"]


def gen_c1(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(1, 51, n, dtype=np.int32)
    b = rng.uniform(900.0, 105000.0, n)
    c = np.array([float(f"0.{i:02d}") for i in range(11)])[rng.integers(0, 11, n, dtype=np.int8)]
    d = rng.integers(0, 2556, n, dtype=np.int32)
    return a, b, c, d


def algorithmic_bytes(n, kept):
    return 24 * n + 12 * kept


def stage_bytes(n, kept, tile_rows, compiled):
    """Algorithmic bytes of each kernel of the filter+project program (four slots
    per tile, one count each)."""
    tiles = (n + tile_rows - 1) // tile_rows
    if compiled:
        pred = 16 * n + n / 8 + 16 * tiles + 12 * kept   # a, c, d read; keep words, counts, staged a, c
        scan = 32 * tiles + n / 8 + 4 * kept             # counts read, offsets written, keep words; indices
        gather = 4 * kept + 20 * kept + 12 * kept + kept / 4  # indices; staged a, c and b; outputs + bitmaps
        return [pred, scan, gather]
    pred = 16 * n + n / 8 + 16 * tiles          # a, c, d read; keep words + slot counts written
    scan = 32 * tiles + n / 8 + 4 * kept        # counts read, offsets written, keep words; kept-row indices
    gather = 4 * kept + 20 * kept + 12 * kept   # indices; a, b, c of kept rows; outputs
    return [pred, scan, gather]


def traffic_from_profile(kernel, profile):
    """HBM bytes per launch of `kernel` measured with rocprofv3 counters (profile file)."""
    if not os.path.exists(profile):
        return None
    for line in open(profile):
        m = re.match(r"^traffic_bytes_per_launch\[(.+)\]:\s*([0-9.eE+]+)", line.strip())
        if m and m.group(1).startswith(kernel):
            return float(m.group(2))
    return None


def stage_kernel_names(prog):
    """Kernel name prefixes (as rocprofv3 prints them) of a filter program's three
    launches: predicate, slot scan, gather."""
    if prog.kernel == 5:
        return ["blz_cg_pred", "slot_scan_kernel", "blz_cg_gather"]
    pred = "pred_direct_kernel" if prog.direct_pred else f"vm_kernel<{prog.r}, 1,"
    return [pred, "slot_scan_kernel", f"vm_kernel<{prog.r}, 2,"]


def program_desc(prog):
    if prog.kernel == 5:
        return f"compiled tier: {prog.r} rows per lane" + (f", tile {prog.tile_rows} rows" if prog.tile_rows else "")
    if prog.kernel == 4:
        return f"tile VM R={prog.r} tile={prog.tile_rows} rows"
    return f"register-file kernel S={prog.s} R={prog.r}"


def _query_bytes(frame, proj, outs, kept=None):
    """Algorithmic bytes of one query: the input columns its expressions reference
    (data + bitmaps) and its outputs (data + bitmaps) for the rows kept."""
    used = set(int(m) for m in re.findall(r"\$(\d+)", proj))
    n = frame[0].col.size
    k = n if kept is None else kept
    b = 0
    for i in used:
        c = frame[i].col
        b += n * np.dtype(__import__("blz_ral").NP_OF[c.dtype]).itemsize + ((n + 7) // 8 if c.valid else 0)
    for o in outs:
        b += k * np.dtype(__import__("blz_ral").NP_OF[o.col.dtype]).itemsize + ((k + 7) // 8 if o.col.valid else 0)
    return b


def run_batch(B, n_plans, rows_lo, rows_hi, seed=404):
    """Batched execution of n_plans independent projections (4-40 ops each) over
    their own 4-8 column tables of log-uniform row counts, one persistent launch
    per kernel kind, shapes chosen by the autotuner (which also moves the largest
    plans to the compiled tier where they run faster), replayed as a HIP graph;
    against the same queries run one after another through the drop-in path
    (execute_project: plan, compile, allocate, launch, wait) and as prepared
    per-query programs on the batch's tier for each (launch only)."""
    from blz_ral import workloads as W
    rng = np.random.default_rng(seed)
    specs = [W.table_spec(rng, W.log_uniform_rows(rng, rows_lo, rows_hi)) for _ in range(n_plans)]
    tables = [W.make_table(sp) for sp in specs]
    queries = [(t, None, W.rand_query(rng, sp, False)[1]) for t, sp in zip(tables, specs)]
    print(f"[bench] batch: {n_plans} plans, {sum(sp['rows'] for sp in specs)} rows", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    batch = B.Batch(queries, r=0, use_graph=True)
    t_build = time.perf_counter() - t0
    batch.launch()
    batch_ms = batch.time(iters=5)
    total_bytes = 0
    for q, (frame, _, proj) in enumerate(queries):
        kept, outs, _ = batch.result(q)
        total_bytes += _query_bytes(frame, proj, outs, kept)
    batch_r, batch_t, tuned = batch.r, batch.t, batch.tuned
    kinds = [k for k, _ in batch.kernels]
    n_compiled = batch.n_compiled
    batch.close()
    # per-query: each query's prepared program on the tier the batch runs it on (the
    # compiled kernels are in the cache), launched one after another
    progs = [B.Program(f, fq, pq, kernel=5 if k == 5 else 4) for (f, fq, pq), k in zip(queries, kinds)]
    for p in progs:
        p.launch()
    progs[-1].result()
    each = [p.time(iters=3) for p in progs]
    t0 = time.perf_counter()
    for p in progs:
        p.launch()
    for p in progs:
        p.result()
    seq_wall_ms = (time.perf_counter() - t0) * 1e3
    for p in progs:
        p.close()
    done = np.cumsum(each)  # completion time of each query when all arrive at once
    # the drop-in path: each query through execute_project of the C ABI (plan,
    # compile, allocate the outputs, launch, wait), one after another, as the
    # reference runs a query; outputs are freed after the timed loop
    def drop_in():
        t0 = time.perf_counter()
        fin, keep = [], []
        for f, _, pq in queries:
            keep.append(B.execute_project(f, pq))
            fin.append(time.perf_counter() - t0)
        for outs in keep:
            for c in outs:
                c.free()
        return np.array(fin) * 1e3
    drop_in()  # first use of each plan shape
    drop = drop_in()
    for t in tables:
        for c in t:
            c.free()
    return {"workload": f"{n_plans} projection plans, 4-40 ops, 4-8 column tables (int32/int64/f32/f64, "
                        f"0-10% nulls), rows log-uniform [{rows_lo:.0e}, {rows_hi:.0e}]",
            "rows": int(sum(sp["rows"] for sp in specs)), "algorithmic_bytes": int(total_bytes),
            "batched": {"ms": batch_ms, "GBps": total_bytes / (batch_ms * 1e-3) / 1e9, "r": batch_r,
                        "tile": batch_t, "autotuner_candidates": tuned, "graph": True,
                        "p50_ms": batch_ms, "p99_ms": batch_ms, "build_s": t_build,
                        "compiled_plans": n_compiled},
            "per_query": {"ms": float(drop[-1]), "GBps": total_bytes / (drop[-1] * 1e-3) / 1e9,
                          "p50_ms": float(np.percentile(drop, 50)), "p99_ms": float(np.percentile(drop, 99)),
                          "path": "execute_project per query (plan, compile, allocate, launch, wait)"},
            "per_query_prepared": {"ms": float(done[-1]), "wall_ms": seq_wall_ms,
                                   "GBps": total_bytes / (done[-1] * 1e-3) / 1e9,
                                   "p50_ms": float(np.percentile(done, 50)),
                                   "p99_ms": float(np.percentile(done, 99)),
                                   "path": "each query's prepared program (the batch's tier for it), "
                                           "launched one after another"},
            "speedup": float(drop[-1] / batch_ms),
            "speedup_vs_prepared": float(done[-1] / batch_ms)}


def run_serve(B, total_gb, rates, seconds, seed=505, r=4, r_served=8):
    """Served query loop over device-resident tables (wide-expression column mix,
    5% nulls) totalling total_gb: Poisson arrivals at each rate, filter + project
    queries of 4-40 ops, admitted by a 200 GB HBM budget and continuously batched
    over 4 streams with 8 rows per lane; against per-query execution (one query
    per launch, one stream, 4 rows per lane: each mode at the shape it runs best
    at). p50 / p99 latency from submission to completion."""
    from blz_ral import workloads as W
    rng = np.random.default_rng(seed)
    mix = [B.GDF_INT32, B.GDF_INT32, B.GDF_INT64, B.GDF_INT64, B.GDF_FLOAT32, B.GDF_FLOAT32,
           B.GDF_FLOAT64, B.GDF_FLOAT64]
    specs, tables, have = [], [], 0
    while have < total_gb * 1e9:
        sp = W.table_spec(rng, W.log_uniform_rows(rng, 1e4, 5e7), dtypes=mix, null_max=0.0)
        sp["nulls"] = [0.05] * len(mix)
        specs.append(sp)
        tables.append(W.make_table(sp))
        have += W.table_bytes(sp)
    print(f"[bench] serve: {len(tables)} resident tables, {have / 1e9:.1f} GB", file=sys.stderr, flush=True)
    out = {"workload": f"{len(tables)} resident tables, {have / 1e9:.1f} GB (2 i32, 2 i64, 2 f32, 2 f64, "
                       f"5% nulls), filter + project queries of 4-40 ops, HBM budget 200 GB", "rates": {}}
    # untimed warm-up of both modes (kernel modules, and the stream-ordered memory
    # pool grown to the outputs and workspaces of a query on every table)
    warm = []
    for i in range(len(tables)):
        f, p = W.rand_query(rng, specs[i], True)
        warm.append((tables[i], f, p))
    for mb, ns, mr in ((64, 4, r_served), (1, 1, r)):
        srv = B.Server(200 << 30, max_batch=mb, n_streams=ns, r=mr)
        srv.replay(warm, np.zeros(len(warm)))
        srv.close()
    for qps in rates:
        n = max(20, int(qps * seconds))
        arrivals = np.cumsum(rng.exponential(1.0 / qps, n))
        qs = []
        for _ in range(n):
            i = int(rng.integers(0, len(tables)))
            f, p = W.rand_query(rng, specs[i], True)
            qs.append((tables[i], f, p))
        row = {}
        for mode, mb, ns, mr in (("served", 64, 4, r_served), ("per_query", 1, 1, r)):
            srv = B.Server(200 << 30, max_batch=mb, n_streams=ns, r=mr)
            # untimed: the first queries of this rate's schedule (the pool settles to
            # this mix's allocation sizes)
            k = min(n, 256)
            srv.replay(qs[:k], arrivals[:k] - arrivals[0])
            st0 = srv.stats()
            lat, kept, wall = srv.replay(qs, arrivals)
            st = srv.stats()
            st["batches"] -= st0["batches"]
            srv.close()
            rows = sum(specs_rows for specs_rows in (t[0].col.size for t, _, _ in qs))
            row[mode] = {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                         "rows_per_s": rows / wall, "queries": n, "wall_s": wall, "batches": st["batches"],
                         "peak_inflight": st["peak_inflight"]}
        out["rates"][str(qps)] = row
        print(f"[bench] serve {qps} QPS: {row}", file=sys.stderr, flush=True)
    for t in tables:
        for c in t:
            c.free()
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sample_rows=10_000_000, min_seconds=10.0, max_threads=16):
    """The scalar oracle on a bounded sample of the same workload: one thread, and
    all cores (row partitions on a thread pool; the oracle's C calls release the
    GIL), at most max_threads (this host's share of a shared box)."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import pyoracle as O
    from oracle import planner as P
    a, b, c, d = gen_c1(sample_rows, 99)
    host = [O.HostColumn(a, P.GDF_INT32), O.HostColumn(b, P.GDF_FLOAT64),
            O.HostColumn(c, P.GDF_FLOAT64), O.HostColumn(d, P.GDF_INT32)]

    def run(cols):
        return len(O.execute_project(O.process_filter(cols, FILTER), PROJECT)[0])

    def timed(fn):
        times, kept = [], 0
        t_end = time.perf_counter() + min_seconds / 2
        while len(times) < 3 or time.perf_counter() < t_end:
            t0 = time.perf_counter()
            kept = fn()
            times.append(time.perf_counter() - t0)
        return float(np.median(times)), kept, len(times)

    t1, kept, n1 = timed(lambda: run(host))
    threads = max(1, min(os.cpu_count() or 1, max_threads))
    bounds = np.linspace(0, sample_rows, threads + 1).astype(np.int64)
    parts = [[O.HostColumn(h.data[bounds[i]:bounds[i + 1]], h.dtype) for h in host] for i in range(threads)]
    with ThreadPoolExecutor(threads) as pool:
        tn, kept_n, nn = timed(lambda: sum(pool.map(run, parts)))
    if kept_n != kept:
        raise SystemExit("cpu baseline: partitioned run kept a different row count")
    gbs = lambda t: algorithmic_bytes(sample_rows, kept) / t / 1e9
    return {"value": gbs(tn), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"{sample_rows} rows of the same workload, scalar oracle (-O2 -ffp-contract=off) on "
                      f"{threads} threads over row partitions, median of {nn} passes ({tn:.2f} s each); "
                      f"one thread: {gbs(t1):.3f} GB/s, median of {n1} passes ({t1:.2f} s each)",
            "single_thread_value": gbs(t1), "cpu": cpu_model(), "nproc": os.cpu_count()}


def run_drop_in_call(B, frame, n, iters=20):
    """The same filter + project through the drop-in one-shot call (blz_filter_project:
    plan, compile, allocate the kept rows' outputs, run the tile VM, wait, as the
    reference's process_filter + execute_project_plan return): wall milliseconds per
    call, outputs freed after each."""
    for _ in range(3):
        outs, _ = B.filter_project(frame, FILTER, PROJECT)
        for c in outs:
            c.free()
    t0 = time.perf_counter()
    for _ in range(iters):
        outs, _ = B.filter_project(frame, FILTER, PROJECT)
        for c in outs:
            c.free()
    ms = (time.perf_counter() - t0) / iters * 1e3
    return {"ms": ms, "GBps": algorithmic_bytes(n, int(outs[0].col.size)) / (ms * 1e-3) / 1e9,
            "path": "blz_filter_project one-shot (plan, compile, allocate the kept rows, run the tile VM, wait)"}


def run_aggregate(B, frame, prog, kept, n, iters=20):
    """Aggregation without GROUP BY through the drop-in call (process_aggregate: parse,
    reduce, one-row results, blocking): SUM over the 100M-row float64 column b (one
    reduction, 8 B per row read), and TPC-H Q6's SUM(b*(1-c)) over the filter+project
    program's kept rows (with COUNT and MAX). Wall milliseconds per call."""
    def timed(cols, q):
        for _ in range(3):
            for c in B.process_aggregate(cols, q):
                c.free()
        t0 = time.perf_counter()
        for _ in range(iters):
            for c in B.process_aggregate(cols, q):
                c.free()
        return (time.perf_counter() - t0) / iters * 1e3
    col_ms = timed(frame, "LogicalAggregate(group=[{}], s=[SUM($1)])")
    prog.launch()
    prog.result()
    q6_ms = timed(prog.outputs, "LogicalAggregate(group=[{}], revenue=[SUM($0)], n=[COUNT()], m=[MAX($1)])")
    return {"sum_100M_f64": {"ms": col_ms, "GBps": 8 * n / (col_ms * 1e-3) / 1e9, "rows": n},
            "q6_sum_of_projection": {"ms": q6_ms, "rows": int(kept),
                                     "path": "process_aggregate over the filter+project outputs"}}


def run_gather(B, n=100_000_000, iters=5):
    """Join / order-by materialisation: 6 columns (3 int64, 3 float64) of n rows
    gathered by a random permutation, through one blz_materialize_columns call and
    through the reference's loop of materialize_column calls. Bytes per row: 4 B of
    index + 6 x (8 B read + 8 B written); wall milliseconds per gather of the table."""
    rng = np.random.default_rng(77)
    perm = rng.permutation(n).astype(np.int32)
    idx = B.DeviceColumn.from_numpy(perm, B.GDF_INT32)
    del perm
    cols = []
    for k in range(6):
        c = B.DeviceColumn(B.GDF_INT64 if k < 3 else B.GDF_FLOAT64, n)
        B.synth_column(c, 900 + k, 5 if k < 3 else 1, 0.0, 1.0, 0.0)
        cols.append(c)
    outs = [B.DeviceColumn(c.col.dtype, n) for c in cols]

    def timed(f):
        f()
        t0 = time.perf_counter()
        for _ in range(iters):
            f()
        return (time.perf_counter() - t0) / iters * 1e3

    multi = timed(lambda: B.materialize_columns(cols, outs, idx))
    loop = timed(lambda: [B.materialize_column(c, o, idx) for c, o in zip(cols, outs)])
    for c in cols + outs + [idx]:
        c.free()
    by = n * (4 + 6 * 16)
    return {"workload": f"{n} rows x 6 columns (3 int64, 3 f64), random permutation", "bytes": by,
            "materialize_columns": {"ms": multi, "GBps": by / (multi * 1e-3) / 1e9},
            "materialize_column_loop": {"ms": loop, "GBps": by / (loop * 1e-3) / 1e9}}


def run_wide_expression(B, args):
    """Wide-expression evaluation: 100M rows x 8 nullable columns (generated on the
    device), the 24-op four-output projection, one launch of the tile VM per step;
    one JSON line. value = algorithmic bytes (the referenced input columns and their
    bitmaps read once, the outputs and their bitmaps written) per step over the
    step's wall time."""
    from blz_ral import workloads as W
    n = args.rows
    print(f"[bench] generating {n} rows x 8 columns on the device", file=sys.stderr, flush=True)
    frame = []
    for i, (dt, dist, a, b) in enumerate(W.WIDE_COLUMNS):
        c = B.DeviceColumn(dt, n, with_valid=True)
        B.synth_column(c, 7001 + i, dist, a, b, W.WIDE_NULL_RATE)
        B.update_null_count(c)
        frame.append(c)
    prog = B.Program(frame, None, W.WIDE_PROJECT, kernel=args.kernel)
    for _ in range(args.warmup):
        prog.launch()
    prog.result()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        prog.launch()
    prog.result()
    dt = (time.perf_counter() - t0) / args.steps
    kernel_ms = prog.time(iters=max(10, args.steps // 2))
    width = lambda d: np.dtype(B.NP_OF[d]).itemsize
    used = sorted({int(m) for m in re.findall(r"\$(\d+)", W.WIDE_PROJECT)})
    bytes_in = sum(n * width(frame[i].col.dtype) + (n + 7) // 8 for i in used)
    bytes_out = sum(n * width(o.col.dtype) + (n + 7) // 8 for o in prog.outputs)
    bytes_step = bytes_in + bytes_out
    kernel_name = "blz_cg_project" if prog.kernel == 5 else f"vm_kernel<{prog.r}, 0,"
    out = {
        "metric": "wide-expression GB/s at 100M rows x 8 columns", "value": bytes_step / dt / 1e9, "unit": "GB/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64/f64",
        "data": "synthetic",
        "config": {"workload": "wide expression: 2 i32, 2 i64, 2 f32, 2 f64 columns, 5% nulls, 24-op plan, "
                               "4 outputs", "rows": n, "ops": prog.n_ops, "algorithmic_bytes_per_step": bytes_step,
                   "program": program_desc(prog)},
        "roofline": {"bound": "hbm", "achieved": bytes_step / (kernel_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": bytes_step / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": traffic_from_profile(kernel_name, profile_path("wide", prog.kernel)),
                     "kernel": kernel_name,
                     "kernel_ms": kernel_ms, "algorithmic_bytes": bytes_step},
    }
    prog.close()
    for c in frame:
        c.free()
    if args.calibrate:
        buf_a = B.DeviceBuffer(1 << 30)
        buf_b = B.DeviceBuffer(1 << 30)
        for _ in range(3):
            B._check(B.lib().blz_device_copy(buf_b.ptr, buf_a.ptr, 1 << 30, None), "copy")
        print("[bench] calibration copy: 3 x 1 GiB", file=sys.stderr, flush=True)
    print(json.dumps(out))


PROFILE_CSV = os.path.join(ROOT, "profiles", "csv_fields_kernel_lineitem_20M.txt")


def gen_lineitem_text(n, seed=77):
    """A lineitem-shaped pipe-delimited file of n fixed-width rows (59 bytes each):
    orderkey (10 digits), partkey (8), quantity NN.NN, extendedprice NNNNNN.NN,
    discount 0.NN, tax 0.0N, returnflag (one letter), shipdate YYYY-MM-DD; and the
    values it was written from (reals in hundredths)."""
    rng = np.random.default_rng(seed)
    v = {"okey": rng.integers(0, 10 ** 10, n), "pkey": rng.integers(0, 10 ** 8, n),
         "qty": rng.integers(100, 5001, n), "price": rng.integers(90_000, 99_999_999, n),
         "disc": rng.integers(0, 11, n), "tax": rng.integers(0, 9, n), "days": rng.integers(8035, 10591, n)}
    fields = [("okey", 10, None), ("pkey", 8, None), ("qty", 5, 2), ("price", 9, 2), ("disc", 4, 2),
              ("tax", 4, 2), ("flag", 1, None), ("date", 10, None)]
    width = sum(w for _, w, _ in fields) + len(fields)
    buf = np.empty((n, width), np.uint8)
    col = 0

    def digits(vals, c0, w):
        x = vals.copy()
        for k in range(w - 1, -1, -1):
            buf[:, c0 + k] = 48 + (x % 10)
            x //= 10

    for name, w, dec in fields:
        if name == "flag":
            buf[:, col] = np.array([ord("A"), ord("N"), ord("R")], np.uint8)[v["okey"] % 3]
        elif name == "date":
            d = v["days"].astype("datetime64[D]")
            y = d.astype("datetime64[Y]").astype(np.int64) + 1970
            m = d.astype("datetime64[M]").astype(np.int64) % 12 + 1
            dd = (d - d.astype("datetime64[M]")).astype(np.int64) + 1
            digits(y, col, 4)
            buf[:, col + 4] = ord("-")
            digits(m, col + 5, 2)
            buf[:, col + 7] = ord("-")
            digits(dd, col + 8, 2)
        elif dec:
            digits(v[name] // 100, col, w - 3)
            buf[:, col + w - 3] = ord(".")
            digits(v[name] % 100, col + w - 2, 2)
        else:
            digits(v[name], col, w)
        col += w
        buf[:, col] = ord("|") if name != "date" else ord("\n")
        col += 1
    return buf.reshape(-1), v


def run_csv_ingest(B, args):
    """Ingest: a 20M-row lineitem-shaped text file resident in HBM parsed into seven
    device columns (int64, int32, four binary64, date32; the one-letter flag is
    skipped) per step: the four kernels and the output allocations. value = text
    bytes per step over the step's wall time. roofline: the field kernel (one
    thread per row), text bytes read + column bytes written over its own duration.
    cpu_baseline: the oracle's reader (pure Python) on the first 20,000 rows."""
    from oracle import planner as P
    n = args.csv_rows
    print(f"[bench] generating a {n}-row text file", file=sys.stderr, flush=True)
    text, vals = gen_lineitem_text(n)
    dtypes = [P.GDF_INT64, P.GDF_INT32, P.GDF_FLOAT64, P.GDF_FLOAT64, P.GDF_FLOAT64, P.GDF_FLOAT64,
              P.GDF_STRING, P.GDF_DATE32]
    want = [0, 1, 2, 3, 4, 5, 7]
    buf = B.DeviceBuffer(len(text))
    buf.upload(text)
    nbytes = len(text)

    def step():
        cols, rows = B.read_csv((buf.value, nbytes), dtypes, want, on_device=True)
        return cols, rows

    cols, rows = step()
    assert rows == n
    checks = [cols[0].data()[:1000] == vals["okey"][:1000], cols[3].data()[-1000:] == vals["price"][-1000:] / 100.0,
              cols[6].data()[:1000] == vals["days"][:1000]]
    if not all(c.all() for c in checks):
        raise SystemExit("csv ingest: parsed values differ from the values written")
    out_bytes = sum(c.col.size * np.dtype(B.NP_OF[c.col.dtype]).itemsize for c in cols)
    del cols
    for _ in range(args.warmup):
        step()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cols, _ = step()
        del cols
    dt = (time.perf_counter() - t0) / args.steps
    stage = np.zeros(4)
    reps = max(3, args.steps // 4)
    for _ in range(reps):
        cols, _, ms = B.parse_csv_device_timed(buf.value, nbytes, dtypes, want)
        stage += np.array(ms)
        del cols
    stage /= reps
    names = ["csv_count_kernel", "csv_scan_kernel", "csv_rows_kernel", "csv_fields_kernel"]
    k = int(np.argmax(stage))
    alg = {0: nbytes, 1: 0, 2: nbytes + rows * 8, 3: nbytes + out_bytes}[k]
    cpu = None
    if not args.no_cpu:
        from oracle import csv_oracle as CO
        sample = bytes(text[: 20_000 * (nbytes // n)])
        t1 = time.perf_counter()
        CO.read_csv(sample, dtypes, want)
        ct = time.perf_counter() - t1
        cpu = {"value": len(sample) / ct / 1e9, "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": f"oracle/csv_oracle.read_csv (pure Python) on the first 20000 rows ({len(sample)} bytes) "
                         f"of the same file, {ct:.2f} s", "cpu": cpu_model(), "nproc": os.cpu_count()}
    out = {
        "metric": "csv ingest GB/s of text at 20M lineitem rows", "value": nbytes / dt / 1e9, "unit": "GB/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8->int64/f64",
        "data": "synthetic",
        "config": {"workload": "lineitem-shaped text, 8 fields, 7 parsed (int64, int32, 4 binary64, date32)",
                   "rows": n, "text_bytes": nbytes, "output_bytes": out_bytes,
                   "stage_ms": dict(zip(names, [round(x, 4) for x in stage]))},
        "roofline": {"bound": "hbm", "achieved": alg / (stage[k] * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg / (stage[k] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": traffic_from_profile(names[k], PROFILE_CSV), "kernel": names[k],
                     "kernel_ms": float(stage[k]), "algorithmic_bytes": alg},
        "cpu_baseline": cpu,
    }
    buf.free()
    if args.calibrate:
        buf_a = B.DeviceBuffer(1 << 30)
        buf_b = B.DeviceBuffer(1 << 30)
        for _ in range(3):
            B._check(B.lib().blz_device_copy(buf_b.ptr, buf_a.ptr, 1 << 30, None), "copy")
        print("[bench] calibration copy: 3 x 1 GiB", file=sys.stderr, flush=True)
    print(json.dumps(out))


def run_filter_gather(B, args):
    """500M-row filter + gather: one JSON line, value = (52 N + 48 kept) bytes per step
    over the step's wall time (inputs generated on the device, resident); roofline
    for the dominant kernel (the gather of the kept rows at high selectivity)."""
    n = args.rows if args.rows != 100_000_000 else 500_000_000
    s = args.selectivity

    def synth(dtype, seed, dist, a, b):
        c = B.DeviceColumn(dtype, n, with_valid=False)
        B.synth_column(c, seed, dist, a, b, 0.0)
        return c

    print(f"[bench] generating {n} rows on the device", file=sys.stderr, flush=True)
    frame = [synth(B.GDF_INT32, 7201, 0, 0, 2.0**31 - 1), synth(B.GDF_INT64, 0, 5, 0, 0),
             synth(B.GDF_INT64, 0, 5, 1e12, 0), synth(B.GDF_INT64, 0, 5, -5, 0),
             synth(B.GDF_FLOAT64, 7202, 1, 0.0, 1.0), synth(B.GDF_FLOAT64, 7203, 1, 0.0, 1.0),
             synth(B.GDF_FLOAT64, 7204, 1, 0.0, 1.0)]
    filt = f"LogicalFilter(condition=[<($0, {int(s * 2**31)})])"
    proj = "LogicalProject(p1=[$1], p2=[$2], p3=[$3], f1=[$4], f2=[$5], f3=[$6])"
    prog = B.Program(frame, filt, proj, kernel=args.kernel)
    for _ in range(args.warmup):
        prog.launch()
    kept, _ = prog.result()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        prog.launch()
    kept, _ = prog.result()
    dt = (time.perf_counter() - t0) / args.steps
    stage_ms = prog.time_stages(iters=max(10, args.steps // 2))
    tiles = (n + prog.tile_rows - 1) // prog.tile_rows
    sbytes = [4 * n + n / 8 + 16 * tiles, 16 * tiles + n / 8 + 4 * kept, 4 * kept + 48 * kept + 48 * kept]
    dom = int(np.argmax(stage_ms))
    kernel_name = stage_kernel_names(prog)[dom]
    achieved = sbytes[dom] / (stage_ms[dom] * 1e-3) / 1e9
    bytes_step = 52 * n + 48 * kept
    desc = program_desc(prog)
    kind = prog.kernel
    prog.close()
    for c in frame:
        c.free()
    if args.calibrate:
        buf_a = B.DeviceBuffer(1 << 30)
        buf_b = B.DeviceBuffer(1 << 30)
        for _ in range(3):
            B._check(B.lib().blz_device_copy(buf_b.ptr, buf_a.ptr, 1 << 30, None), "copy")
        print("[bench] calibration copy: 3 x 1 GiB", file=sys.stderr, flush=True)
    print(json.dumps({
        "metric": "filter + gather GB/s at 500M rows", "value": bytes_step / dt / 1e9, "unit": "GB/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64/f64",
        "data": "synthetic",
        "config": {"workload": "filter + gather: int32 key, 3 int64 + 3 f64 payload, key < s*2^31",
                   "rows": n, "selectivity": s, "rows_kept": int(kept),
                   "min_traffic_GBps": (4 * n + 96 * kept) / dt / 1e9, "program": desc},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_from_profile(kernel_name, profile_path("fg", kind)),
                     "kernel": kernel_name, "stage": STAGES[dom], "kernel_ms": stage_ms[dom],
                     "algorithmic_bytes": sbytes[dom]},
        "pipeline": {"stages": dict(zip(STAGES, stage_ms)), "kernel_ms": float(sum(stage_ms))},
    }))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--rows", type=int, default=100_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a 1 GiB plain copy kernel (counter calibration)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-plans measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-queries measurement")
    ap.add_argument("--batch-plans", type=int, default=256)
    ap.add_argument("--serve-gb", type=float, default=100.0)
    ap.add_argument("--serve-seconds", type=float, default=1.0)
    ap.add_argument("--workload", choices=["filter-project", "filter-gather", "wide-expression", "csv-ingest"],
                    default="filter-project")
    ap.add_argument("--csv-rows", type=int, default=20_000_000, help="csv-ingest: rows of the text file")
    ap.add_argument("--selectivity", type=float, default=0.9, help="filter-gather: fraction of rows kept")
    ap.add_argument("--kernel", type=int, default=0, choices=[0, 3, 4, 5],
                    help="program kind: 0 automatic (the compiled tier), 4 tile VM, 3 register-file kernel")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import blz_ral as B
    if B.device_count() < 1:
        raise SystemExit("no HIP device visible")
    if args.workload == "filter-gather":
        run_filter_gather(B, args)
        return
    if args.workload == "wide-expression":
        run_wide_expression(B, args)
        return
    if args.workload == "csv-ingest":
        run_csv_ingest(B, args)
        return

    n = args.rows
    print(f"[bench] generating {n} rows", file=sys.stderr, flush=True)
    a, b, c, d = gen_c1(n, 2024)
    frame = [B.DeviceColumn.from_numpy(x) for x in (a, b, c, d)]
    del a, b, c, d
    t_prep = time.perf_counter()
    prog = B.Program(frame, FILTER, PROJECT, kernel=args.kernel)
    t_prep = time.perf_counter() - t_prep
    print(f"[bench] program kernel={prog.kernel} R={prog.r} tile={prog.tile_rows} ops={prog.n_ops} "
          f"grid={prog.grid}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        prog.launch()
    kept, _ = prog.result()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        prog.launch()
    kept, _ = prog.result()  # synchronises the stream
    dt = (time.perf_counter() - t0) / args.steps
    print(f"[bench] {args.steps} steps, {dt * 1e3:.3f} ms/step, kept {kept}", file=sys.stderr, flush=True)
    stage_ms = prog.time_stages(iters=max(10, args.steps // 2))
    names = STAGES if len(stage_ms) == 3 else ["program"]
    sbytes = (stage_bytes(n, kept, prog.tile_rows, prog.kernel == 5) if len(stage_ms) == 3
              else [algorithmic_bytes(n, kept)])
    dom = int(np.argmax(stage_ms))
    # kernel name prefixes as rocprofv3 prints them: blz_cg_* (compiled tier), or
    # pred_direct_kernel / vm_kernel<rows per lane, mode (1 predicate, 2 gather), op set>
    kernel_name = stage_kernel_names(prog)[dom] if len(stage_ms) == 3 else f"kernel {prog.kernel}"
    desc = program_desc(prog)
    kind = prog.kernel
    achieved = sbytes[dom] / (stage_ms[dom] * 1e-3) / 1e9
    bytes_step = algorithmic_bytes(n, kept)
    total_ms = float(sum(stage_ms))
    if args.calibrate:
        buf_a = B.DeviceBuffer(1 << 30)
        buf_b = B.DeviceBuffer(1 << 30)
        for _ in range(3):
            B._check(B.lib().blz_device_copy(buf_b.ptr, buf_a.ptr, 1 << 30, None), "copy")
        print("[bench] calibration copy: 3 x 1 GiB", file=sys.stderr, flush=True)
    aggregate = run_aggregate(B, frame, prog, kept, n)
    prog.close()
    drop_in = run_drop_in_call(B, frame, n)
    for c in frame:
        c.free()
    gather = run_gather(B)
    batch = None if args.no_batch else run_batch(B, args.batch_plans, 1e4, 5e7)
    serve = None if args.no_serve else run_serve(B, args.serve_gb, [200, 1000, 5000], args.serve_seconds)
    cpu = None if args.no_cpu else cpu_baseline()
    line = {
        "metric": METRIC,
        "value": bytes_step / dt / 1e9,
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "C1 filter+project: a i32, b f64, c f64, d i32, 5-term AND "
                               "predicate, project b*(1-c), a+1",
                   "rows": n, "rows_kept": int(kept), "selectivity": kept / n,
                   "algorithmic_bytes_per_step": bytes_step,
                   "program": desc, "prepare_s": t_prep},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_from_profile(kernel_name, profile_path("c1", kind)),
                     "kernel": kernel_name, "stage": names[dom], "kernel_ms": stage_ms[dom],
                     "algorithmic_bytes": sbytes[dom]},
        "pipeline": {"stages": dict(zip(names, stage_ms)), "kernel_ms": total_ms,
                     "achieved": bytes_step / (total_ms * 1e-3) / 1e9,
                     "frac": bytes_step / (total_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
        "drop_in_call": drop_in,
        "aggregate": aggregate,
        "gather": gather,
        "batch": batch,
        "serve": serve,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()

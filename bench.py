"""Benchmark: fused filter+project throughput on one MI355X.

Workload (the metric's "filter+project at 100M rows"): 100,000,000 rows of
a int32 U[1,50], b float64 U[900,105000], c float64 in {0.00,...,0.10},
d int32 U[0,2555] (no nulls), filter
  AND(>=($3, 730), <($3, 1095), >=($2, 0.03), <=($2, 0.05), <($0, 30))
and project  b*(1-c), a+1  over the rows kept. One step = one execution of the
compiled filter+project program over the device-resident table: the predicate
kernel (keep words + per-slot counts), the slot scan, and the gather kernel that
evaluates the projection on the kept rows.

value = algorithmic bytes per step / wall time per step (GB/s), where algorithmic
bytes = 24 B per input row (a, b, c, d read once) + 12 B per kept row (f64 + i32
written). roofline is for the dominant kernel (the predicate kernel): its
algorithmic bytes (predicate columns a, c, d read once: 16 B/row, keep words
written: 1/8 B/row, slot counts: 16 B per 1024-row tile) over its own duration
(HIP events on the launch stream); "pipeline" gives the same figure for all three
kernels together. traffic is the dominant kernel's HBM bytes per launch from the
rocprofv3 FETCH_SIZE / WRITE_SIZE summary in profiles/. cpu_baseline times the
scalar oracle on a 10M-row sample of the same workload on this host.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--rows N] [--calibrate]
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
PROFILE = os.path.join(ROOT, "profiles", "vm_kernel_c1_filter_project_100M.txt")
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


def stage_bytes(n, kept, tile_rows):
    """Algorithmic bytes of each kernel of the filter+project program."""
    tiles = (n + tile_rows - 1) // tile_rows
    pred = 16 * n + n / 8 + 16 * tiles          # a, c, d read; keep words + slot counts written
    scan = 8 * 4 * tiles                         # slot counts read, prefixes written
    gather = n / 8 + 20 * kept + 12 * kept      # keep words; a, b, c of kept rows; outputs
    return [pred, scan, gather]


def traffic_from_profile(kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 counters (profile file)."""
    if not os.path.exists(PROFILE):
        return None
    for line in open(PROFILE):
        m = re.match(r"^traffic_bytes_per_launch\[(.+)\]:\s*([0-9.eE+]+)", line.strip())
        if m and m.group(1) in kernel:
            return float(m.group(2))
    return None


def cpu_baseline(sample_rows=10_000_000, min_seconds=10.0):
    from oracle import pyoracle as O
    from oracle import planner as P
    a, b, c, d = gen_c1(sample_rows, 99)
    host = [O.HostColumn(a, P.GDF_INT32), O.HostColumn(b, P.GDF_FLOAT64),
            O.HostColumn(c, P.GDF_FLOAT64), O.HostColumn(d, P.GDF_INT32)]
    times, kept = [], 0
    t_end = time.perf_counter() + min_seconds
    while len(times) < 3 or time.perf_counter() < t_end:
        t0 = time.perf_counter()
        out = O.execute_project(O.process_filter(host, FILTER), PROJECT)
        times.append(time.perf_counter() - t0)
        kept = len(out[0])
    t = float(np.median(times))
    return {"value": algorithmic_bytes(sample_rows, kept) / t / 1e9, "unit": "GB/s", "cores": 1,
            "kind": "port",
            "sample": f"{sample_rows} rows of the same workload, scalar oracle (1 thread, -O2 "
                      f"-ffp-contract=off), median of {len(times)} passes ({t:.2f} s each)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--rows", type=int, default=100_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a 1 GiB plain copy kernel (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import blz_ral as B
    if B.device_count() < 1:
        raise SystemExit("no HIP device visible")

    n = args.rows
    print(f"[bench] generating {n} rows", file=sys.stderr, flush=True)
    a, b, c, d = gen_c1(n, 2024)
    frame = [B.DeviceColumn.from_numpy(x) for x in (a, b, c, d)]
    del a, b, c, d
    prog = B.Program(frame, FILTER, PROJECT)
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
    sbytes = stage_bytes(n, kept, prog.tile_rows) if len(stage_ms) == 3 else [algorithmic_bytes(n, kept)]
    dom = int(np.argmax(stage_ms))
    kernel_name = (f"vm_kernel<{prog.r}, {1 if dom == 0 else 2}, false>" if names[dom] != "slot_scan"
                   else "slot_scan_kernel") if len(stage_ms) == 3 else f"kernel {prog.kernel}"
    achieved = sbytes[dom] / (stage_ms[dom] * 1e-3) / 1e9
    bytes_step = algorithmic_bytes(n, kept)
    total_ms = float(sum(stage_ms))
    if args.calibrate:
        buf_a = B.DeviceBuffer(1 << 30)
        buf_b = B.DeviceBuffer(1 << 30)
        for _ in range(3):
            B._check(B.lib().blz_device_copy(buf_b.ptr, buf_a.ptr, 1 << 30, None), "copy")
        print("[bench] calibration copy: 3 x 1 GiB", file=sys.stderr, flush=True)
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
                   "program": f"tile VM R={prog.r} tile={prog.tile_rows} rows"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profile(kernel_name),
                     "kernel": kernel_name, "stage": names[dom], "kernel_ms": stage_ms[dom],
                     "algorithmic_bytes": sbytes[dom]},
        "pipeline": {"stages": dict(zip(names, stage_ms)), "kernel_ms": total_ms,
                     "achieved": bytes_step / (total_ms * 1e-3) / 1e9,
                     "frac": bytes_step / (total_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
    }
    prog.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()

"""Benchmark of the fused scan-aggregate hot path on one MI355X.

Workload (one step = one pass of the hot path over one batch): two 2^28-row
f64 columns a, b resident in HBM, uniform in [0, 1000) (SplitMix64 rows,
seeds 1 and 2), compound predicate a > t1 AND b < t2 at 50 % selectivity
(t1 = 292.8932188, t2 = 707.1067812), returning COUNT, SUM(a), MIN(a), MAX(a)
in one fused launch (voxel_gpu_scan_launch).  A breakdown over f64/f32/i64/i32
and 1/50/99 % selectivity is printed in the same JSON line.

    python bench.py [--steps K] [--warmup W] [--no-cpu] [--no-breakdown]

value = % of the 8 TB/s HBM peak achieved by the f64 compound scan, from the
whole timed region (K launches bracketed by device synchronisation);
roofline.achieved = algorithmic bytes / average kernel duration from HIP
events recorded on the launch stream; cpu_baseline = the reference's own JIT
fused filter+sum (oracle/_ref, single core) on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "filter+sum %HBM peak; sort Gkeys/s learned vs LSD radix; join Gprobes/s"

SEL = {  # floats: t1, t2 with independent factors sqrt(s)
    0.01: (900.0, 100.0),
    0.50: (292.8932188, 707.1067812),
    0.99: (5.0125628, 994.9874371),
}


def int_thresholds(s: float):
    r = math.sqrt(s)
    return int(math.floor((1 - r) * 2**31)), int(math.ceil(r * 2**31))


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--rows-log2", type=int, default=28)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-breakdown", action="store_true")
    return p.parse_args()


def time_scan(desc, steps, warmup, stream_handle):
    import torch
    from voxel_amd import scan
    for _ in range(warmup):
        scan.launch(desc, stream_handle)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()                       # torch's current stream == the launch stream
    for _ in range(steps):
        scan.launch(desc, stream_handle)
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ev0.elapsed_time(ev1) / steps
    return wall, kernel_ms


def cpu_baseline(n_rows_log2: int = 27):
    """The reference's fastest CPU path for filter+sum: its x86-64 JIT fused
    loop (codegen/x86_64.hpp:1632-1716), built from the reference sources into
    oracle/_ref, on one core; 2 warmups + median of 3 (validate.cpp:20)."""
    import numpy as np
    from oracle import oracle as o
    if o.R is None:
        return None
    n = 1 << n_rows_log2
    x = (np.random.default_rng(1).random(n) * 1000.0)
    code = np.array([0x00001070, 0x000301C5, 0x000015D0, 0x0005002A, 0x00041120, 0x00021040, 0xFFA00052,
                     0x00000001], np.uint32)
    fn = o.R.vr_jit_compile(o.ptr(code), code.size, None)
    if not fn:
        return None
    ts = []
    for i in range(5):
        t0 = time.perf_counter()
        o.R.vr_jit_entry(fn, o.ptr(x), n, n, 500.0)
        ts.append(time.perf_counter() - t0)
    o.R.vr_jit_release(fn)
    med = statistics.median(ts[2:])
    gbs = n * 8 / med / 1e9
    return {"value": round(100.0 * gbs / HBM_PEAK_GBS, 4), "unit": "% of 8 TB/s HBM peak",
            "gb_per_s": round(gbs, 3), "rows_per_s": round(n / med, 1), "cores": 1, "kind": "reference",
            "sample": "reference JIT fused filter+sum (oracle/_ref vr_jit_entry = voxel_jit_run loop), "
                      "2^%d f64 rows uniform [0,1000), x > 500, single column (8 B/row), "
                      "median of 3 after 2 warmups, one core" % n_rows_log2,
            "host_cpu": _cpu_model(), "nproc": os.cpu_count()}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1).", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("no GPU available", file=sys.stderr)
        sys.exit(1)
    torch.cuda.init()
    torch.empty(1, device="cuda")
    from voxel_amd import scan
    from voxel_amd._lib import lib, last_error
    if lib.voxel_gpu_available() != 1:
        print("libvoxel_c: " + last_error(), file=sys.stderr)
        sys.exit(1)
    stream = torch.cuda.current_stream().cuda_stream
    n = 1 << args.rows_log2
    cols = {}

    def column(dt_name, seed):
        key = (dt_name, seed)
        if key not in cols:
            tdt = {"f64": torch.float64, "f32": torch.float32, "i64": torch.int64, "i32": torch.int32}[dt_name]
            code = {"f64": 0, "f32": 1, "i64": 2, "i32": 3}[dt_name]
            t = torch.empty(n, dtype=tdt, device="cuda")
            hi = 1000.0 if dt_name in ("f64", "f32") else float(2**31)
            if lib.voxel_gpu_fill_uniform(code, t.data_ptr(), n, seed, 0.0, hi, stream) != 0:
                raise RuntimeError("fill failed")
            cols[key] = t
        return cols[key]

    def desc_for(dt_name, sel):
        a, b = column(dt_name, 1), column(dt_name, 2)
        if dt_name in ("f64", "f32"):
            t1, t2 = SEL[sel]
        else:
            t1, t2 = int_thresholds(sel)
        return scan.make_desc(a, "gt", t1, b, "lt", t2), a.element_size()

    # headline: f64 compound at 50 %
    desc, es = desc_for("f64", 0.50)
    wall, kernel_ms = time_scan(desc, args.steps, args.warmup, stream)
    res = scan.result(0, stream)
    algo_bytes = 2 * n * es
    ms_per_step = wall * 1e3 / args.steps
    gbs_wall = algo_bytes / (wall / args.steps) / 1e9
    gbs_kernel = algo_bytes / (kernel_ms * 1e-3) / 1e9

    breakdown = {}
    if not args.no_breakdown:
        for dt_name in ("f64", "f32", "i64", "i32"):
            for sel in (0.01, 0.50, 0.99):
                d, e = desc_for(dt_name, sel)
                _, kms = time_scan(d, max(5, args.steps // 2), 2, stream)
                r = scan.result(d.dtype, stream)
                gb = 2 * n * e / (kms * 1e-3) / 1e9
                breakdown["%s_sel%02d" % (dt_name, int(sel * 100))] = {
                    "kernel_ms": round(kms, 4), "gb_per_s": round(gb, 1),
                    "pct_hbm_peak": round(100 * gb / HBM_PEAK_GBS, 2),
                    "selectivity_measured": round(r["count"] / n, 5)}
            del cols[(dt_name, 1)], cols[(dt_name, 2)]
            torch.cuda.empty_cache()

    out = {
        "metric": METRIC,
        "value": round(100.0 * gbs_wall / HBM_PEAK_GBS, 3),
        "unit": "%% of 8 TB/s HBM peak (f64 compound filter+SUM/COUNT/MIN/MAX, 2 columns x 2^%d rows)" % args.rows_log2,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (SplitMix64 uniform [0,1000), seeds 1 and 2, generated in HBM)",
        "config": {"workload": "fused scan-aggregate: a > 292.8932188 AND b < 707.1067812 (50%% selectivity), "
                               "SUM/COUNT/MIN/MAX of a, f64, 2^%d rows per column" % args.rows_log2,
                   "rows": n, "columns": 2, "selectivity_target": 0.5,
                   "selectivity_measured": round(res["count"] / n, 5)},
        "gb_per_s": round(gbs_wall, 1),
        "rows_per_s": round(n / (wall / args.steps), 1),
        "roofline": {"bound": "hbm", "achieved": round(gbs_kernel, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs_kernel / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_scan<double,2,...> (vx_scan.hip)", "algorithmic_bytes_per_launch": algo_bytes,
                     "avg_kernel_ms": round(kernel_ms, 4)},
        "breakdown": breakdown,
        "result_check": {"count": res["count"], "sum": res["sum"], "min": res["min"], "max": res["max"]},
    }
    del cols
    torch.cuda.empty_cache()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()

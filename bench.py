"""Benchmark of the VoxelVM hot path on one MI355X.

Headline (one step = one pass of the hot path over one batch): two 2^28-row
f64 columns a, b resident in HBM, uniform in [0, 1000) (SplitMix64 rows,
seeds 1 and 2), compound predicate a > t1 AND b < t2 at 50 % selectivity
(t1 = 292.8932188, t2 = 707.1067812), returning COUNT, SUM(a), MIN(a), MAX(a)
in one fused launch (voxel_gpu_scan_launch).  The same JSON line carries
* breakdown: f64/f32/i64/i32 x 1/50/99 % selectivity;
* sort: 2^29 u64 keys + u32 row-index payload, plain LSD radix vs the
  learned-CDF sort, uniform / Zipf(1.1, 2^32) / lognormal(20, 2) /
  95 %-sorted keys, Gkeys/s and output equality;
* join: 2^24 build keys (i64 payload), 2^28 probes (90 % drawn Zipf(1.0)
  over the build keys, 10 % absent), match count + SUM(payload), Gprobes/s;
* groupby: SUM/AVG(fare) GROUP BY zone over 3e8 rows (i32 zone 1..265,
  f64 fare lognormal(2.5, 0.6));
* serve: 4096 filter+aggregate queries over 64 resident f64/i64 columns
  (lengths lognormal, 1K..64M rows) batched vs one launch per query, at
  saturation and as a 20k queries/s Poisson arrival stream.
Zipf keys use the continuous inverse CDF of the power law (rank =
floor((1 - u(1 - N^(1-s)))^(1/(1-s)))), lognormal keys floor(exp(20 + 2Z)).

    python bench.py [--steps K] [--warmup W] [--no-cpu] [--no-breakdown]
                    [--no-sort] [--no-join] [--no-groupby] [--no-serve] [--no-window] [--no-nulls]
                    [--no-encodings] [--no-sql] [--no-expr]

value = % of the 8 TB/s HBM peak achieved by the f64 compound scan, from the
whole timed region (K launches bracketed by device synchronisation);
roofline.achieved = algorithmic bytes / average kernel duration from HIP
events recorded on the launch stream; cpu_baseline = the reference's own
compound-predicate operator path (oracle/_ref, single core) on a bounded
sample, with its JIT single-column filter+sum beside it.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import struct
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
    p.add_argument("--no-sort", action="store_true")
    p.add_argument("--no-join", action="store_true")
    p.add_argument("--no-groupby", action="store_true")
    p.add_argument("--no-serve", action="store_true")
    p.add_argument("--no-window", action="store_true")
    p.add_argument("--no-nulls", action="store_true")
    p.add_argument("--no-encodings", action="store_true")
    p.add_argument("--no-sql", action="store_true")
    p.add_argument("--no-expr", action="store_true")
    p.add_argument("--sort-log2", type=int, default=29)
    p.add_argument("--settle", type=float, default=1.0,
                   help="seconds of untimed headline launches before the warm-up (clocks at steady state)")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="also run a device-to-device copy of one 2^rows-log2 f64 column (the known-size "
                        "kernel that scales rocprofv3 FETCH_SIZE/WRITE_SIZE to bytes)")
    return p.parse_args()


_MARK = {}


def mark(stream_handle):
    """A one-row i32 fill (k_fill<int> with a one-workgroup grid) on the launch
    stream: in a rocprofv3 kernel trace the timed scans are the ones between
    two such dispatches.  Launched outside the timed region."""
    import torch
    from voxel_amd._lib import lib
    if "t" not in _MARK:
        _MARK["t"] = torch.empty(1, dtype=torch.int32, device="cuda")
    if lib.voxel_gpu_fill_uniform(3, _MARK["t"].data_ptr(), 1, 0, 0.0, 1.0, stream_handle) != 0:
        raise RuntimeError("marker fill failed")


def time_scan(desc, steps, warmup, stream_handle, settle_s=0.0, marks=False, settle_max_s=10.0, settle_chunks=80):
    """settle_s: untimed launches for at least that long first, and then on
    until the rate stops improving (40-launch chunks, settle_chunks in a row
    not 0.5 % faster than the best, at most settle_max_s in all): right after another
    process freed a lot of device memory, the first seconds run about 5 %
    slower (the driver clears the freed memory in the background), and the
    timed region is meant to see the kernel at steady state.  Then the W
    warm-up launches, then K timed launches (between two marker dispatches
    when marks is set)."""
    import torch
    from voxel_amd import scan
    if settle_s > 0:
        t_start = time.perf_counter()
        best, stable = float("inf"), 0
        while True:
            t0 = time.perf_counter()
            for _ in range(40):
                scan.launch(desc, stream_handle)
            torch.cuda.synchronize()
            now = time.perf_counter()
            dt = now - t0
            if dt < best * 0.995:
                best, stable = dt, 0
            else:
                stable += 1
            if now - t_start >= settle_max_s or (now - t_start >= settle_s and stable >= settle_chunks):
                break
    for _ in range(warmup):
        scan.launch(desc, stream_handle)
    if marks:
        mark(stream_handle)
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
    if marks:
        mark(stream_handle)
        torch.cuda.synchronize()
    return wall, kernel_ms


def _sync_time(fn, reps=3, warmup=1, settle_s=0.0):
    """Median wall time of reps calls of fn, each followed by a device
    synchronisation; warmup calls first, then further calls until settle_s
    seconds have passed (the clocks reach their steady state under load)."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t_end = time.perf_counter() + settle_s
    while time.perf_counter() < t_end:
        fn()
        torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts)


def sort_keys(dist: str, n: int, seed: int):
    """Synthetic u64 keys (as int64 tensors) of the four sort workloads."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    if dist == "uniform":
        return torch.randint(-2**63, 2**63 - 1, (n,), dtype=torch.int64, device="cuda", generator=g)
    if dist == "zipf":
        s_, N = 1.1, float(2**32)
        u = torch.rand(n, dtype=torch.float64, device="cuda", generator=g)
        r = torch.pow(1.0 - u * (1.0 - N ** (1.0 - s_)), 1.0 / (1.0 - s_))
        return torch.clamp(torch.floor(r) - 1.0, 0.0, N - 1.0).to(torch.int64)
    if dist == "lognormal":
        z = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
        return torch.clamp(torch.floor(torch.exp(20.0 + 2.0 * z)), 0.0, 2.0**62).to(torch.int64)
    if dist == "sorted95":
        k = torch.randint(0, 2**62, (n,), dtype=torch.int64, device="cuda", generator=g)
        k = torch.sort(k).values
        m = int(0.05 * n / 2)
        i = torch.randint(0, n, (m,), device="cuda", generator=g)
        j = torch.randint(0, n, (m,), device="cuda", generator=g)
        ki, kj = k[i].clone(), k[j].clone()
        k[i] = kj
        k[j] = ki
        return k
    raise ValueError(dist)


def bench_sort(log2n: int) -> dict:
    """Plain LSD radix vs learned-CDF sort of 2^log2n (u64 key, u32 row index)."""
    import ctypes as C
    import torch
    from voxel_amd._lib import LearnedStats, lib
    n = 1 << log2n
    out = {"n": n, "payload": "u32 row index", "algorithmic_bytes": 2 * n * 12,
           "note": "Gkeys/s = n / wall time of one sort call (inputs resident); bytes = one read + one write "
                   "of key+payload"}
    ko_l = torch.empty(n, dtype=torch.int64, device="cuda")
    vo_l = torch.empty(n, dtype=torch.int32, device="cuda")
    ko_r = torch.empty(n, dtype=torch.int64, device="cuda")
    vo_r = torch.empty(n, dtype=torch.int32, device="cuda")
    speedups = []
    for dist in ("uniform", "zipf", "lognormal", "sorted95"):
        k = sort_keys(dist, n, 7)
        torch.cuda.synchronize()
        st = LearnedStats()

        def lsd():
            if lib.voxel_gpu_sort_pairs_u64(k.data_ptr(), None, ko_r.data_ptr(), vo_r.data_ptr(), n, None) != 0:
                raise RuntimeError("lsd sort failed")

        def learned():
            if lib.voxel_gpu_sort_pairs_u64_learned(k.data_ptr(), None, ko_l.data_ptr(), vo_l.data_ptr(), n, 0,
                                                    C.byref(st), None) != 0:
                raise RuntimeError("learned sort failed")

        t_lsd = _sync_time(lsd)
        t_learned = _sync_time(learned)
        same = bool(torch.equal(ko_l, ko_r) and torch.equal(vo_l, vo_r))
        d = st.as_dict()
        speedups.append(t_lsd / t_learned)
        out[dist] = {"lsd_gkeys_per_s": round(n / t_lsd / 1e9, 3), "learned_gkeys_per_s": round(n / t_learned / 1e9, 3),
                     "lsd_ms": round(t_lsd * 1e3, 3), "learned_ms": round(t_learned * 1e3, 3),
                     "speedup": round(t_lsd / t_learned, 3), "bit_exact": same,
                     "learned_achieved_gb_per_s": round(2 * n * 12 / t_learned / 1e9, 1),
                     "partition_passes": d["partition_passes"], "buckets": d["buckets"], "fell_back": d["fell_back"],
                     "big_rows": d["big_rows"], "final_logit_mse": round(d["final_loss"], 5),
                     "phase_ms": {k: round(d[k + "_ms"], 3) for k in ("sample", "train", "table", "partition",
                                                                       "bucket", "fixup")}}
        del k
        torch.cuda.empty_cache()
    out["speedup_geomean"] = round(math.exp(sum(math.log(x) for x in speedups) / len(speedups)), 3)
    del ko_l, vo_l, ko_r, vo_r
    torch.cuda.empty_cache()
    # read + write rate of a plain device copy of the same n x 12 bytes: the
    # practical ceiling of a digit pass, which reads and writes every row
    src = torch.empty(n * 12, dtype=torch.uint8, device="cuda")
    src.fill_(1)
    dst = torch.empty_like(src)
    t_copy = _sync_time(lambda: dst.copy_(src), reps=5)
    out["device_copy_rw_gb_per_s"] = round(2 * n * 12 / t_copy / 1e9, 1)
    del src, dst
    torch.cuda.empty_cache()
    return out


def bench_join() -> dict:
    """Build 2^24 distinct i64 keys + i64 payloads; probe 2^28 keys (90 % hits
    drawn Zipf(1.0) over the build rows through a fixed permutation, 10 %
    absent: top bit set, build keys have it clear)."""
    import ctypes as C
    import torch
    from voxel_amd._lib import lib
    nb, npr = 1 << 24, 1 << 28
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    while True:
        bk = torch.randint(0, 2**63 - 1, (nb,), dtype=torch.int64, device="cuda", generator=g)
        if torch.unique(bk).numel() == nb:
            break
    pay = torch.randint(-2**63, 2**63 - 1, (nb,), dtype=torch.int64, device="cuda", generator=g)
    perm = torch.randperm(nb, device="cuda", generator=g)
    u = torch.rand(npr, dtype=torch.float64, device="cuda", generator=g)
    rank = torch.clamp(torch.floor(torch.pow(float(nb), u)), 1, nb).to(torch.int64)      # Zipf(1.0) ranks
    row = perm[rank - 1]
    hit = torch.rand(npr, device="cuda", generator=g) < 0.9
    miss = torch.randint(0, 2**63 - 1, (npr,), dtype=torch.int64, device="cuda", generator=g) | (-2**63)
    probe = torch.where(hit, bk[row], miss)
    exp_count = int(hit.sum().item())
    exp_sum = int(torch.where(hit, pay[row], torch.zeros_like(pay[row])).sum().item())
    del u, rank, miss
    torch.cuda.synchronize()
    from voxel_amd._lib import JoinStats
    layouts = {}
    for mode, name in ((0, "auto"), (1, "one_table"), (3, "learned_range"), (2, "hash")):
        holder = {}
        st = JoinStats()

        def build():
            if "j" in holder:
                lib.voxel_gpu_join_destroy(holder.pop("j"))
            holder["j"] = lib.voxel_gpu_join_build_ex(bk.data_ptr(), pay.data_ptr(), nb, mode, C.byref(st), None)
            if not holder["j"]:
                raise RuntimeError("join build failed")

        t_build = _sync_time(build)
        cnt, sm = C.c_uint64(), C.c_int64()

        def probe_fn():
            if lib.voxel_gpu_join_probe_sum(holder["j"], probe.data_ptr(), npr, C.byref(cnt), C.byref(sm), None) != 0:
                raise RuntimeError("join probe failed")

        t_probe = _sync_time(probe_fn)
        lib.voxel_gpu_join_destroy(holder.pop("j"))
        d = st.as_dict()
        layouts[name] = {"gprobes_per_s": round(npr / t_probe / 1e9, 3), "probe_ms": round(t_probe * 1e3, 3),
                         "build_ms": round(t_build * 1e3, 3), "mode": d["mode"], "fell_back": d["fell_back"],
                         "partitions": d["partitions"], "max_partition_keys": d["max_partition_keys"],
                         "min_partition_keys": d["min_partition_keys"],
                         "matches_expected": cnt.value == exp_count and sm.value == exp_sum}
        if mode == 3:
            layouts[name]["model_train_ms"] = round(d["model"]["train_ms"], 3)
            layouts[name]["model_final_logit_mse"] = round(d["model"]["final_loss"], 5)
    a = layouts["auto"]
    out = {"build_rows": nb, "probe_rows": npr, "gprobes_per_s": a["gprobes_per_s"],
           "probe_ms": a["probe_ms"], "build_ms": a["build_ms"],
           "build_mrows_per_s": round(nb / (a["build_ms"] * 1e-3) / 1e6, 1),
           "probe_algorithmic_gb_per_s": round(npr * 24 / (a["probe_ms"] * 1e-3) / 1e9, 1),
           "match_count": exp_count, "payload_sum": exp_sum,
           "matches_expected": all(v["matches_expected"] for v in layouts.values()),
           "layouts": layouts,
           "note": "auto = one 2^25-slot table, direct probes; hash / learned_range = 256 sub-tables of 2^17 "
                   "slots, probes partitioned by sub-table (one digit pass) before probing"}
    del bk, pay, perm, row, hit, probe
    torch.cuda.empty_cache()
    return out


def bench_groupby() -> dict:
    import torch
    from voxel_amd._lib import lib
    out = {}
    for rows in (3_000_000, 300_000_000):
        zone = torch.empty(rows, dtype=torch.int32, device="cuda")
        assert lib.voxel_gpu_fill_uniform(3, zone.data_ptr(), rows, 21, 1.0, 266.0, None) == 0
        g = torch.Generator(device="cuda")
        g.manual_seed(22)
        fare = torch.exp(2.5 + 0.6 * torch.randn(rows, dtype=torch.float64, device="cuda", generator=g))
        G = 512
        res = [torch.zeros(G, dtype=torch.int64 if i < 2 else torch.float64, device="cuda") for i in range(6)]
        import numpy as np
        ng = np.zeros(1, np.uint64)

        def run():
            if lib.voxel_gpu_groupby_f64(3, zone.data_ptr(), fare.data_ptr(), rows, G, *(t.data_ptr() for t in res),
                                         ng.ctypes.data, None) != 0:
                raise RuntimeError("groupby failed")

        t = _sync_time(run, reps=30, warmup=5, settle_s=0.25)
        out["rows_%d" % rows] = {"ms": round(t * 1e3, 3), "groups": int(ng[0]),
                                 "gb_per_s": round(rows * 12 / t / 1e9, 1),
                                 "frac_of_8tbs": round(rows * 12 / t / 8e12, 4),
                                 "count_total_ok": int(res[1][: int(ng[0])].sum().item()) == rows}
        del zone, fare, res
        torch.cuda.empty_cache()
    return out


def bench_nulls(rows_log2: int = 28) -> dict:
    """Scans over nullable columns: the headline compound filter + aggregate
    (2 f64 columns of 2^28 rows) with a validity bitmap per column (about 10 %
    nulls), NullableSegment SumValid/MinValid/MaxValid and CompactValid."""
    import ctypes as C
    import torch
    from voxel_amd import scan as vs
    from voxel_amd._lib import Agg, lib
    n = 1 << rows_log2
    a = torch.empty(n, dtype=torch.float64, device="cuda")
    b = torch.empty(n, dtype=torch.float64, device="cuda")
    assert lib.voxel_gpu_fill_uniform(0, a.data_ptr(), n, 1, 0.0, 1000.0, None) == 0
    assert lib.voxel_gpu_fill_uniform(0, b.data_ptr(), n, 2, 0.0, 1000.0, None) == 0
    nw = (n + 63) // 64
    g = torch.Generator(device="cuda")
    g.manual_seed(5)

    def words():
        bits = torch.rand(nw, 64, device="cuda", generator=g) < 0.9
        w = torch.zeros(nw, dtype=torch.int64, device="cuda")
        for k in range(64):
            w |= bits[:, k].to(torch.int64) << k
        return w

    wa, wb = words(), words()
    d = vs.make_desc(a, vs.GT, 292.8932188, b, vs.LT, 707.1067812)
    r = Agg()

    def run():
        if lib.voxel_gpu_scan_valid(C.byref(d), C.c_void_p(wa.data_ptr()), C.c_void_p(wb.data_ptr()), None,
                                    C.byref(r)) != 0:
            raise RuntimeError("scan_valid failed")

    def run_plain():
        if lib.voxel_gpu_scan(C.byref(d), None, C.byref(r)) != 0:
            raise RuntimeError("scan failed")

    t = _sync_time(run, reps=5)
    tp = _sync_time(run_plain, reps=5)
    bytes_ = n * 16 + 2 * nw * 8
    out = {"scan_valid_compound_2x2^%d" % rows_log2: {
        "ms": round(t * 1e3, 4), "gb_per_s": round(bytes_ / t / 1e9, 1),
        "frac_of_peak": round(bytes_ / t / 1e9 / HBM_PEAK_GBS, 4), "count": int(r.count),
        "same_scan_without_bitmaps_ms": round(tp * 1e3, 4)}}
    t = _sync_time(lambda: vs.nullable_reduce(a, wa), reps=5)
    out["nullable_reduce_2^%d" % rows_log2] = {"ms": round(t * 1e3, 4),
                                               "gb_per_s": round((n * 8 + nw * 8) / t / 1e9, 1)}
    o_ = torch.empty_like(a)
    m = [0]
    t = _sync_time(lambda: m.__setitem__(0, vs.nullable_compact(a, wa, o_)), reps=5)
    out["nullable_compact_2^%d" % rows_log2] = {"ms": round(t * 1e3, 4), "kept": m[0],
                                                "gb_per_s": round((n * 8 + nw * 8 + m[0] * 8) / t / 1e9, 1)}
    del a, b, wa, wb, o_
    torch.cuda.empty_cache()
    return out


def bench_encodings(rows_log2: int = 28) -> dict:
    """Filter + aggregate scans that read encoded columns (2^28 rows): the
    dictionary (1000 distinct f64, u32 indices), frame-of-reference (i32
    codes), 12-bit packed u32, RLE (runs of 16) and delta (i64) forms, beside
    the plain f64 scan of the same rows."""
    import ctypes as C
    import torch
    from voxel_amd import encoding as enc
    from voxel_amd import scan as vs
    from voxel_amd._lib import lib
    n = 1 << rows_log2
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    out = {}
    d = torch.sort(torch.randn(1000, dtype=torch.float64, device="cuda", generator=g) * 100).values
    ix = torch.randint(0, 1000, (n,), dtype=torch.int32, device="cuda", generator=g)
    plain = d[ix.long()]
    t = float(d[500].item())

    def entry(name, fn, nbytes, rows=n):
        ms = _sync_time(fn, reps=5) * 1e3
        out[name] = {"ms": round(ms, 4), "rows_per_s": round(rows / ms * 1e3, 1),
                     "encoded_gb_per_s": round(nbytes / ms / 1e6, 1),
                     "frac_of_peak": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 4)}

    dcol = enc.dictionary(d, ix)
    entry("plain_f64_scan", lambda: vs.scan(plain, vs.GT, t), n * 8)
    entry("dictionary_scan", lambda: dcol.scan(vs.GT, t), n * 4 + 8000)
    del plain
    torch.cuda.empty_cache()
    codes = torch.randint(0, 2**31 - 1, (n,), dtype=torch.int32, device="cuda", generator=g)
    fcol = enc.frame_of_reference(codes, 1000.0, 1e-6)
    entry("frame_of_reference_scan", lambda: fcol.scan(vs.GT, 2000.0), n * 4)
    del codes, fcol
    words = torch.randint(-2**62, 2**62, (n * 12 // 64,), dtype=torch.int64, device="cuda", generator=g)
    bcol = enc.bitpacked(words, 12, n)
    entry("bitpacked12_scan", lambda: bcol.scan(vs.LT, 1000), n * 12 // 8)
    del words, bcol
    nr = n // 16
    vals = torch.randn(nr, dtype=torch.float64, device="cuda", generator=g)
    lens = torch.full((nr,), 16, dtype=torch.int32, device="cuda")
    rcol = enc.rle(vals, lens)
    entry("rle16_scan", lambda: rcol.scan(vs.GT, 0.0), nr * 12)
    del vals, lens, rcol
    deltas = torch.randint(-1000, 1000, (n,), dtype=torch.int64, device="cuda", generator=g)
    dl = enc.delta(deltas, 0)
    entry("delta_i64_decode_and_scan", lambda: dl.scan(vs.GT, 0), n * 8)
    o_ = torch.empty(n, dtype=torch.float64, device="cuda")
    entry("dictionary_decode", lambda: dcol.decode(0, n, o_), n * 12)
    del deltas, dl, o_, dcol, d, ix
    torch.cuda.empty_cache()
    return out


def bench_sql(rows_log2: int = 28) -> dict:
    """SQL over HBM-resident columns (2 x 2^28 f64): a compound WHERE on two
    columns lowered to one fused filter + aggregate pass, and a three-column
    condition through bitmaps + the validity-word scan; wall time per
    statement including parsing and planning."""
    import torch
    from voxel_amd import voxel_sql as vsql
    from voxel_amd._lib import lib
    n = 1 << rows_log2
    a = torch.empty(n, dtype=torch.float64, device="cuda")
    b = torch.empty(n, dtype=torch.float64, device="cuda")
    c = torch.empty(n, dtype=torch.float64, device="cuda")
    for i, col in enumerate((a, b, c)):
        assert lib.voxel_gpu_fill_uniform(0, col.data_ptr(), n, i + 1, 0.0, 1000.0, None) == 0
    data = {"t": {"a": a, "b": b, "c": c}}
    out = {}
    for name, sql, nbytes in (
            ("two_column_fused", "SELECT SUM(a), COUNT(*) FROM t WHERE a > 292.8932188 AND b < 707.1067812", n * 16),
            ("three_column_fused", "SELECT SUM(a) FROM t WHERE a > 292.89 AND b < 707.1 AND c >= 500", n * 24),
            ("three_column_or", "SELECT SUM(a) FROM t WHERE a > 292.89 AND (b < 707.1 OR c >= 500)", n * 24)):
        res = [None]
        t = _sync_time(lambda: res.__setitem__(0, vsql.execute(sql, data)), reps=5)
        out[name] = {"ms": round(t * 1e3, 4), "gb_per_s": round(nbytes / t / 1e9, 1), "sql": sql,
                     "result": res[0] if not isinstance(res[0], tuple) else list(res[0])}
    del a, b, c, data
    torch.cuda.empty_cache()
    return out


# The reference's own figures for its window family (docs/Interpreter.md:58-65:
# 1M f64 rows, window 20; quantile window 100), milliseconds on its CPU.
WINDOW_DOC_MS = {"sum": 12.0, "mean": 15.0, "delta": 2.0, "stddev": 39.0, "variance": 37.0, "quantile": 1000.0}


def _window_calls(lib, rows: int):
    return (("delta", lambda x, o: lib.voxel_gpu_window_delta_f64(x, rows, o, rows, None)),
            ("sum", lambda x, o: lib.voxel_gpu_window_sum_f64(x, rows, o, rows, 20, None)),
            ("mean", lambda x, o: lib.voxel_gpu_window_mean_f64(x, rows, o, rows, 20, None)),
            ("stddev", lambda x, o: lib.voxel_gpu_window_stddev_f64(x, rows, o, rows, 20, None)),
            ("variance", lambda x, o: lib.voxel_gpu_window_variance_f64(x, rows, o, rows, 20, None)),
            ("quantile", lambda x, o: lib.voxel_gpu_window_quantile_f64(x, rows, o, rows, 100, 0.05, None)))


def bench_window() -> dict:
    """EngineF64 window family on HBM-resident columns: the reference's 1M-row
    table (window 20, quantile window 100) and a 2^27-row column."""
    import torch
    from voxel_amd._lib import lib
    out = {}
    for rows in (1_000_000, 1 << 27):
        x = torch.empty(rows, dtype=torch.float64, device="cuda")
        assert lib.voxel_gpu_fill_uniform(0, x.data_ptr(), rows, 31, 0.0, 1000.0, None) == 0
        y = torch.empty(rows, dtype=torch.float64, device="cuda")
        res = {}
        for kind, call in _window_calls(lib, rows):
            if kind == "quantile" and rows > 1_000_000:
                continue

            def run():
                if call(x.data_ptr(), y.data_ptr()) < 0:
                    raise RuntimeError("window %s failed" % kind)

            t = _sync_time(run, reps=5)
            r = {"ms": round(t * 1e3, 4), "gb_per_s": round(rows * 16 / t / 1e9, 1)}
            if rows == 1_000_000:
                r["reference_doc_ms"] = WINDOW_DOC_MS[kind]
            res[kind] = r
        out["rows_%d" % rows] = res
        del x, y
        torch.cuda.empty_cache()
    return out


def _word(op, rd=0, ra=0, rb=0, imm=0):
    """One bytecode word [op:8][rd:4][ra:4][rb:4][imm12:12]."""
    return (op & 0xFF) | (rd & 0xF) << 8 | (ra & 0xF) << 12 | (rb & 0xF) << 16 | (imm & 0xFFF) << 20


def expr_programs() -> dict:
    """Tier-0 loops with vector words over segment 0 (R1 offset, R2 rows,
    R4 = 4 lanes, R3 a scalar operand): name -> (words, reduced-register
    accumulator)."""
    def loop(body, acc_words):
        n = 2 + len(body) + len(acc_words) + 3
        return ([_word(0x10, rd=4, imm=4), _word(0x70, rd=0, ra=1)] + body + acc_words +
                [_word(0x20, rd=1, ra=1, rb=4), _word(0x40, ra=1, rb=2), _word(0x52, imm=(1 - (n - 1)) & 0xFFF),
                 _word(0x01)])
    vsum = [_word(0xD0, rd=5, ra=2), _word(0x2A, rd=0, ra=0, rb=5)]
    return {
        # sum(x * x + x)
        "sum_x2_plus_x": (loop([_word(0x82, rd=1, ra=0, rb=0), _word(0x80, rd=2, ra=1, rb=0)], vsum), 0),
        # count(x * x > R3)
        "count_x2_gt_t": (loop([_word(0x82, rd=1, ra=0, rb=0), _word(0xC5, rd=2, ra=1, rb=3)],
                               [_word(0xD7, rd=5, ra=2), _word(0x20, rd=7, ra=7, rb=5)]), 7),
        # sum(sqrt(|x * R3 - R3|) where x * R3 - R3 >= R3), six words
        "sum_sqrt_abs_affine_filtered": (loop([_word(0x92, rd=1, ra=0, rb=3), _word(0x91, rd=2, ra=1, rb=3),
                                               _word(0xC6, rd=3, ra=2, rb=3), _word(0x86, rd=6, ra=3),
                                               _word(0x8C, rd=2 + 5, ra=6), _word(0x80, rd=2 + 6, ra=7, rb=0)],
                                              [_word(0xD0, rd=5, ra=8), _word(0x2A, rd=0, ra=0, rb=5)]), 0),
    }


def bench_expr(rows_log2: int = 27) -> dict:
    """Tier-0 loops with vector words on an HBM-resident f64 column (2^27
    rows): compiled whole at run time (the default for loops of 2^20 rows or
    more; compile_ms is the first run's extra time, paid once per program
    shape and process) and on the interpreter kernels, beside the plain VSUM
    loop.  Algorithmic bytes: 8 per row (one read of the column)."""
    import ctypes as C
    import torch
    from voxel_amd._lib import check, lib
    n = 1 << rows_log2
    x = torch.empty(n, dtype=torch.float64, device="cuda")
    assert lib.voxel_gpu_fill_uniform(0, C.c_void_p(x.data_ptr()), n, 41, -2.0, 2.0, None) == 0
    torch.cuda.synchronize()
    out = {"rows": n}

    def run_loop(words, acc_reg, mode, reps=7):
        prev = lib.voxel_gpu_set_expr_jit(mode)
        e = lib.voxel_gpu_engine_create(0)
        try:
            lib.voxel_gpu_engine_add_device_segment(e, C.c_void_p(x.data_ptr()), n)
            arr = (C.c_uint32 * len(words))(*words)
            lib.voxel_engine_load_program(e, arr, len(words))
            ts = []
            for _ in range(reps):
                for r in (0, 1, 5, 7):
                    lib.voxel_engine_set_scalar(e, r, 0)
                lib.voxel_engine_set_scalar(e, 2, n)
                lib.voxel_engine_set_scalar_f64(e, 3, 0.5)
                t0 = time.perf_counter()
                lib.voxel_engine_run(e)
                ts.append(time.perf_counter() - t0)
                check("run")
            return ts[0], statistics.median(ts[1:]), lib.voxel_engine_get_scalar(e, acc_reg)
        finally:
            lib.voxel_engine_destroy(e)
            lib.voxel_gpu_set_expr_jit(prev)

    plain = [_word(0x10, rd=4, imm=4), _word(0x70, rd=0, ra=1), _word(0xD0, rd=5, ra=0), _word(0x2A, rd=0, ra=0, rb=5),
             _word(0x20, rd=1, ra=1, rb=4), _word(0x40, ra=1, rb=2), _word(0x52, imm=(1 - 6) & 0xFFF), _word(0x01)]
    _, t_plain, _ = run_loop(plain, 0, 1)
    out["plain_vsum"] = {"ms": round(t_plain * 1e3, 4), "gb_per_s": round(n * 8 / t_plain / 1e9, 1)}
    for name, (words, acc) in expr_programs().items():
        first, tc, rc = run_loop(words, acc, 1)
        _, ti, ri = run_loop(words, acc, 0, reps=4)
        out[name] = {"compiled_ms": round(tc * 1e3, 4), "compiled_gb_per_s": round(n * 8 / tc / 1e9, 1),
                     "frac_of_peak": round(n * 8 / tc / 1e9 / HBM_PEAK_GBS, 4),
                     "compile_ms": round((first - tc) * 1e3, 1),
                     "interpreted_ms": round(ti * 1e3, 4), "speedup_vs_interpreted": round(ti / tc, 2),
                     "same_result": rc == ri}
        if acc != 7:                                   # f64 sum register: also how far apart, relative
            fc, fi = (struct.unpack("<d", struct.pack("<Q", v & (2**64 - 1)))[0] for v in (rc, ri))
            out[name]["rel_diff_vs_interpreted"] = abs(fc - fi) / abs(fi) if fi else abs(fc - fi)
    del x
    torch.cuda.empty_cache()
    return out


def cpu_reference_paths() -> dict:
    """The reference's other hot-path CPU implementations (built from its
    sources into oracle/_ref), each on a bounded sample of its workload on
    this host: the interpreter fast path (one core), RunParallel (all the
    host cores given to this process, at most 16), the Tier-0 interpreter on
    a loop with vector words (one core), RadixSort<u64> (one core),
    HashJoin<i64> Build + Probe (one core) and HashAggregator<i32,f64> (one
    core); median of 3 after a warm-up."""
    import numpy as np
    from oracle import oracle as o
    if o.R is None:
        return None

    def med(fn, reps=3):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return statistics.median(ts)

    out = {}
    rng = np.random.default_rng(5)
    code = np.array([0x00001070, 0x000301C5, 0x000015D0, 0x0005002A, 0x00041120, 0x00021040, 0xFFA00052,
                     0x00000001], np.uint32)
    n = 1 << 25
    x = rng.random(n) * 1000.0

    def fast():
        r = np.zeros(64, np.uint64)
        r[2] = n
        r[3] = o.f64_bits(500.0)
        r[4] = 4
        o.R.vr_engine_run_f64(o.ptr(x), n, o.ptr(code), code.size, o.ptr(r))
    t = med(fast)
    out["fast_path_f64"] = {"rows_per_s": round(n / t, 1), "gb_per_s": round(n * 8 / t / 1e9, 3), "cores": 1,
                            "sample": "Engine<f64>::Run fast path, 2^25 rows, x > 500"}
    threads = min(os.cpu_count() or 1, 16)

    def par():
        r = np.zeros(64, np.uint64)
        r[2] = n
        r[3] = o.f64_bits(500.0)
        r[4] = 4
        o.R.vr_engine_run_parallel_f64(o.ptr(x), n, o.ptr(code), code.size, o.ptr(r), threads, 0)
    t = med(par)
    out["run_parallel_f64"] = {"rows_per_s": round(n / t, 1), "gb_per_s": round(n * 8 / t / 1e9, 3),
                               "cores": threads, "sample": "Engine<f64>::RunParallel, 2^25 rows, x > 500"}
    # config 2's compound predicate the reference's operator way, one core
    b = rng.random(n) * 1000.0
    res = np.zeros(4)
    t = med(lambda: o.R.vr_compound_f64(o.ptr(x), o.ptr(b), n, 292.8932188, 707.1067812, o.ptr(res)))
    out["compound_filter_f64"] = {
        "rows_per_s": round(n / t, 1), "gb_per_s": round(2 * n * 8 / t / 1e9, 3), "cores": 1,
        "pct_of_8tbs": round(100 * 2 * n * 8 / t / 1e9 / HBM_PEAK_GBS, 4),
        "selectivity_measured": round(float(res[0]) / n, 5),
        "sample": "VectorFilter<f64>::ApplyGT(a) + ApplyLT(b), ColumnFilter::Init + AndWith x2, "
                  "SelectionVector::BuildFromBitmap, then COUNT/SUM/MIN/MAX of a over the selected rows; "
                  "2^25 rows per column, a > 292.8932188 AND b < 707.1067812 (50 %)"}
    del b
    # a Tier-0 loop with vector words (sum(x * x + x), the bench's
    # tier0_expression loop) in the reference's interpreter, one core
    words = np.array(expr_programs()["sum_x2_plus_x"][0], np.uint32)
    ne = 1 << 24
    xe = x[:ne].copy()

    def expr():
        r = np.zeros(64, np.uint64)
        r[2] = ne
        o.R.vr_engine_run_f64(o.ptr(xe), ne, o.ptr(words), words.size, o.ptr(r))
    t = med(expr)
    out["tier0_expression_f64"] = {"rows_per_s": round(ne / t, 1), "gb_per_s": round(ne * 8 / t / 1e9, 3), "cores": 1,
                                   "sample": "Engine<f64>::Run of the loop V1 = V0 * V0, V2 = V1 + V0, VSUM, ADDF "
                                             "(Tier-0 interpreter), 2^24 rows"}
    del x, xe
    ns = 1 << 24
    keys0 = rng.integers(0, 2**63, ns, dtype=np.int64).astype(np.uint64) * np.uint64(2)

    def rsort():
        k = keys0.copy()
        idx = np.arange(ns, dtype=np.uint32)
        o.R.vr_radix_sort_u64(o.ptr(k), o.ptr(idx), ns)
    t = med(rsort)
    out["radix_sort_u64"] = {"gkeys_per_s": round(ns / t / 1e9, 4), "cores": 1,
                             "sample": "RadixSort<u64>::Sort with u32 row ids, 2^24 uniform keys (array copy included)"}
    del keys0
    nb, npr = 1 << 19, 1 << 24          # the reference's arena fails above 2^19 build rows (16 MiB blocks)
    build = rng.integers(0, 2**62, nb, dtype=np.int64)
    pay = rng.integers(-2**40, 2**40, nb, dtype=np.int64)
    probe = np.where(rng.random(npr) < 0.9, build[rng.integers(0, nb, npr)],
                     rng.integers(0, 2**62, npr, dtype=np.int64) | np.int64(2**62)).astype(np.int64)
    left = np.zeros(npr, np.uint32)
    right = np.zeros(npr, np.uint32)
    ps = np.zeros(1, np.int64)
    t = med(lambda: o.R.vr_hash_join_i64(o.ptr(build), nb, o.ptr(pay), o.ptr(probe), npr, o.ptr(left), o.ptr(right),
                                         o.ptr(ps)), reps=2)
    out["hash_join_i64"] = {"gprobes_per_s": round(npr / t / 1e9, 4), "cores": 1,
                            "sample": "HashJoin<i64> Build(2^19) + Probe(2^24, 90 % hits) + payload sum; larger builds "
                                      "crash the reference (its Arena allocates at most 16 MiB per block)"}
    del build, pay, probe, left, right
    m = 3_000_000
    zone = rng.integers(1, 266, m).astype(np.int32)
    fare = np.exp(2.5 + 0.6 * rng.standard_normal(m))
    G = 1024
    bufs = [np.zeros(G, t_) for t_ in (np.int32, np.uint64, np.float64, np.float64, np.float64, np.float64,
                                       np.float64)]
    t = med(lambda: o.R.vr_hash_agg_i32_f64(o.ptr(zone), o.ptr(fare), m, 512, G, *(o.ptr(b_) for b_ in bufs)))
    out["hash_agg_i32_f64"] = {"rows_per_s": round(m / t, 1), "gb_per_s": round(m * 12 / t / 1e9, 3), "cores": 1,
                               "sample": "HashAggregator<i32,f64> Init + Accumulate + Finalize, 3e6 rows, 265 zones"}
    return out


def cpu_window_baseline() -> dict:
    """The reference's window methods (Engine<f64>::Window*, built from its
    sources into oracle/_ref) on one host core: 1M rows, window 20 (quantile:
    window 100 on the first 100K rows, scaled to 1M)."""
    import numpy as np
    from oracle import oracle as o
    if o.R is None:
        return None
    x = np.random.default_rng(3).random(1_000_000) * 1000.0
    y = np.zeros_like(x)
    res = {}
    for k, kind in enumerate(("delta", "sum", "mean", "stddev", "variance", "quantile")):
        n, w, scale = (100_000, 100, 10.0) if kind == "quantile" else (x.size, 20, 1.0)
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            o.R.vr_window_f64(k, o.ptr(x), n, o.ptr(y), n, w, 0.05)
            ts.append(time.perf_counter() - t0)
        res[kind] = {"ms_1m": round(statistics.median(ts) * scale * 1e3, 3)}
    return {"cores": 1, "kind": "reference", "host_cpu": _cpu_model(), "kinds": res}


def bench_serve() -> dict:
    import numpy as np
    import torch
    from voxel_amd import scan
    from voxel_amd._lib import lib
    from voxel_amd.serve import Server
    rng = np.random.default_rng(2024)
    ncol, nq = 64, 4096
    lengths = np.clip(np.floor(np.exp(rng.normal(math.log(2**20), 2, ncol))), 1000, 64 * 2**20).astype(np.int64)
    cols = []
    for i, n in enumerate(lengths):
        f = i < 32
        c = torch.empty(int(n), dtype=torch.float64 if f else torch.int64, device="cuda")
        assert lib.voxel_gpu_fill_uniform(0 if f else 2, c.data_ptr(), int(n), 100 + i, 0.0,
                                          1000.0 if f else float(2**31), None) == 0
        cols.append(c)
    torch.cuda.synchronize()
    pick = rng.integers(0, ncol, nq)
    sel = rng.uniform(0.01, 0.99, nq)
    qs = [scan.make_desc(cols[int(i)], "lt", float(s * 1000.0) if int(i) < 32 else int(s * 2**31))
          for i, s in zip(pick, sel)]
    arrivals = np.cumsum(rng.exponential(1 / 20000.0, nq))
    srv = Server(max_batch=nq, memory_budget=256 << 20, inflight=2)
    tune = srv.autotune(qs, reps=1)
    res = {}
    # batched: the live request loop fed at the arrivals; unbatched: one
    # synchronous launch per query; unbatched_async: one launch per query
    # queued back to back without waiting, results read through events
    for mode, name in ((0, "batched"), (1, "unbatched"), (2, "unbatched_async")):
        best = None
        for _ in range(2):                       # saturated: all queries at t = 0; best of two replays
            out, lat, st = srv.replay_raw(qs, np.zeros(nq), mode=mode)
            if best is None or st["makespan_s"] < best["makespan_s"]:
                best = st
        for load, st in (("saturated", best), ("poisson_20k_qps", srv.replay_raw(qs, arrivals, mode=mode)[2])):
            res["%s_%s" % (name, load)] = {"queries_per_s": round(st["qps"], 1), "gb_per_s": round(st["gbps"], 1),
                                          "p50_us": round(st["p50_s"] * 1e6, 1), "p99_us": round(st["p99_s"] * 1e6, 1),
                                          "makespan_ms": round(st["makespan_s"] * 1e3, 3), "batches": st["batches"],
                                          "mean_batch": round(st["mean_batch"], 2)}
    res["throughput_ratio_saturated"] = round(res["batched_saturated"]["queries_per_s"] /
                                              res["unbatched_saturated"]["queries_per_s"], 3)
    res["throughput_ratio_saturated_vs_async"] = round(res["batched_saturated"]["queries_per_s"] /
                                                       res["unbatched_async_saturated"]["queries_per_s"], 3)
    res["autotune"] = tune
    # one query alone, no arrival schedule: a batch of one (the server's
    # small-batch kernels) against one synchronous voxel_gpu_scan, wall time
    # of the call, on the columns of about 1M and 8M rows
    from voxel_amd.serve import Batch
    single = {}
    b1 = Batch(16)
    for want in (1 << 20, 1 << 23):
        ci = int(np.argmin(np.abs(lengths - want) + (np.arange(ncol) >= 32) * 2**40))   # an f64 column
        q1 = scan.make_desc(cols[ci], "lt", 500.0)

        def lat(fn, reps=300):
            for _ in range(20):
                fn()
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                fn()
                ts.append(time.perf_counter() - t0)
            ts.sort()
            return round(ts[reps // 2] * 1e6, 1), round(ts[int(reps * 0.99)] * 1e6, 1)
        pb = lat(lambda: b1.run_raw([q1]))
        ps = lat(lambda: scan.scan(cols[ci], "lt", 500.0))
        single["rows_%d" % int(lengths[ci])] = {"batch_of_one_p50_p99_us": pb, "voxel_gpu_scan_p50_p99_us": ps}
    b1.close()
    res["single_query_latency"] = single
    # a saturated batch holds all 4096 queries: queries on one column share a
    # scan, up to 4 per job, so the rows read are fewer than the rows queried
    per_col = np.bincount(pick, minlength=ncol)
    physical = sum(int(math.ceil(c / 4)) * int(lengths[i]) * 8 for i, c in enumerate(per_col))
    logical = sum(int(lengths[int(i)]) * 8 for i in pick)
    res["bytes_note"] = ("gb_per_s = bytes queried (column bytes x queries) / makespan; a saturated batch reads "
                         "%.1f GB of the %.1f GB queried (shared scans of up to 4 queries per column)" %
                         (physical / 1e9, logical / 1e9))
    res["batched_saturated"]["physical_gb_per_s"] = round(physical / (res["batched_saturated"]["makespan_ms"] * 1e-3) / 1e9, 1)
    res["workload"] = ("4096 queries (x < t, selectivity U[1%%,99%%]) over 64 resident columns (32 f64, 32 i64), "
                       "lengths floor(exp(N(ln 2^20, 2))) clipped to [1e3, 2^26], %.1f GB queried" % (logical / 1e9))
    srv.close()
    del cols, qs
    torch.cuda.empty_cache()
    return res


def cpu_baseline(n_rows_log2: int = 26):
    """The reference's CPU path for the headline workload (config 2's
    compound predicate): its operators VectorFilter<f64>::ApplyGT / ApplyLT,
    ColumnFilter AndWith and SelectionVector, then COUNT/SUM/MIN/MAX of a
    over the selected rows, built from the reference sources into
    oracle/_ref, on one core, 2^n_rows_log2 rows per column; median of 3
    after a warm-up.  Beside it, the reference's fastest single-column path,
    its x86-64 JIT fused filter+sum loop (codegen/x86_64.hpp:1632-1716), on
    2^27 rows (2 warmups + median of 3, validate.cpp:20)."""
    import numpy as np
    from oracle import oracle as o
    if o.R is None:
        return None
    n = 1 << n_rows_log2
    rng = np.random.default_rng(1)
    a = rng.random(n) * 1000.0
    b = rng.random(n) * 1000.0
    res = np.zeros(4)
    ts = []
    for i in range(4):
        t0 = time.perf_counter()
        o.R.vr_compound_f64(o.ptr(a), o.ptr(b), n, 292.8932188, 707.1067812, o.ptr(res))
        ts.append(time.perf_counter() - t0)
    med = statistics.median(ts[1:])
    gbs = 2 * n * 8 / med / 1e9
    ref = o.relational(a, 4, 292.8932188, b, 2, 707.1067812)
    out = {"value": round(100.0 * gbs / HBM_PEAK_GBS, 4), "unit": "% of 8 TB/s HBM peak",
           "gb_per_s": round(gbs, 3), "rows_per_s": round(n / med, 1), "cores": 1, "kind": "reference",
           "sample": "reference operators for a > 292.8932188 AND b < 707.1067812 (VectorFilter<f64> ApplyGT / "
                     "ApplyLT, ColumnFilter Init + AndWith x2, SelectionVector::BuildFromBitmap, then COUNT/SUM/"
                     "MIN/MAX of a over the selected rows in row order), 2^%d f64 rows per column uniform "
                     "[0,1000) (16 B/row), median of 3 after a warm-up, one core" % n_rows_log2,
           "host_cpu": _cpu_model(), "nproc": os.cpu_count(),
           "result_check": {"count": int(res[0]), "count_exact": ref["count"],
                            "sum_rel_diff": abs(res[1] - float(ref["sum"])) / max(float(ref["sum"]), 1.0)}}
    del a, b
    # the fastest single-column reference path, for scale
    m = 1 << 27
    x = (np.random.default_rng(1).random(m) * 1000.0)
    code = np.array([0x00001070, 0x000301C5, 0x000015D0, 0x0005002A, 0x00041120, 0x00021040, 0xFFA00052,
                     0x00000001], np.uint32)
    fn = o.R.vr_jit_compile(o.ptr(code), code.size, None)
    if fn:
        ts = []
        got = 0.0
        for i in range(5):
            t0 = time.perf_counter()
            got = o.R.vr_jit_entry(fn, o.ptr(x), m, m, 500.0)
            ts.append(time.perf_counter() - t0)
        o.R.vr_jit_release(fn)
        jm = statistics.median(ts[2:])
        want = float(x[x > 500.0].sum())
        out["jit_single_column"] = {
            "pct_of_8tbs": round(100.0 * m * 8 / jm / 1e9 / HBM_PEAK_GBS, 4), "gb_per_s": round(m * 8 / jm / 1e9, 3),
            "rows_per_s": round(m / jm, 1), "cores": 1,
            "sample": "reference JIT fused filter+sum (oracle/_ref vr_jit_entry = voxel_jit_run loop), 2^27 f64 "
                      "rows, x > 500, one column (8 B/row), median of 3 after 2 warmups, one core",
            "result_check": {"jit_sum": got, "numpy_sum": want, "rel_diff": abs(got - want) / abs(want)}}
    return out


def _traffic(rows_log2: int):
    """HBM bytes per launch of the headline kernel measured by rocprofv3
    (FETCH_SIZE + WRITE_SIZE, each scaled by the same counter on a copy of
    known size), as recorded in profiles/scan_f64_compound_traffic.json for
    this workload; None when no such measurement is committed."""
    path = os.path.join(ROOT, "profiles", "scan_f64_compound_traffic.json")
    try:
        d = json.load(open(path))
    except (OSError, ValueError):
        return None
    if d.get("rows_log2") != rows_log2:
        return None
    return d.get("traffic_bytes_per_launch")


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
    wall, kernel_ms = time_scan(desc, args.steps, args.warmup, stream, settle_s=args.settle, marks=True)
    res = scan.result(0, stream)
    algo_bytes = 2 * n * es
    ms_per_step = wall * 1e3 / args.steps
    gbs_wall = algo_bytes / (wall / args.steps) / 1e9
    gbs_kernel = algo_bytes / (kernel_ms * 1e-3) / 1e9

    if args.calibrate_copy:
        src = column("f64", 1)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del dst

    breakdown = {}
    if not args.no_breakdown:
        for dt_name in ("f64", "f32", "i64", "i32"):
            for sel in (0.01, 0.50, 0.99):
                d, e = desc_for(dt_name, sel)
                _, kms = time_scan(d, max(5, args.steps // 2), 2, stream, settle_s=0.2, settle_chunks=20)
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
                     "frac": round(gbs_kernel / HBM_PEAK_GBS, 4), "traffic": _traffic(args.rows_log2),
                     "kernel": "k_scan<double,2,...> + k_scan_final (vx_scan.hip), one op", "algorithmic_bytes_per_launch": algo_bytes,
                     "avg_kernel_ms": round(kernel_ms, 4)},
        "breakdown": breakdown,
        "result_check": {"count": res["count"], "sum": res["sum"], "min": res["min"], "max": res["max"]},
    }
    del cols
    torch.cuda.empty_cache()
    if not args.no_sort:
        out["sort"] = bench_sort(args.sort_log2)
    if not args.no_join:
        out["join"] = bench_join()
    if not args.no_groupby:
        out["groupby"] = bench_groupby()
    if not args.no_window:
        out["window"] = bench_window()
    if not args.no_expr:
        out["tier0_expression"] = bench_expr()
    if not args.no_nulls:
        out["nulls"] = bench_nulls(args.rows_log2)
    if not args.no_encodings:
        out["encodings"] = bench_encodings(args.rows_log2)
    if not args.no_sql:
        out["sql"] = bench_sql(args.rows_log2)
    if not args.no_serve:
        out["serve"] = bench_serve()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
        if not args.no_window and out["cpu_baseline"] is not None:
            out["cpu_baseline"]["window"] = cpu_window_baseline()
        if out["cpu_baseline"] is not None:
            out["cpu_baseline"]["paths"] = cpu_reference_paths()
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""GPU tests of the batched query kernel (vx_batch.hip) and the
continuous-batching server (vx_serve.cpp) through the C ABI.

Bars:
* every query of a batch equals the single-query kernel (voxel_gpu_scan):
  count, integer sum, min and max identical; float sums within 2 ulp of
  the single-query result (both are compensated sums of the same rows; only
  the split into partials differs);
* a query's result is bit-identical alone and inside any mix (fixed
  tile_rows), including on repeated graph launches;
* the autotuner's configurations all agree (checked inside the call);
* the server returns, for every query, the same bits as the batch engine
  run on that query alone, and the one-launch-per-query mode returns the
  single-query kernel's results.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TORCH_DT = {0: "float64", 1: "float32", 2: "int64", 3: "int32"}


def _ulp(x: float, dt: int) -> float:
    if dt == 1:
        return float(np.spacing(np.float32(abs(x)))) if x else float(np.finfo(np.float32).tiny)
    return float(np.spacing(abs(x))) if x else float(np.finfo(np.float64).tiny)


def same_as_single(res, ref, dt):
    assert res["count"] == ref["count"]
    assert res["min"] == ref["min"] or (res["min"] != res["min"] and ref["min"] != ref["min"])
    assert res["max"] == ref["max"] or (res["max"] != res["max"] and ref["max"] != ref["max"])
    if dt in (0, 1):
        assert abs(res["sum"] - ref["sum"]) <= 2 * _ulp(ref["sum"], dt), (res["sum"], ref["sum"])
    else:
        assert res["sum"] == ref["sum"]


def make_columns(rng, lengths, dtypes):
    import torch
    cols = []
    for n, dt in zip(lengths, dtypes):
        if dt in (0, 1):
            x = torch.from_numpy((rng.random(n) * 1000).astype(np.float64 if dt == 0 else np.float32))
        else:
            x = torch.from_numpy(rng.integers(-2**31, 2**31, n).astype(np.int64 if dt == 2 else np.int32))
        cols.append(x.cuda())
    return cols


_KEEP = []                                                   # second columns stay alive for the module


def make_queries(rng, cols, two_frac=0.3):
    from voxel_amd import scan
    qs = []
    for c in cols:
        dt = {"torch.float64": 0, "torch.float32": 1, "torch.int64": 2, "torch.int32": 3}[str(c.dtype)]
        op = int(rng.integers(0, 7))
        if dt in (0, 1):
            t = float(rng.random() * 1000)
        else:
            t = int(rng.integers(-2**31, 2**31))
        if rng.random() < two_frac and c.numel() > 0:
            other = c.flip(0).contiguous()
            _KEEP.append(other)
            qs.append(scan.make_desc(c, op, t, other, int(rng.integers(0, 6)), t))
        else:
            qs.append(scan.make_desc(c, op, t))
    return qs


def single(q):
    import ctypes as C
    from voxel_amd._lib import Agg, lib, check
    from voxel_amd.scan import _result
    r = Agg()
    if lib.voxel_gpu_scan(C.byref(q), None, C.byref(r)) != 0:
        check("voxel_gpu_scan")
    return _result(q.dtype, r)


def test_batch_matches_single_query_kernel(gpu):
    from voxel_amd.serve import Batch
    rng = np.random.default_rng(5)
    lengths = [0, 1, 3, 7, 64, 1000, 4095, 4096, 65535, 65537, 1 << 20, 3_000_001, 250_000]
    dtypes = [int(rng.integers(0, 4)) for _ in lengths]
    cols = make_columns(rng, lengths, dtypes)
    # misaligned starts (scalar path inside tiles)
    cols += [cols[-2][1:], cols[-3][3:]]
    qs = make_queries(rng, cols)
    b = Batch(64)
    res = b.run(qs)
    for q, r in zip(qs, res):
        same_as_single(r, single(q), q.dtype)


def test_batch_alone_equals_mix_bitwise_and_graph_replays(gpu):
    import torch
    from voxel_amd.serve import Batch
    rng = np.random.default_rng(6)
    lengths = [int(x) for x in np.clip(np.exp(rng.normal(np.log(2**16), 2, 200)), 1000, 1 << 22)]
    dtypes = [int(rng.integers(0, 4)) for _ in lengths]
    cols = make_columns(rng, lengths, dtypes)
    qs = make_queries(rng, cols)
    b = Batch(256)
    b.config(tile_rows=32768)
    mix, _ = b.run_raw(qs)
    for rep in range(3):                                   # graph launched again, eager work between
        torch.cuda.synchronize()
        _ = (cols[0].double() * 2).sum().item()
        again, _ = b.run_raw(qs)
        for x, y in zip(mix, again):
            assert bytes(x) == bytes(y)
    for i in range(0, len(qs), 17):
        alone, _ = b.run_raw([qs[i]])
        assert bytes(alone[0]) == bytes(mix[i]), i
    sub = [qs[i] for i in range(5, len(qs), 3)]
    part, _ = b.run_raw(sub)
    for j, i in enumerate(range(5, len(qs), 3)):
        assert bytes(part[j]) == bytes(mix[i])


def test_batch_autotune_and_errors(gpu):
    from voxel_amd.serve import Batch
    rng = np.random.default_rng(7)
    lengths = [int(x) for x in np.clip(np.exp(rng.normal(np.log(2**18), 1.5, 64)), 1000, 1 << 22)]
    cols = make_columns(rng, lengths, [int(rng.integers(0, 4)) for _ in lengths])
    qs = make_queries(rng, cols)
    b = Batch(64)
    info = b.autotune(qs, reps=1)
    assert info["configs_tried"] == 45 and info["best_ms"] > 0
    assert info["unroll"] in (2, 4, 8) and info["tile_rows"] % 4096 == 0
    res = b.run(qs)
    for q, r in zip(qs, res):
        same_as_single(r, single(q), q.dtype)
    with pytest.raises(RuntimeError):
        b.config(unroll=3)
    with pytest.raises(RuntimeError):
        b.config(tile_rows=5000)
    with pytest.raises(RuntimeError):
        Batch(2).run(qs[:3])                               # more queries than the engine holds


def test_server_continuous_batching_results_and_latency(gpu):
    from voxel_amd.serve import Batch, Server
    rng = np.random.default_rng(8)
    n = 512
    ncol = 16
    lengths = [int(x) for x in np.clip(np.exp(rng.normal(np.log(2**18), 2, ncol)), 1000, 1 << 23)]
    dtypes = [0 if i % 2 == 0 else 2 for i in range(ncol)]
    cols = make_columns(rng, lengths, dtypes)
    pick = rng.integers(0, ncol, n)
    qs = make_queries(rng, [cols[i] for i in pick], two_frac=0.0)
    arrivals = np.cumsum(rng.exponential(50e-6, n))       # 20k queries/s
    srv = Server(max_batch=128, memory_budget=1 << 16, inflight=2)   # small budget: admission waits
    srv.config(tile_rows=65536)
    out, lat, st = srv.replay_raw(qs, arrivals, mode=0)
    assert st["queries"] == n and st["batches"] >= 1 and 0 < st["p50_s"] <= st["p99_s"] <= st["max_s"]
    assert np.all(lat > 0)
    b = Batch(128)
    b.config(tile_rows=65536)
    for i in range(0, n, 7):                               # same bits as that query alone
        alone, _ = b.run_raw([qs[i]])
        assert bytes(alone[0]) == bytes(out[i]), i
    out1, lat1, st1 = srv.replay(qs[:64], arrivals[:64], mode=1)
    for q, r in zip(qs[:64], out1):
        ref = single(q)
        assert r == ref or (r["count"] == ref["count"] and r["sum"] == ref["sum"])
    print("serve: batched p50 %.1f us p99 %.1f us, %d batches (mean %.1f), waits %d; unbatched p50 %.1f us" % (
        st["p50_s"] * 1e6, st["p99_s"] * 1e6, st["batches"], st["mean_batch"], st["admission_waits"],
        st1["p50_s"] * 1e6))

"""GPU tests of the batched query kernel (vx_batch.hip) and the
continuous-batching server (vx_serve.cpp) through the C ABI.

Bars:
* every query of a batch equals the single-query kernel (voxel_gpu_scan):
  count, integer sum, min and max identical; float sums within 2 ulp of
  the single-query result (both are compensated sums of the same rows; only
  the split into partials differs);
* a query's result is bit-identical alone, inside any mix and under every
  configuration (tile rows, unroll, workgroups per CU), including on
  repeated graph launches (fixed 8192-row reduction units);
* the autotuner's configurations all agree bit for bit (checked inside
  the call);
* batch results against the oracle directly on a sample of queries: count,
  integer sums, min and max exact; float sums within 2 ulp (f64) / 4 ulp
  (f32) of the exact sum of the selected rows;
* the server (live submit / poll / wait from several threads, and replays)
  returns, for every query, the same bits as the batch engine run on that
  query alone; the one-launch-per-query modes return the single-query
  kernel's bits.
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
    from voxel_amd import scan
    for c in cols[5:9]:                                  # both predicates on one column (b == a)
        lo, hi = (250.0, 750.0) if c.dtype.is_floating_point else (-2**30, 2**30)
        qs.append(scan.make_desc(c, "gt", lo, c, "lt", hi))
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
    for tile_rows, unroll, bpc in ((16384, 2, 1), (262144, 8, 4), (65536, 4, 2)):
        b.config(unroll=unroll, tile_rows=tile_rows, blocks_per_cu=bpc)
        other, _ = b.run_raw(qs)
        for x, y in zip(mix, other):
            assert bytes(x) == bytes(y), (tile_rows, unroll, bpc)
        alone, _ = b.run_raw([qs[3]])                      # one query: smaller tiles, same bits
        assert bytes(alone[0]) == bytes(mix[3])


def test_batch_results_against_the_oracle(gpu):
    """A sample of batch queries against the oracle's exact relational
    result on the host copy of the same rows."""
    import math
    from oracle import oracle as o
    from voxel_amd.serve import Batch
    from voxel_amd.scan import _result
    rng = np.random.default_rng(9)
    lengths = [int(x) for x in np.clip(np.exp(rng.normal(np.log(2**17), 1.5, 48)), 1000, 1 << 21)]
    dtypes = [int(rng.integers(0, 4)) for _ in lengths]
    cols = make_columns(rng, lengths, dtypes)
    qs = make_queries(rng, cols, two_frac=0.4)
    b = Batch(64)
    raw, _ = b.run_raw(qs)
    for q, r in zip(qs, raw):
        got = _result(q.dtype, r)
        import ctypes as C
        a = _host(q.a, q.n, q.dtype)
        bb = _host(q.b, q.n, q.dtype) if q.b else None
        ta, tb = _thr(q.t_a, q.dtype), _thr(q.t_b, q.dtype)
        ref = o.relational(a, q.op_a, ta, bb, q.op_b if bb is not None else 6, tb)
        assert got["count"] == ref["count"]
        if ref["count"] == 0:
            continue
        if q.dtype in (0, 1):
            ex = float(ref["sum"])
            ulps = 2 if q.dtype == 0 else 4
            assert abs(got["sum"] - ex) <= ulps * _ulp(ex, q.dtype) + (0 if q.dtype == 0 else 1e-30), (got["sum"], ex)
        else:
            assert got["sum"] == ref["sum"]
        if ref["min"] is not None:
            assert got["min"] == ref["min"] and got["max"] == ref["max"]


def _host(ptr, n, dt):
    import ctypes as C
    from voxel_amd._lib import lib
    npdt = {0: np.float64, 1: np.float32, 2: np.int64, 3: np.int32}[dt]
    out = np.zeros(n, npdt)
    if n:
        assert lib.voxel_gpu_memcpy_d2h(out.ctypes.data, C.c_void_p(ptr), out.nbytes) == 0
    return out


def _thr(bits, dt):
    if dt == 0:
        return np.array([bits], np.uint64).view(np.float64)[0]
    if dt == 1:
        return np.array([bits & 0xFFFFFFFF], np.uint32).view(np.float32)[0]
    if dt == 2:
        return np.array([bits], np.uint64).view(np.int64)[0]
    return np.array([bits & 0xFFFFFFFF], np.uint32).view(np.int32)[0]


def test_batch_autotune_and_errors(gpu):
    from voxel_amd.serve import Batch
    rng = np.random.default_rng(7)
    lengths = [int(x) for x in np.clip(np.exp(rng.normal(np.log(2**18), 1.5, 64)), 1000, 1 << 22)]
    cols = make_columns(rng, lengths, [int(rng.integers(0, 4)) for _ in lengths])
    qs = make_queries(rng, cols)
    b = Batch(64)
    info = b.autotune(qs, reps=1)
    assert info["configs_tried"] == 45 and info["best_ms"] > 0
    assert info["unroll"] in (2, 4, 8) and info["tile_rows"] % 16384 == 0
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
    out2, lat2, st2 = srv.replay_raw(qs[:200], arrivals[:200], mode=2)   # queued without waiting
    import ctypes as C
    from voxel_amd._lib import Agg, lib
    for q, r in zip(qs[:200], out2):
        ref = Agg()
        assert lib.voxel_gpu_scan(C.byref(q), None, C.byref(ref)) == 0
        assert bytes(r) == bytes(ref)
    assert st2["queries"] == 200 and np.all(lat2 > 0)
    print("serve: batched p50 %.1f us p99 %.1f us, %d batches (mean %.1f), waits %d; unbatched p50 %.1f us" % (
        st["p50_s"] * 1e6, st["p99_s"] * 1e6, st["batches"], st["mean_batch"], st["admission_waits"],
        st1["p50_s"] * 1e6))


def test_live_server_submit_poll_wait_from_threads(gpu):
    """start(); two client threads submit queries as they go and wait for
    their tickets while a third polls; every result equals the same query
    run alone in a batch; stop() reports every query; submit after stop is
    refused."""
    import threading
    import time
    from voxel_amd.serve import Batch, Server
    rng = np.random.default_rng(10)
    ncol = 12
    lengths = [int(x) for x in np.clip(np.exp(rng.normal(np.log(2**17), 2, ncol)), 1000, 1 << 22)]
    cols = make_columns(rng, lengths, [0 if i % 2 == 0 else 2 for i in range(ncol)])
    qs = make_queries(rng, [cols[i] for i in rng.integers(0, ncol, 300)], two_frac=0.2)
    srv = Server(max_batch=64, memory_budget=1 << 20, inflight=2)
    srv.start()
    got = {}
    lock = threading.Lock()
    errors = []

    def client(idx):
        try:
            mine = []
            for i in idx:
                mine.append((i, srv.submit(qs[i])))
                time.sleep(float(rng.exponential(30e-6)))
            for i, t in mine:
                r, lat = srv.wait_raw(t)
                assert lat > 0
                with lock:
                    got[i] = bytes(r)
        except Exception as e:                              # noqa: BLE001
            errors.append(e)

    th = [threading.Thread(target=client, args=(range(k, 300, 2),)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    extra = srv.submit(qs[0])
    while srv.poll_raw(extra) is None:
        time.sleep(1e-4)
    st = srv.stop()
    assert st["queries"] == 301 and st["batches"] >= 1 and st["p50_s"] > 0
    with pytest.raises(RuntimeError, match="not running"):
        srv.submit(qs[0])
    b = Batch(8)
    for i in range(0, 300, 11):
        alone, _ = b.run_raw([qs[i]])
        assert bytes(alone[0]) == got[i], i
    from oracle import oracle as o
    from voxel_amd._lib import Agg
    from voxel_amd.scan import _result
    for i in range(0, 300, 29):                                # served results against the oracle
        q = qs[i]
        r = _result(q.dtype, Agg.from_buffer_copy(got[i]))
        a = _host(q.a, q.n, q.dtype)
        bb = _host(q.b, q.n, q.dtype) if q.b and q.op_b != 6 else None
        ref = o.relational(a, q.op_a, _thr(q.t_a, q.dtype), bb, q.op_b if bb is not None else 6, _thr(q.t_b, q.dtype))
        assert r["count"] == ref["count"], i
        if q.dtype in (0, 1) and ref["count"]:
            ex = float(ref["sum"])
            assert abs(r["sum"] - ex) <= (2 if q.dtype == 0 else 4) * _ulp(ex, q.dtype) + 1e-30, i
        elif ref["count"]:
            assert r["sum"] == ref["sum"], i
    print("live server: %d batches (mean %.2f), p50 %.1f us p99 %.1f us" % (
        st["batches"], st["mean_batch"], st["p50_s"] * 1e6, st["p99_s"] * 1e6))

"""GPU parity of bitmap filters, ColumnFilter algebra, compaction (vx_filter.hip)
and the hash join (vx_join.hip) through the C ABI.  Bar: bit-exact (bitmap
words, counts, indices, values, pairs, wrapping payload sums)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import oracle as o

pytestmark = pytest.mark.gpu

DT = {np.dtype(np.float64): 0, np.dtype(np.float32): 1, np.dtype(np.int64): 2, np.dtype(np.int32): 3}


def tbits(dt, v):
    return int(np.array([v], dt).view(np.uint64 if np.dtype(dt).itemsize == 8 else np.uint32)[0])


def dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def u64(n):
    import torch
    return torch.zeros(max(n, 1), dtype=torch.int64, device="cuda")


def host(t, n, dt):
    return t[:n].cpu().numpy().view(dt)


def oracle_bitmap(x, op, t, t2=0):
    nw = (x.size + 63) // 64
    w = np.zeros(max(nw, 1), np.uint64)
    fn = {np.float64: o.O.vo_bitmap_f64, np.float32: o.O.vo_bitmap_f32, np.int64: o.O.vo_bitmap_i64,
          np.int32: o.O.vo_bitmap_i32}[x.dtype.type]
    pc = fn(o.ptr(x), x.size, op, t, t2, o.ptr(w))
    return w[:nw], pc


@pytest.mark.parametrize("dt", [np.float64, np.float32, np.int64, np.int32])
@pytest.mark.parametrize("n", [1, 63, 64, 65, 1000, 130_001])
def test_bitmap_filter_all_ops(gpu, golden, dt, n):
    from voxel_amd._lib import lib
    rng = np.random.default_rng(n)
    if np.dtype(dt).kind == "f":
        x = np.round(rng.standard_normal(n) * 10).astype(dt)
        x[::13] = np.nan
    else:
        x = rng.integers(-20, 20, n).astype(dt)
    t = dt(3)
    d = dev(x)
    nw = (n + 63) // 64
    for op in range(8):
        words = u64(nw)
        pc = np.zeros(1, np.uint64)
        t2 = dt(9)
        rc = lib.voxel_gpu_filter_bitmap(DT[np.dtype(dt)], d.data_ptr(), n, op, tbits(dt, t), tbits(dt, t2),
                                         words.data_ptr(), pc.ctypes.data, None)
        assert rc == 0
        if op == 6:                      # no predicate: every row
            w_ref = np.zeros(nw, np.uint64)
            bits = np.zeros(nw * 64, np.uint8)
            bits[:n] = 1
            w_ref = np.packbits(bits, bitorder="little").view(np.uint64)
            pc_ref = n
        else:
            w_ref, pc_ref = oracle_bitmap(x, 6 if op == 7 else op, t, t2)
        assert int(pc[0]) == pc_ref, op
        assert np.array_equal(host(words, nw, np.uint64), w_ref), op


def test_bitmap_golden_and_compound(gpu, golden):
    from voxel_amd._lib import lib
    g = golden["bitmap_gt4"]
    x = np.array(g["data"], np.float64)
    words, pc = u64(1), np.zeros(1, np.uint64)
    dx = dev(x)
    lib.voxel_gpu_filter_bitmap(0, dx.data_ptr(), x.size, 4, tbits(np.float64, 4.0), 0, words.data_ptr(),
                                pc.ctypes.data, None)
    assert int(words[0].item()) == g["word"] and int(pc[0]) == g["pass"]
    rng = np.random.default_rng(1)
    n = 200_003
    a, b = rng.random(n) * 1000, rng.random(n) * 1000
    nw = (n + 63) // 64
    words = u64(nw)
    da, db = dev(a), dev(b)                # keep both alive for the call
    lib.voxel_gpu_filter_bitmap2(0, da.data_ptr(), db.data_ptr(), n, 4, tbits(np.float64, 292.8932188), 2,
                                 tbits(np.float64, 707.1067812), words.data_ptr(), pc.ctypes.data, None)
    wa, _ = oracle_bitmap(a, 4, 292.8932188)
    wb, _ = oracle_bitmap(b, 2, 707.1067812)
    cnt = o.O.vo_bitmap_combine(o.ptr(wa), o.ptr(wb), n, 0)
    assert int(pc[0]) == cnt and np.array_equal(host(words, nw, np.uint64), wa)


def test_bitmap_combine_and_selection(gpu):
    from voxel_amd._lib import lib
    rng = np.random.default_rng(2)
    n = 100_037
    nw = (n + 63) // 64
    x = rng.integers(0, 100, n).astype(np.int32)
    for combine in range(3):
        wa, _ = oracle_bitmap(x, 2, 60)
        wb, _ = oracle_bitmap(x, 5, 30)
        da, db = dev(wa.view(np.int64)), dev(wb.view(np.int64))
        cnt = np.zeros(1, np.uint64)
        assert lib.voxel_gpu_bitmap_combine(da.data_ptr(), db.data_ptr(), n, combine, cnt.ctypes.data, None) == 0
        ref_cnt = o.O.vo_bitmap_combine(o.ptr(wa), o.ptr(wb), n, combine)
        assert int(cnt[0]) == ref_cnt and np.array_equal(host(da, nw, np.uint64), wa)
        import torch
        idx = torch.zeros(n, dtype=torch.int32, device="cuda")
        k = np.zeros(1, np.uint64)
        assert lib.voxel_gpu_select_from_bitmap(da.data_ptr(), n, idx.data_ptr(), k.ctypes.data, None) == 0
        ref_idx = np.zeros(n, np.uint32)
        kr = o.O.vo_select_from_bitmap(o.ptr(wa), n, o.ptr(ref_idx))
        assert int(k[0]) == kr and np.array_equal(host(idx, kr, np.uint32), ref_idx[:kr])


@pytest.mark.parametrize("dt", [np.float64, np.float32, np.int64, np.int32])
@pytest.mark.parametrize("n", [0, 1, 4095, 4096, 4097, 1_000_003])
def test_compaction_values_and_indices(gpu, dt, n):
    import torch
    from voxel_amd._lib import lib
    rng = np.random.default_rng(n + 3)
    x = (rng.random(n) * 1000).astype(dt) if np.dtype(dt).kind == "f" else rng.integers(0, 1000, n).astype(dt)
    d = dev(x) if n else torch.zeros(1, dtype={8: torch.int64, 4: torch.int32}[np.dtype(dt).itemsize], device="cuda")
    vals = torch.zeros(max(n, 1) * np.dtype(dt).itemsize, dtype=torch.uint8, device="cuda")
    idx = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
    k = np.zeros(1, np.uint64)
    rc = lib.voxel_gpu_compact(DT[np.dtype(dt)], d.data_ptr(), n, 4, tbits(dt, 500), vals.data_ptr(),
                               idx.data_ptr(), k.ctypes.data, None)
    assert rc == 0
    ref = np.zeros(max(n, 1), dt)
    fn = {np.float64: o.O.vo_compact_f64, np.float32: o.O.vo_compact_f32, np.int64: o.O.vo_compact_i64,
          np.int32: o.O.vo_compact_i32}[np.dtype(dt).type]
    kr = fn(o.ptr(x), n, 4, dt(500), o.ptr(ref))
    assert int(k[0]) == kr
    got = vals.cpu().numpy().view(dt)[:kr]
    assert np.array_equal(got, ref[:kr])
    assert np.array_equal(host(idx, kr, np.uint32), np.nonzero(x > dt(500))[0].astype(np.uint32))
    if o.R is not None and dt in (np.float64, np.float32) and n:
        rr = np.zeros(n, dt)
        fr = o.R.vr_scalar_filter_gt_f64 if dt == np.float64 else o.R.vr_scalar_filter_gt_f32
        assert fr(o.ptr(rr), o.ptr(x), dt(500), n) == kr and np.array_equal(rr[:kr], ref[:kr])


def gpu_join(build, payload, probe, pairs=False):
    import torch
    from voxel_amd._lib import lib
    bd, pd = dev(build.astype(np.int64)), dev(payload.astype(np.int64)) if payload is not None else None
    pr = dev(probe.astype(np.int64)) if probe.size else torch.zeros(1, dtype=torch.int64, device="cuda")
    j = lib.voxel_gpu_join_build(bd.data_ptr() if build.size else None, pd.data_ptr() if pd is not None else None,
                                 build.size, None)
    assert j, lib.voxel_last_error()
    m = np.zeros(1, np.uint64)
    s = np.zeros(1, np.int64)
    assert lib.voxel_gpu_join_probe_sum(j, pr.data_ptr(), probe.size, m.ctypes.data, s.ctypes.data, None) == 0
    out = {"matches": int(m[0]), "sum": int(s[0])}
    if pairs:
        n = max(probe.size, 1)
        lo = torch.zeros(n, dtype=torch.int32, device="cuda")
        ro = torch.zeros(n, dtype=torch.int32, device="cuda")
        k = np.zeros(1, np.uint64)
        assert lib.voxel_gpu_join_probe_pairs(j, pr.data_ptr(), probe.size, lo.data_ptr(), ro.data_ptr(),
                                              k.ctypes.data, None) == 0
        kk = int(k[0])
        out.update(left=host(lo, kk, np.uint32).tolist(), right=host(ro, kk, np.uint32).tolist(), pairs=kk)
    lib.voxel_gpu_join_destroy(j)
    return out


def test_join_golden(gpu, golden):
    for name in ("join_small", "join_dups"):
        g = golden[name]
        r = gpu_join(np.array(g["build"]), np.array(g["payload"]), np.array(g["probe"]), pairs=True)
        assert r["matches"] == r["pairs"] == g["matches"]
        assert r["left"] == g["left"] and r["right"] == g["right"]
        assert r["sum"] == g["payload_sum"]


@pytest.mark.parametrize("nb,np_", [(0, 10), (1, 1), (1000, 5000), (50_000, 300_001)])
def test_join_random_vs_oracle(gpu, nb, np_):
    rng = np.random.default_rng(nb + np_)
    build = rng.integers(-5000, 5000, nb).astype(np.int64)       # duplicates on purpose
    if nb:
        build[rng.integers(0, nb, 3)] = np.iinfo(np.int64).min      # the empty-slot sentinel as a real key
    payload = rng.integers(-2**62, 2**62, nb).astype(np.int64)
    probe = rng.integers(-6000, 6000, np_).astype(np.int64)
    probe[::97] = np.iinfo(np.int64).min
    r = gpu_join(build, payload, probe, pairs=True)
    lo, ro = np.zeros(max(np_, 1), np.uint32), np.zeros(max(np_, 1), np.uint32)
    s = np.zeros(1, np.int64)
    m = o.O.vo_hashjoin_i64(o.ptr(build), nb, o.ptr(payload), o.ptr(probe), np_, o.ptr(lo), o.ptr(ro), o.ptr(s))
    assert r["matches"] == m == r["pairs"] and r["sum"] == int(s[0])
    assert r["left"] == lo[:m].tolist() and r["right"] == ro[:m].tolist()


def join_mode(build, payload, probe, mode, partition_probes=1):
    import ctypes as C
    import torch
    from voxel_amd._lib import JoinStats, lib
    bd, pd, pr = dev(build), dev(payload), dev(probe)
    st = JoinStats()
    j = lib.voxel_gpu_join_build_ex(bd.data_ptr(), pd.data_ptr(), build.size, mode, C.byref(st), None)
    assert j, lib.voxel_last_error()
    assert lib.voxel_gpu_join_set_probe_partitioning(j, partition_probes) == 0
    m, s = np.zeros(1, np.uint64), np.zeros(1, np.int64)
    assert lib.voxel_gpu_join_probe_sum(j, pr.data_ptr(), probe.size, m.ctypes.data, s.ctypes.data, None) == 0
    lo = torch.zeros(probe.size, dtype=torch.int32, device="cuda")
    ro = torch.zeros(probe.size, dtype=torch.int32, device="cuda")
    k = np.zeros(1, np.uint64)
    assert lib.voxel_gpu_join_probe_pairs(j, pr.data_ptr(), probe.size, lo.data_ptr(), ro.data_ptr(),
                                          k.ctypes.data, None) == 0
    lib.voxel_gpu_join_destroy(j)
    kk = int(k[0])
    return {"matches": int(m[0]), "sum": int(s[0]), "pairs": kk, "left": host(lo, kk, np.uint32),
            "right": host(ro, kk, np.uint32), "stats": st.as_dict()}


def oracle_join(build, payload, probe):
    lo, ro = np.zeros(max(probe.size, 1), np.uint32), np.zeros(max(probe.size, 1), np.uint32)
    s = np.zeros(1, np.int64)
    m = o.O.vo_hashjoin_i64(o.ptr(build), build.size, o.ptr(payload), o.ptr(probe), probe.size, o.ptr(lo),
                            o.ptr(ro), o.ptr(s))
    return m, int(s[0]), lo[:m], ro[:m]


@pytest.mark.parametrize("mode", [1, 2, 3, 4])
def test_join_partitioned_layouts_vs_oracle(gpu, mode):
    """One table, hash sub-tables, learned and equal-width range sub-tables:
    match count, payload sum and (probe, build) pairs equal the oracle's, with
    the probes partitioned by sub-table (2^20 + 3 rows) and in input order."""
    rng = np.random.default_rng(40 + mode)
    nb, np_ = 300_000, (1 << 20) + 3
    build = rng.integers(-2**40, 2**40, nb).astype(np.int64)
    build[rng.integers(0, nb, 20_000)] = build[rng.integers(0, nb, 20_000)]    # duplicate keys
    build[7] = np.iinfo(np.int64).min                                        # the empty-slot sentinel
    payload = rng.integers(-2**62, 2**62, nb).astype(np.int64)
    probe = np.where(rng.random(np_) < 0.7, build[rng.integers(0, nb, np_)], rng.integers(-2**41, 2**41, np_))
    probe[::1001] = np.iinfo(np.int64).min
    m, s, lo, ro = oracle_join(build, payload, probe)
    for pp in (1, 0):
        r = join_mode(build, payload, probe, mode, pp)
        assert r["stats"]["mode"] == mode and r["stats"]["fell_back"] == 0
        assert r["stats"]["partitions"] == (1 if mode == 1 else 8)
        assert r["matches"] == m == r["pairs"] and r["sum"] == s
        assert np.array_equal(r["left"], lo) and np.array_equal(r["right"], ro)


def test_join_learned_partition_gain_and_equal_width_fallback(gpu):
    """Build keys floor(exp(20 + 2Z)): equal-width key ranges put nearly all
    keys into the first sub-table, which overflows, and the table falls back to
    one table; the learned CDF splits them evenly (fullest sub-table at most
    1.5x the mean) and needs no fallback.  Both give the oracle's result."""
    rng = np.random.default_rng(77)
    nb, np_ = 300_000, 400_000
    build = np.unique(np.floor(np.exp(20 + 2 * rng.standard_normal(nb))).astype(np.int64))
    payload = rng.integers(-2**62, 2**62, build.size).astype(np.int64)
    probe = np.where(rng.random(np_) < 0.9, build[rng.integers(0, build.size, np_)], -rng.integers(1, 2**40, np_))
    m, s, lo, ro = oracle_join(build, payload, probe)
    eq = join_mode(build, payload, probe, 4)
    assert eq["stats"]["fell_back"] == 4 and eq["stats"]["mode"] == 1
    le = join_mode(build, payload, probe, 3)
    st = le["stats"]
    assert st["fell_back"] == 0 and st["mode"] == 3 and st["partitions"] == 8
    assert st["max_partition_keys"] <= 1.5 * build.size / st["partitions"], st
    assert st["model"]["train_steps"] > 0 and st["model"]["log_grid"] == 1
    for r in (eq, le):
        assert r["matches"] == m and r["sum"] == s
        assert np.array_equal(r["left"], lo) and np.array_equal(r["right"], ro)


def test_join_config_shape_properties(gpu):
    """Build 2^24 distinct keys (bit 63 clear) with i64 payloads, probe 2^28:
    a probe set built from build keys (all match) and their bit-63-set copies
    (never match): matches = hits exactly, payload sum = sum of gathered
    payloads (computed by torch on the same device data)."""
    import torch
    from voxel_amd._lib import lib
    nb, npr = 1 << 24, 1 << 28
    keys = torch.empty(nb, dtype=torch.int64, device="cuda")
    pay = torch.empty(nb, dtype=torch.int64, device="cuda")
    assert lib.voxel_gpu_fill_uniform(2, keys.data_ptr(), nb, 11, 0.0, 2.0**62, None) == 0
    assert lib.voxel_gpu_fill_uniform(2, pay.data_ptr(), nb, 12, -2.0**40, 2.0**40, None) == 0
    torch.cuda.synchronize()
    keys = torch.unique(keys)                        # distinct keys
    nb = keys.numel()
    pay = pay[:nb].contiguous()
    g = torch.Generator(device="cuda").manual_seed(3)
    pick = torch.randint(0, nb, (npr,), device="cuda", generator=g)
    miss = torch.rand(npr, device="cuda", generator=g) < 0.1
    probe = keys[pick] | (miss.to(torch.int64) << 63)
    j = lib.voxel_gpu_join_build(keys.data_ptr(), pay.data_ptr(), nb, None)
    assert j
    m, s = np.zeros(1, np.uint64), np.zeros(1, np.int64)
    assert lib.voxel_gpu_join_probe_sum(j, probe.data_ptr(), npr, m.ctypes.data, s.ctypes.data, None) == 0
    lib.voxel_gpu_join_destroy(j)
    hit = ~miss
    assert int(m[0]) == int(hit.sum().item())
    assert int(s[0]) == int(pay[pick[hit]].sum().item())   # |sum| < 2^40 * 2^28: no wrap

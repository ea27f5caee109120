"""GPU parity of the columnar encodings (vx_encoding.hip) through the C ABI.

Encoded inputs come from the oracle's builders (vo_*_build, pinned against
the reference's Build methods).  Bars: DecodeBatch outputs bit-identical with
the oracle's DecodeBatch (vo_*_decode, pinned against the reference), same
row count, outputs past the written rows untouched; fused scans over the
encoded bytes: COUNT, MIN, MAX and integer SUM exact against the decoded
column (NaN-free inputs), f64 SUM within 2 ulp of the exact sum of the
selected decoded rows.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as o

pytestmark = pytest.mark.gpu

OPS = {0: np.equal, 1: np.not_equal, 2: np.less, 3: np.less_equal, 4: np.greater, 5: np.greater_equal}


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def expect(x, op, t):
    sel = OPS[op](x, t) if op < 6 else np.ones(x.size, bool)
    return x[sel]


def check_scan(got, v, is_float):
    assert got["count"] == v.size
    if v.size == 0:
        return
    if is_float:
        ex = o.exact_sum(v.astype(np.float64))
        assert abs(got["sum"] - float(ex)) <= 2 * o.ulp(float(ex)) + 1e-300, (got["sum"], float(ex))
        assert got["min"] == v.min() and got["max"] == v.max()
    else:
        assert got["sum"] == int(np.sum(v.astype(np.int64), dtype=np.int64)) or \
            (got["sum"] - int(v.astype(object).sum())) % 2**64 == 0
        assert got["min"] == int(v.min()) and got["max"] == int(v.max())


def decode_cases(n):
    return ((0, n), (n // 3, n // 2 + 5), (max(n - 2, 0), 9), (n + 4, 3))


@pytest.mark.parametrize("kind", ["f64", "i64"])
def test_dictionary_and_rle_decode_and_scan(gpu, kind):
    import torch
    from voxel_amd import encoding as enc
    rng = np.random.default_rng(61 if kind == "f64" else 62)
    dt = np.float64 if kind == "f64" else np.int64
    for trial in range(12):
        n = int(rng.choice([1, 7, 1000, 4097, 300_001]))
        if kind == "f64":
            x = np.round(rng.standard_normal(n) * 50, int(rng.integers(0, 3)))
        else:
            x = rng.integers(-300, 300, n).astype(np.int64) * (2**33 if trial % 2 else 1)
        if trial % 3 == 0:
            x = np.repeat(x[: (n + 15) // 16], 16)[:n].copy()
        d, ix = np.zeros(n, dt), np.zeros(n, np.uint32)
        g = getattr(o.O, "vo_dict_build_" + kind)(o.ptr(x), n, o.ptr(d), o.ptr(ix))
        v, ln = np.zeros(n, dt), np.zeros(n, np.uint32)
        r = getattr(o.O, "vo_rle_build_" + kind)(o.ptr(x), n, o.ptr(v), o.ptr(ln))
        cols = {"dict": enc.dictionary(dev(d[:g]), dev(ix)), "rle": enc.rle(dev(v[:r]), dev(ln[:r]))}
        for name, col in cols.items():
            for off, cnt in decode_cases(n):
                want = np.full(cnt + 3, 7, dt)
                if name == "dict":
                    getattr(o.O, "vo_dict_decode_" + kind)(o.ptr(d), o.ptr(ix), n, off, cnt, o.ptr(want))
                    wn = max(0, min(cnt, n - off))
                else:
                    getattr(o.O, "vo_rle_decode_" + kind)(o.ptr(v), o.ptr(ln), r, off, cnt, o.ptr(want))
                    wn = cnt
                out = torch.full((cnt + 3,), 7, dtype=torch.float64 if kind == "f64" else torch.int64, device="cuda")
                got_n = col.decode(off, cnt, out)
                assert got_n == wn, (name, off, cnt)
                assert np.array_equal(out.cpu().numpy().view(np.uint64), want.view(np.uint64)), (name, trial, off)
            for op in range(7):
                t = x[int(rng.integers(0, n))] if op != 6 else 0
                got = col.scan(op, t)
                check_scan(got, expect(x, op, t), kind == "f64")


@pytest.mark.parametrize("nd", [10, 5000])                      # dictionary in LDS / in HBM
def test_dictionary_code_out_of_range_is_refused(gpu, nd):
    """An index at or past the dictionary's size (a caller-built column; the
    reference's Build never writes one) makes DecodeBatch and the fused scan
    fail with an error; the kernels read no entry past the dictionary (such
    a code is taken as 0 and flagged).  Valid columns still decode after."""
    import torch
    from voxel_amd import encoding as enc
    rng = np.random.default_rng(64)
    d = np.sort(rng.standard_normal(nd))
    for bad in (nd, 2**32 - 1):
        ix = rng.integers(0, nd, 4099).astype(np.uint32)
        ix[int(rng.integers(0, ix.size))] = bad
        col = enc.dictionary(dev(d), dev(ix.view(np.int32)))
        out = torch.zeros(ix.size, dtype=torch.float64, device="cuda")
        with pytest.raises(RuntimeError, match="not below the dictionary"):
            col.decode(0, ix.size, out)
        with pytest.raises(RuntimeError, match="not below the dictionary"):
            col.scan("gt", 0.0)
    ix = rng.integers(0, nd, 4099).astype(np.uint32)
    col = enc.dictionary(dev(d), dev(ix.view(np.int32)))
    out = torch.zeros(ix.size, dtype=torch.float64, device="cuda")
    assert col.decode(0, ix.size, out) == ix.size
    assert np.array_equal(out.cpu().numpy(), d[ix])
    check_scan(col.scan("gt", 0.0), expect(d[ix], 4, 0.0), True)


def test_delta_decode_and_scan(gpu):
    import torch
    from voxel_amd import encoding as enc
    rng = np.random.default_rng(63)
    for trial in range(10):
        n = int(rng.choice([1, 100, 4095, 4096, 4097, 1_000_003]))
        for suf, dt, tdt, lo, hi in (("i64", np.int64, torch.int64, -2**62, 2**62),
                                     ("i32", np.int32, torch.int32, -2**30, 2**30)):
            x = np.cumsum(rng.integers(-1000, 1000, n)).astype(dt) if trial % 2 else rng.integers(lo, hi, n).astype(dt)
            dl = np.zeros(n, dt)
            getattr(o.O, "vo_delta_build_" + suf)(o.ptr(x), n, o.ptr(dl))
            base = int(rng.integers(-5, 5))
            col = enc.delta(dev(dl), base)
            for off, cnt in decode_cases(n):
                want = np.full(cnt + 3, 9, dt)
                getattr(o.O, "vo_delta_decode_" + suf)(base, o.ptr(dl), n, off, cnt, o.ptr(want))
                out = torch.full((cnt + 3,), 9, dtype=tdt, device="cuda")
                assert col.decode(off, cnt, out) == max(0, min(cnt, n - off))
                assert np.array_equal(out.cpu().numpy(), want), (suf, trial, off, cnt)
            full = np.zeros(n, dt)
            getattr(o.O, "vo_delta_decode_" + suf)(base, o.ptr(dl), n, 0, n, o.ptr(full))
            t = full[int(rng.integers(0, n))]
            for op in (2, 5, 6):
                check_scan(col.scan(op, int(t)), expect(full, op, t), False)


def test_bitpacked_and_for_decode_and_scan(gpu):
    import torch
    from voxel_amd import encoding as enc
    rng = np.random.default_rng(64)
    for trial in range(12):
        n = int(rng.choice([1, 63, 64, 65, 10_000, 1 << 20]))
        bits = int(rng.integers(0, 33))
        u = rng.integers(0, 2**bits, n, dtype=np.uint64).astype(np.uint32) if bits else np.zeros(n, np.uint32)
        w = np.zeros(max(n, 1), np.uint64)
        bw = np.zeros(1, np.uint32)
        k = o.O.vo_bitpack_build(o.ptr(u), n, o.ptr(w), o.ptr(bw))
        col = enc.bitpacked(dev(w[:k]), int(bw[0]), n)
        for off, cnt in decode_cases(n):
            want = np.full(cnt + 3, 5, np.uint32)
            o.O.vo_bitpack_decode(o.ptr(w), k, int(bw[0]), n, off, cnt, o.ptr(want))
            out = torch.full((cnt + 3,), 5, dtype=torch.int32, device="cuda")
            assert col.decode(off, cnt, out) == max(0, min(cnt, n - off))
            assert np.array_equal(out.cpu().numpy().view(np.uint32), want), (trial, off, cnt)
        ui = u.astype(np.int64)
        for op in range(7):
            t = int(u[int(rng.integers(0, n))]) if op != 6 else 0
            check_scan(col.scan(op, t), expect(ui, op, t), False)
        f = rng.standard_normal(n) * 1000 + float(rng.choice([0.0, 5e6]))
        e = np.zeros(max(n, 1), np.int32)
        bs = np.zeros(2)
        o.O.vo_for_build(o.ptr(f), n, o.ptr(bs[0:]), o.ptr(bs[1:]), o.ptr(e))
        col = enc.frame_of_reference(dev(e[:n]), float(bs[0]), float(bs[1]))
        full = np.zeros(n)
        o.O.vo_for_decode(bs[0], bs[1], o.ptr(e), n, 0, n, o.ptr(full))
        for off, cnt in decode_cases(n):
            want = np.full(cnt + 3, 2.5)
            o.O.vo_for_decode(bs[0], bs[1], o.ptr(e), n, off, cnt, o.ptr(want))
            out = torch.full((cnt + 3,), 2.5, dtype=torch.float64, device="cuda")
            assert col.decode(off, cnt, out) == max(0, min(cnt, n - off))
            assert np.array_equal(out.cpu().numpy().view(np.uint64), want.view(np.uint64)), (trial, off)
        for op in range(7):
            t = full[int(rng.integers(0, n))] if op != 6 else 0.0
            check_scan(col.scan(op, float(t)), expect(full, op, t), True)


def test_encoded_columns_at_scale(gpu):
    """2^26 rows: dictionary (1000 distinct f64), FOR and 12-bit packed
    columns scanned from their encoded bytes agree with the decoded column."""
    import torch
    from voxel_amd import encoding as enc
    rng = np.random.default_rng(65)
    n = 1 << 26
    d = np.sort(rng.standard_normal(1000) * 100)
    ix = rng.integers(0, 1000, n).astype(np.uint32)
    col = enc.dictionary(dev(d), dev(ix))
    x = d[ix]
    check_scan(col.scan(4, float(d[500])), expect(x, 4, d[500]), True)
    u = rng.integers(0, 4096, n).astype(np.uint32)
    w = np.zeros(n // 64 * 12 + 1, np.uint64)
    bw = np.zeros(1, np.uint32)
    k = o.O.vo_bitpack_build(o.ptr(u), n, o.ptr(w), o.ptr(bw))
    assert bw[0] == 12
    col = enc.bitpacked(dev(w[:k]), 12, n)
    check_scan(col.scan(2, 1000), expect(u.astype(np.int64), 2, 1000), False)
    del col
    torch.cuda.empty_cache()


def test_bitpacked_scan_every_width(gpu):
    """The packed scan at every bit width 1..32: row counts around whole
    32-row groups, the words 16-byte aligned and 8 bytes past alignment, and
    thresholds at the ends of the u32 range (empty, full and wrapped
    complement ranges of the predicate)."""
    import torch
    from voxel_amd import encoding as enc
    rng = np.random.default_rng(66)
    for bits in range(1, 33):
        n = int(rng.choice([31, 32, 33, 4095, 70_001]))
        top = (1 << bits) - 1
        u = rng.integers(0, top + 1, n, dtype=np.uint64).astype(np.uint32)
        u[: min(n, 3)] = [top, 0, top][: min(n, 3)]
        w = np.zeros(max(n, 1), np.uint64)
        bw = np.zeros(1, np.uint32)
        k = o.O.vo_bitpack_build(o.ptr(u), n, o.ptr(w), o.ptr(bw))
        assert bw[0] == bits
        width = bits
        buf = torch.zeros(k + 2, dtype=torch.int64, device="cuda")
        ui = u.astype(np.int64)
        for shift in (0, 1):                                  # aligned words, then 8 bytes past
            buf[shift:shift + k] = dev(w[:k].view(np.int64))
            col = enc.bitpacked(buf[shift:shift + k], width, n)
            cases = [(op, int(u[int(rng.integers(0, n))])) for op in range(6)]
            cases += [(2, 0), (5, 0), (4, 0xFFFFFFFF), (3, 0xFFFFFFFF), (1, 0), (1, 0xFFFFFFFF), (0, top), (6, 0)]
            for op, t in cases:
                check_scan(col.scan(op, t), expect(ui, op, t), False)


def test_delta_scan_across_workgroup_spans(gpu):
    """Delta scans long enough that each workgroup walks several 4096-row
    tiles, with the deltas 16-byte aligned and 8 (4) bytes past alignment."""
    import torch
    from voxel_amd import encoding as enc
    rng = np.random.default_rng(67)
    n = (1 << 25) + 12_345
    for suf, dt, tdt in (("i64", np.int64, torch.int64), ("i32", np.int32, torch.int32)):
        x = np.cumsum(rng.integers(-1000, 1000, n)).astype(dt)
        dl = np.zeros(n, dt)
        getattr(o.O, "vo_delta_build_" + suf)(o.ptr(x), n, o.ptr(dl))
        base = int(rng.integers(-5, 5))
        full = np.zeros(n, dt)
        getattr(o.O, "vo_delta_decode_" + suf)(base, o.ptr(dl), n, 0, n, o.ptr(full))
        buf = torch.zeros(n + 1, dtype=tdt, device="cuda")
        for shift in (0, 1):
            buf[shift:shift + n] = dev(dl)
            col = enc.delta(buf[shift:shift + n], base)
            for op in (1, 2, 4, 6):
                t = full[int(rng.integers(0, n))] if op != 6 else 0
                check_scan(col.scan(op, int(t)), expect(full, op, t), False)
        del buf, col
        torch.cuda.empty_cache()


@pytest.mark.parametrize("dt", [np.float64, np.float32, np.int64, np.int32])
def test_dictionary_scan_by_code_counts(gpu, dt):
    """Dictionaries of at most 4096 values are scanned by per-code row counts
    (the aggregate taken once per code as count x value): every predicate,
    dictionaries of 1, 3 and 4096 values, infinite and huge f64 values, i64
    sums that wrap; COUNT, MIN, MAX exact, integer SUM exact modulo 2^64,
    f64 SUM within 2 ulp and f32 SUM within 4 f32 ulp of the exact sum."""
    from voxel_amd import encoding as enc
    rng = np.random.default_rng(68)
    for nd in (1, 3, 4096):
        n = (1 << 22) + int(rng.integers(0, 1000))
        if dt == np.float64:
            d = np.sort(rng.standard_normal(nd) * 1e3)
            if nd >= 3:
                d[0], d[-1] = -np.inf, np.inf
                d[nd // 2] = 3e300
            d = np.sort(d)
        elif dt == np.float32:
            d = np.sort(rng.standard_normal(nd).astype(np.float32) * np.float32(100))
        elif dt == np.int64:
            d = np.sort(rng.integers(-2**62, 2**62, nd)).astype(np.int64)
        else:
            d = np.sort(rng.integers(-2**30, 2**30, nd)).astype(np.int32)
        ix = rng.integers(0, nd, n).astype(np.uint32)
        x = d[ix]
        col = enc.dictionary(dev(d), dev(ix.view(np.int32)))
        for op in range(7):
            t = d[int(rng.integers(0, nd))] if op != 6 else dt(0)
            if dt == np.float64 and not np.isfinite(t):
                t = dt(0.5)
            got = col.scan(op, t.item() if hasattr(t, "item") else t)
            v = expect(x, op, t)
            assert got["count"] == v.size, (nd, op)
            if v.size == 0:
                continue
            assert got["min"] == v.min() and got["max"] == v.max(), (nd, op)
            if dt == np.float64:
                if np.all(np.isfinite(v)):
                    ex = float(o.exact_sum(v))
                    assert abs(got["sum"] - ex) <= 2 * o.ulp(ex) + 1e-300, (nd, op, got["sum"], ex)
                else:
                    assert math.isnan(got["sum"]) or math.isinf(got["sum"])
            elif dt == np.float32:
                ex = float(o.exact_sum(v.astype(np.float64)))
                assert abs(got["sum"] - ex) <= 4 * float(np.spacing(np.float32(abs(ex)))) + 1e-30, (nd, op)
            else:
                assert (got["sum"] - int(v.astype(object).sum())) % 2**64 == 0, (nd, op)

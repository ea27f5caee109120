"""GPU parity of the validity-bitmap paths (vx_nulls.hip, k_scan_valid in
vx_scan.hip) through the C ABI.

Bars: NullableSegment<T> ValidCount / NullCountVal / MinValid / MaxValid and
CompactValid bit-identical with the oracle (vo_nullable_*, pinned against the
reference's NullableSegment built from its sources); SumValid within 2 ulp of
the exact sum of the valid rows (f64), exact (i64, wrapping); scans over
nullable columns: COUNT, MIN, MAX exact and SUM within 2 ulp of the exact sum
of the rows selected by (predicate AND validity), the selection computed on
the host from the same words.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as o

pytestmark = pytest.mark.gpu


def random_validity(rng, n, kind):
    nw = max((n + 63) // 64, 1)
    if kind == 0:
        return np.full(nw, np.uint64(0xFFFFFFFFFFFFFFFF), np.uint64)
    if kind == 1:
        return np.zeros(nw, np.uint64)
    w = rng.integers(0, 2**63, nw, dtype=np.int64).astype(np.uint64) * np.uint64(2) + \
        rng.integers(0, 2, nw).astype(np.uint64)
    if kind == 3:
        w[rng.integers(0, nw, max(nw // 3, 1))] = 0
    return w


def valid_rows(w, n):
    return np.unpackbits(w.view(np.uint8), bitorder="little")[:n].astype(bool)


def close_f64(got, exact, ulps=2.0):
    e = float(exact)
    return abs(got - e) <= ulps * o.ulp(e) + 1e-300


@pytest.mark.parametrize("is_f64", [True, False])
def test_nullable_reduce_and_compact_vs_oracle(gpu, is_f64):
    import torch
    from voxel_amd import scan as vs
    rng = np.random.default_rng(41 if is_f64 else 42)
    for trial in range(30):
        n = int(rng.choice([1, 63, 64, 65, 1000, 4097, 100_003, 1 << 20]))
        if is_f64:
            x = rng.standard_normal(n) * 100
            if trial % 6 == 5:
                x[rng.integers(0, n, 5)] = np.nan
            if trial % 12 == 11:
                x[0] = np.nan                                 # a NaN first valid value (kind 0: all valid)
        else:
            x = rng.integers(-2**40, 2**40, n).astype(np.int64)
        kind = trial % 5
        w = random_validity(rng, n, 0 if trial % 12 == 11 else kind)
        dt = np.float64 if is_f64 else np.int64
        res, cnt, comp = np.zeros(3, dt), np.zeros(3, np.uint64), np.zeros(n, dt)
        (o.O.vo_nullable_f64 if is_f64 else o.O.vo_nullable_i64)(o.ptr(x), n, o.ptr(w), o.ptr(res), o.ptr(cnt),
                                                                  o.ptr(comp))
        xd = torch.from_numpy(x).cuda()
        wd = torch.from_numpy(w.view(np.int64)).cuda()
        got = vs.nullable_reduce(xd, wd)
        assert got["valid_count"] == int(cnt[0]) and got["null_count"] == int(cnt[1]), trial
        sel = valid_rows(w, n)
        if is_f64:
            for k, key in ((1, "min"), (2, "max")):
                want = float(res[k])
                if math.isnan(want):
                    assert math.isnan(got[key]), (trial, key)
                else:
                    assert got[key] == want, (trial, key, got[key], want)   # == : a -0/+0 tie may differ in sign
            if cnt[0] == 0:
                assert got["sum"] == 0.0
            elif np.isnan(x[sel]).any():
                assert math.isnan(got["sum"])
            else:
                assert close_f64(got["sum"], o.exact_sum(x[sel])), trial
        else:
            assert got["sum"] == int(res[0]) and got["min"] == int(res[1]) and got["max"] == int(res[2]), trial
        out = torch.zeros_like(xd)
        m = vs.nullable_compact(xd, wd, out)
        assert m == int(cnt[2])
        assert np.array_equal(out[:m].cpu().numpy().view(np.uint64), comp[:m].view(np.uint64)), trial


def test_nullable_four_byte_types_and_no_bitmap(gpu):
    import torch
    from voxel_amd import scan as vs
    rng = np.random.default_rng(43)
    n = 70_001
    w = random_validity(rng, n, 2)
    sel = valid_rows(w, n)
    wd = torch.from_numpy(w.view(np.int64)).cuda()
    x32 = rng.integers(-2**31, 2**31, n).astype(np.int32)
    got = vs.nullable_reduce(torch.from_numpy(x32).cuda(), wd)
    want_sum = int(np.int64(x32[sel].astype(np.int64).sum()).astype(np.int32))  # SumValid accumulates in i32
    assert got["valid_count"] == int(sel.sum())
    assert got["sum"] == want_sum and got["min"] == int(x32[sel].min()) and got["max"] == int(x32[sel].max())
    f32 = (rng.standard_normal(n) * 10).astype(np.float32)
    got = vs.nullable_reduce(torch.from_numpy(f32).cuda(), wd)
    assert got["min"] == float(f32[sel].min()) and got["max"] == float(f32[sel].max())
    assert abs(got["sum"] - float(o.exact_sum(f32[sel].astype(np.float64)))) <= 1e-5 * np.abs(f32[sel]).sum()
    out = torch.zeros(n, dtype=torch.float32, device="cuda")
    m = vs.nullable_compact(torch.from_numpy(f32).cuda(), wd, out)
    assert m == int(sel.sum()) and np.array_equal(out[:m].cpu().numpy(), f32[sel])
    # no bitmap: every row valid
    x = rng.standard_normal(5000)
    got = vs.nullable_reduce(torch.from_numpy(x).cuda(), None)
    assert got["valid_count"] == 5000 and got["null_count"] == 0 and got["min"] == x.min()


@pytest.mark.parametrize("dt", ["f64", "i64", "f32", "i32"])
def test_scan_over_nullable_columns(gpu, dt):
    import torch
    from voxel_amd import scan as vs
    rng = np.random.default_rng({"f64": 51, "i64": 52, "f32": 53, "i32": 54}[dt])
    npdt = {"f64": np.float64, "i64": np.int64, "f32": np.float32, "i32": np.int32}[dt]
    ops = {0: np.equal, 1: np.not_equal, 2: np.less, 3: np.less_equal, 4: np.greater, 5: np.greater_equal}
    for trial in range(16):
        n = int(rng.choice([5, 64, 1000, 123_457, 1 << 21]))
        if dt in ("f64", "f32"):
            a = (rng.standard_normal(n) * 100).astype(npdt)
            b = (rng.standard_normal(n) * 100).astype(npdt)
            if trial % 4 == 1:
                a[rng.integers(0, n, 3)] = np.nan
            ta, tb = float(np.round(rng.standard_normal() * 50)), float(np.round(rng.standard_normal() * 50))
        else:
            hi = 2**40 if dt == "i64" else 2**20
            a = rng.integers(-hi, hi, n).astype(npdt)
            b = rng.integers(-hi, hi, n).astype(npdt)
            ta, tb = int(rng.integers(-hi, hi)), int(rng.integers(-hi, hi))
        if trial % 5 == 0:
            a[rng.integers(0, n, 4)] = npdt(ta)
        wa = random_validity(rng, n, trial % 5)
        wb = random_validity(rng, n, (trial + 2) % 5)
        op_a, op_b = int(rng.integers(0, 7)), int(rng.integers(0, 6))
        two = trial % 2 == 1
        with np.errstate(invalid="ignore"):
            sel = valid_rows(wa, n) & (ops[op_a](a, npdt(ta)) if op_a < 6 else True)
            if two:
                sel = sel & ops[op_b](b, npdt(tb)) & valid_rows(wb, n)
        ad, bd = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
        wad, wbd = torch.from_numpy(wa.view(np.int64)).cuda(), torch.from_numpy(wb.view(np.int64)).cuda()
        if two:
            r = vs.scan(ad, op_a, ta, bd, op_b, tb, valid_a=wad, valid_b=wbd)
        else:
            r = vs.scan(ad, op_a, ta, valid_a=wad)
        v = a[sel]
        assert r["count"] == int(sel.sum()), (dt, trial)
        if dt in ("i64", "i32"):
            assert r["sum"] == int(v.astype(np.int64).sum()), (dt, trial)
            if v.size:
                assert r["min"] == int(v.min()) and r["max"] == int(v.max())
        else:
            vn = v[~np.isnan(v)]
            if np.isnan(v).any():
                assert math.isnan(r["sum"])
            elif v.size:
                ex = o.exact_sum(v.astype(np.float64))
                tol = 2 * o.ulp(float(ex)) if dt == "f64" else 1e-5 * float(np.abs(v.astype(np.float64)).sum()) + 1e-30
                assert abs(r["sum"] - float(ex)) <= tol, (dt, trial)
            if vn.size:
                assert r["min"] == float(vn.min()) and r["max"] == float(vn.max()), (dt, trial)

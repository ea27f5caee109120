"""GPU parity of the window family (vx_window.hip) through the C ABI and the
voxel_py EngineF64.window_* methods.

Bars: delta, sum, mean, variance, stddev and quantile outputs bit-identical
with the oracle (vo_window_f64, pinned against Engine<f64>'s static window
methods built from the reference sources), the same return value, outputs
past N untouched; time_window_sum within 2 ulp (+1e-300) of the exact bucket
sum (oracle long accumulator) and the same return value as the oracle
(vo_time_window_sum_f64, pinned against Engine<f64>::TimeWindowSum).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as o

pytestmark = pytest.mark.gpu

KINDS = ("delta", "sum", "mean", "stddev", "variance", "quantile")


def oracle_window(kind, x, out_len, w, q, fill):
    out = np.full(out_len, fill)
    n = o.O.vo_window_f64(KINDS.index(kind), o.ptr(x), x.size, o.ptr(out), out_len, w, q)
    return n, out


def gpu_window(e, kind, x, out, w, q):
    if kind == "delta":
        return e.window_delta(x, out)
    if kind == "quantile":
        return e.window_quantile(x, out, w, q)
    return getattr(e, "window_" + kind)(x, out, w)


@pytest.mark.parametrize("kind", KINDS)
def test_window_family_vs_oracle(gpu, kind):
    from voxel_amd import voxel_py as vx
    e = vx.EngineF64()
    rng = np.random.default_rng(11 + KINDS.index(kind))
    cases = []
    for w in (0, 1, 2, 3, 20, 100, 128, 129, 200, 255):
        for n in (1, 5, 300, 1023, 1024, 1025, 5000):
            cases.append((w, n))
    for trial, (w, n) in enumerate(cases):
        x = rng.standard_normal(n) * 10.0 ** int(rng.integers(-2, 5))
        if trial % 5 == 0:
            x = np.round(x)                                  # ties and zeros for the quantile
        q = float(rng.choice([0.0, 0.05, 0.5, 0.95, 1.0, -0.1, 1.5]))
        out_len = int(rng.choice([0, n // 2, n, n + 7]))
        fill = -12345.678
        want_n, want = oracle_window(kind, x, out_len, w, q, fill)
        out = np.full(out_len, fill)
        got_n = gpu_window(e, kind, x, out, w, q)
        assert got_n == want_n, (kind, w, n, q, out_len)
        if kind == "quantile" and trial % 5 == 0:
            # equal values: nth_element may return either zero of a -0/+0 tie
            assert np.array_equal(out, want), (kind, w, n, q)
        else:
            assert np.array_equal(out.view(np.uint64), want.view(np.uint64)), (kind, w, n, q)


def test_window_large_column(gpu):
    """2^24 rows, windows 20 and 255: sum, mean, stddev bit-identical."""
    from voxel_amd import voxel_py as vx
    e = vx.EngineF64()
    x = o.mt_uniform_f64(7, 1 << 24)
    for kind, w in (("sum", 20), ("mean", 255), ("stddev", 20)):
        want_n, want = oracle_window(kind, x, x.size, w, 0.0, 0.0)
        out = np.zeros(x.size)
        assert gpu_window(e, kind, x, out, w, 0.0) == want_n
        assert np.array_equal(out.view(np.uint64), want.view(np.uint64)), kind


def test_window_quantile_large_window_of_market_data(gpu):
    """voxel_market's VaR shape (window 250 -> first 128 rows of each window,
    quantile 0.05) on a random walk of 1e5 returns."""
    from voxel_amd import voxel_py as vx
    e = vx.EngineF64()
    r = np.diff(np.log(np.cumsum(np.random.default_rng(5).standard_normal(100_001)) + 1e4))
    want_n, want = oracle_window("quantile", r, r.size, 250, 0.05, 0.0)
    out = np.zeros(r.size)
    assert e.window_quantile(r, out, 250, 0.05) == want_n == r.size - 249
    assert np.array_equal(out.view(np.uint64), want.view(np.uint64))


def test_window_errors(gpu):
    from voxel_amd import voxel_py as vx
    e = vx.EngineF64()
    x = np.arange(50, dtype=np.float64)
    out = np.zeros(50)
    for bad in (256, -1, 2.0):
        with pytest.raises(TypeError):
            e.window_sum(x, out, bad)
    x[10] = np.nan
    with pytest.raises(RuntimeError, match="NaN"):
        e.window_quantile(x, out, 5, 0.5)
    # NaN outside every window that is computed: accepted
    assert e.window_quantile(x[11:], out, 5, 0.5) == 35
    # sums propagate NaN as the reference does
    want_n, want = oracle_window("sum", x, 50, 4, 0.0, 0.0)
    out = np.zeros(50)
    assert e.window_sum(x, out, 4) == want_n
    assert np.array_equal(np.isnan(out), np.isnan(want)) and np.array_equal(out[~np.isnan(out)], want[~np.isnan(want)])


def test_time_window_sum_vs_oracle(gpu):
    from voxel_amd import voxel_py as vx
    e = vx.EngineF64()
    rng = np.random.default_rng(21)
    for trial in range(24):
        n = int(rng.choice([0, 1, 17, 1000, 200_000]))
        ts = np.sort(rng.integers(-5000, 400_000, n)).astype(np.int64)
        if trial % 4 == 1:
            rng.shuffle(ts)
        v = rng.standard_normal(n) * 100
        if trial % 6 == 2 and n:
            v[rng.integers(0, n)] = np.inf
        start = int(rng.integers(-1000, 1000))
        dur = int(rng.choice([-3, 0, 1, 60, 3600]))
        out_len = int(rng.choice([0, 1, 5, 120, 4000]))
        out0 = rng.standard_normal(out_len)
        want = out0.copy()
        want_n = o.O.vo_time_window_sum_f64(o.ptr(ts), o.ptr(v), n, start, dur, o.ptr(want), out_len)
        got = out0.copy()
        got_n = e.time_window_sum(ts, v, start, dur, got)
        assert got_n == want_n, (trial, n, dur, out_len)
        if dur <= 0 or n == 0 or out_len == 0:
            assert np.array_equal(got, out0)
            continue
        d = ts - start
        b = np.where(d >= 0, d // dur, -((-d) // dur))           # C++ division truncates toward zero
        order = np.argsort(b, kind="stable")
        bs = b[order]
        for k in range(out_len):
            lo, hi = np.searchsorted(bs, k, "left"), np.searchsorted(bs, k, "right")
            rows = v[order[lo:hi]]
            if rows.size == 0:
                assert got[k] == out0[k]
                continue
            if not np.all(np.isfinite(rows)):
                assert got[k] == want[k] or (math.isnan(got[k]) and math.isnan(want[k])), (trial, k)
                continue
            ex = o.exact_sum(np.concatenate([[out0[k]], rows]))
            assert abs(got[k] - float(ex)) <= 2 * o.ulp(float(ex)) + 1e-300, (trial, k, got[k], float(ex))

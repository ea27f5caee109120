"""GPU runs of the SQL front end (voxel_amd/voxel_sql.py).

The reference plan (compat) must give voxel_query's results, whose words and
results are pinned against the reference elsewhere, including its
last-threshold-wins compound WHERE.  The lowered plan is checked against SQL's
meaning computed on the host: COUNT, MIN, MAX exact; SUM within 2 ulp of the
exact sum of the selected rows (oracle long accumulator); AVG = SUM / COUNT.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as o

pytestmark = pytest.mark.gpu


def close(got, v):
    ex = float(o.exact_sum(v))
    return abs(got - ex) <= 2 * o.ulp(ex) + 1e-300


def test_reference_plan_matches_voxel_query(gpu, x42):
    from voxel_amd import voxel_query as vq
    from voxel_amd import voxel_sql as vs
    x = x42
    assert vs.execute("SELECT SUM(price) FROM trades WHERE price > 500", {"trades": x}) == \
        vq.from_numpy(x).filter_gt(500).sum().run()
    # compound WHERE: each comparison is a filter_* call, the last one wins
    got = vs.execute("SELECT SUM(price) FROM trades WHERE price > 100 AND price < 500", {"trades": x})
    assert got == vq.from_numpy(x).filter_lt(500).sum().run()
    assert got == vs.execute("SELECT SUM(price) FROM trades WHERE price < 500", {"trades": x})
    # OR and NOT act as AND of their comparisons; <> is not compiled
    assert vs.execute("SELECT MAX(p) FROM t WHERE NOT (p >= 10 OR p <= 700)", x) == \
        vq.from_numpy(x).filter_le(700).max().run()
    assert vs.execute("SELECT COUNT(p) FROM t WHERE p <> 3", x) == vq.from_numpy(x).count().run()
    r = vs.execute("SELECT SUM(v) FROM t GROUP BY k", {"t": x[:1000], "k": (np.arange(1000) % 7).astype(np.float64)})
    assert r == {"groups": 7, "function": "sum"}


def test_lowered_plan_over_named_columns(gpu):
    from voxel_amd import voxel_sql as vs
    rng = np.random.default_rng(9)
    n = 1_000_003
    fare = np.round(rng.lognormal(2.5, 0.6, n), 2)
    dist = np.round(rng.exponential(3.0, n), 1)
    zone = rng.integers(1, 266, n).astype(np.float64)
    data = {"trips": {"fare": fare, "dist": dist, "zone": zone}}
    cases = [
        ("fare > 10 AND dist < 5", (fare > 10) & (dist < 5)),                 # fused two-column pass
        ("fare > 10 AND fare < 50", (fare > 10) & (fare < 50)),               # both on the aggregated column
        ("dist BETWEEN 1 AND 2", (dist >= 1) & (dist <= 2)),                  # predicate on another column only
        ("fare > 10 AND dist < 5 AND zone = 132", (fare > 10) & (dist < 5) & (zone == 132)),   # bitmap plan
        ("fare < 5 OR dist > 20", (fare < 5) | (dist > 20)),
        ("NOT (fare >= 8 AND zone <> 7)", ~((fare >= 8) & (zone != 7))),
        ("30 < fare", fare > 30),
        ("fare > 10 AND (dist < 5 OR zone = 7)", (fare > 10) & ((dist < 5) | (zone == 7))),   # fused + OR group words
        ("dist < 3 AND zone > 100 AND fare < 40 AND fare > 5",
         (dist < 3) & (zone > 100) & (fare < 40) & (fare > 5)),                  # fused + two comparisons' words
        ("NOT (zone >= 100) AND dist BETWEEN 1 AND 9", (zone < 100) & (dist >= 1) & (dist <= 9)),
    ]
    for where, sel in cases:
        v = fare[sel]
        s, c, a, mn, mx = vs.execute("SELECT SUM(fare), COUNT(*), AVG(fare), MIN(fare), MAX(fare) FROM trips WHERE "
                                     + where, data)
        assert c == int(sel.sum()), where
        assert close(s, v), where
        assert abs(a - float(o.exact_sum(v)) / v.size) <= 4 * o.ulp(a), where
        assert mn == v.min() and mx == v.max(), where
    assert vs.execute("SELECT COUNT(*) FROM trips WHERE fare > 1e9", data) == 0
    assert vs.execute("SELECT MIN(fare) FROM trips WHERE fare > 1e9", data) is None


def test_lowered_rows_order_limit_and_group_by(gpu):
    import torch
    from voxel_amd import voxel_sql as vs
    rng = np.random.default_rng(10)
    n = 200_001
    fare = np.round(rng.lognormal(2.5, 0.6, n), 2)
    zone = rng.integers(1, 40, n).astype(np.float64)
    data = {"trips": {"fare": torch.from_numpy(fare).cuda(), "zone": zone}}     # one column already in HBM
    got = vs.execute("SELECT fare FROM trips WHERE zone = 3 AND fare > 20 ORDER BY fare DESC LIMIT 25", data)
    want = np.sort(fare[(zone == 3) & (fare > 20)])[::-1][:25]
    assert np.array_equal(got, want)
    got = vs.execute("SELECT fare FROM trips WHERE zone < 3", data)
    assert np.array_equal(got, fare[zone < 3])                                  # row order kept
    r = vs.execute("SELECT zone, SUM(fare), COUNT(fare), MAX(fare) FROM trips WHERE fare < 30 GROUP BY zone", data)
    sel = fare < 30
    keys = np.unique(zone[sel]).astype(np.int64)
    assert np.array_equal(r["zone"], keys)
    for i, k in enumerate(keys):
        v = fare[sel & (zone == k)]
        assert r["count(fare)"][i] == v.size and r["max(fare)"][i] == v.max()
        assert close(r["sum(fare)"][i], v)
    # a one-column table with compat=False: the WHERE clause keeps SQL's meaning
    x = fare
    s = vs.execute("SELECT SUM(p) FROM t WHERE p > 10 AND p < 50", x, compat=False)
    assert close(s, x[(x > 10) & (x < 50)])


def test_lowered_plan_on_library_buffers(gpu):
    """Columns held in the library's own device arrays, NaN rows under NOT,
    negative and fractional GROUP BY keys (truncated toward zero), ORDER BY
    with LIMIT past the selected rows, an empty selection."""
    from voxel_amd import voxel_sql as vs
    from voxel_amd.devarray import DeviceArray
    rng = np.random.default_rng(11)
    n = 70_003
    v = rng.normal(0.0, 50.0, n)
    v[rng.integers(0, n, 300)] = np.nan
    k = np.round(rng.uniform(-6.0, 6.0, n), 1)
    data = {"t": {"v": DeviceArray.from_numpy(v), "k": k}}
    got = vs.execute("SELECT v FROM t WHERE NOT (v > -10 AND v < 10)", data)
    want = v[~((v > -10) & (v < 10))]                     # NaN rows fail both comparisons, so NOT keeps them
    assert np.array_equal(got, want, equal_nan=True)
    got = vs.execute("SELECT v FROM t WHERE v > 140 ORDER BY v ASC LIMIT 100000", data)
    assert np.array_equal(got, np.sort(v[v > 140]))
    assert vs.execute("SELECT v FROM t WHERE v > 1e9", data).size == 0
    r = vs.execute("SELECT k, COUNT(v), MIN(v) FROM t WHERE v > 0 GROUP BY k", data)
    sel = v > 0
    ik = np.trunc(k[sel]).astype(np.int64)
    keys = np.unique(ik)
    assert np.array_equal(r["k"], keys)
    for i, g in enumerate(keys):
        x = v[sel][ik == g]
        assert r["count(v)"][i] == x.size and r["min(v)"][i] == x.min()

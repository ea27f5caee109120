"""The SQL front end's parser and plan choice (voxel_amd/voxel_sql.py), no GPU."""
from __future__ import annotations

import pytest

from voxel_amd import voxel_sql as vs


def test_parse_select_forms():
    q = vs.parse("SELECT SUM(price), COUNT(*) FROM trades WHERE price > 500 AND qty <= -2.5;")
    assert q["targets"] == [("func", "sum", "price"), ("func", "count", "*")]
    assert q["table"] == ("name", "trades")
    assert q["where"] == ("and", [("cmp", ("col", "price"), ">", ("const", 500.0)),
                                  ("cmp", ("col", "qty"), "<=", ("const", -2.5))])
    q = vs.parse("select x from 'data.csv' where not (x = 1 or x != 2) order by x desc limit 10")
    assert q["table"] == ("name", "data.csv") and q["order_by"] == ("x", False) and q["limit"] == 10
    assert q["where"] == ("not", ("or", [("cmp", ("col", "x"), "=", ("const", 1.0)),
                                         ("cmp", ("col", "x"), "<>", ("const", 2.0))]))
    q = vs.parse("SELECT zone, AVG(fare) AS a FROM (SELECT fare FROM t) s WHERE fare BETWEEN 1 AND 2e3 GROUP BY zone")
    assert q["table"][0] == "subquery" and q["group_by"] == "zone"
    assert q["where"] == ("between", ("col", "fare"), ("const", 1.0), ("const", 2000.0))


def test_parse_errors():
    with pytest.raises(NotImplementedError):
        vs.parse("INSERT INTO t VALUES (1)")
    for bad in ("SELECT FROM t", "SELECT x FROM t WHERE x >", "SELECT x FROM t LIMIT", "SELECT x FROM t junk junk"):
        with pytest.raises(ValueError):
            vs.parse(bad)


def test_conjunctions_flatten_for_the_fused_plan():
    w = vs.parse("SELECT SUM(a) FROM t WHERE (a > 1 AND 5 > b) AND a BETWEEN 0 AND 9")["where"]
    assert vs._comparisons(w) == [("a", ">", 1.0), ("b", "<", 5.0), ("a", ">=", 0.0), ("a", "<=", 9.0)]
    assert vs._comparisons(vs.parse("SELECT SUM(a) FROM t WHERE a > 1 OR b < 2")["where"]) is None
    assert vs._comparisons(vs.parse("SELECT SUM(a) FROM t WHERE NOT a > 1")["where"]) is None

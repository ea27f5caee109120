"""SQL front end over the GPU backend (bindings/voxel_sql.py:1-347).

The reference parses with pglast (not available here); this module parses the
same SELECT subset itself: ``SELECT <targets> FROM <table | 'file.csv' |
(subquery)> [WHERE <condition>] [GROUP BY col] [ORDER BY col [ASC|DESC]]
[LIMIT n]``, comparisons ``col op const`` (op one of > >= < <= = <> !=),
``BETWEEN``, AND / OR / NOT and parentheses.

Two plans:

* ``compat=True`` (the default for a one-column table, as the reference
  takes): the reference's plan, step for step (voxel_sql.py:60-82, 208-241):
  the query goes through ``voxel_query`` (its words run on the GPU through
  the C ABI), so every WHERE comparison is applied with the reference's
  ``filter_*`` calls -- the last threshold and comparison win and OR / NOT
  act as AND (voxel_query.py:37-72, voxel_sql.py:242-244) -- and GROUP BY
  returns ``{"groups": n, "function": name}`` from the hash aggregator.
* ``compat=False`` (the default when the table is a dict of columns): the
  WHERE clause is lowered to a predicate plan over named columns with SQL's
  meaning.  A conjunction touching the aggregated column and at most one
  other column (at most two comparisons) runs as one fused filter +
  aggregate pass (voxel_gpu_scan, both predicates on the aggregated column
  when they share it).  A longer conjunction keeps that fused pass over the
  aggregated column and one other, its remaining conjuncts (comparisons,
  OR / NOT groups) becoming selection words the same pass reads
  (voxel_gpu_scan_valid); anything else builds per-comparison bitmaps on
  the GPU, combines them (AND / OR / AND-NOT) and aggregates through the
  validity-word scan.  Row queries (``SELECT col``)
  compact the selected rows, ORDER BY sorts them on the GPU, LIMIT keeps the
  first n; GROUP BY runs the GPU group-by on the selected rows.
"""
from __future__ import annotations

import csv as _csv
import ctypes as C
import os as _os
import re

import numpy as np

from . import voxel_query as _vq
from . import voxel_py as _vx

# ---------------------------------------------------------------------------
# Parser
# ---------------------------------------------------------------------------
_TOKEN = re.compile(r"\s*(?:(?P<num>\d+\.\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?|\d+(?:[eE][-+]?\d+)?)"
                    r"|(?P<str>'(?:[^']|'')*')|(?P<id>[A-Za-z_][A-Za-z_0-9.]*|\"[^\"]+\")"
                    r"|(?P<op><>|!=|>=|<=|[=<>(),*+-]))")
_KEYWORDS = {"select", "from", "where", "and", "or", "not", "group", "by", "order", "asc", "desc", "limit",
             "between", "as"}
_AGGS = {"sum", "count", "avg", "min", "max"}


def _tokens(sql: str):
    out, pos = [], 0
    sql = sql.strip().rstrip(";")
    while pos < len(sql):
        m = _TOKEN.match(sql, pos)
        if not m or m.end() == pos:
            raise ValueError("SQL syntax error near %r" % sql[pos:pos + 20])
        pos = m.end()
        if m.group("num") is not None:
            out.append(("num", float(m.group("num"))))
        elif m.group("str") is not None:
            out.append(("str", m.group("str")[1:-1].replace("''", "'")))
        elif m.group("id") is not None:
            v = m.group("id")
            if v.startswith('"'):
                out.append(("id", v[1:-1]))
            elif v.lower() in _KEYWORDS:
                out.append(("kw", v.lower()))
            else:
                out.append(("id", v))
        else:
            out.append(("op", m.group("op")))
    return out


class _Parser:
    def __init__(self, sql: str):
        self.t = _tokens(sql)
        self.i = 0

    def peek(self, kind=None, val=None, k=0):
        if self.i + k >= len(self.t):
            return False
        tk = self.t[self.i + k]
        return (kind is None or tk[0] == kind) and (val is None or tk[1] == val)

    def take(self, kind=None, val=None):
        if not self.peek(kind, val):
            got = self.t[self.i] if self.i < len(self.t) else ("end", None)
            raise ValueError("SQL syntax error: expected %s, got %r" % (val or kind, got[1]))
        tk = self.t[self.i]
        self.i += 1
        return tk[1]

    def select(self):
        self.take("kw", "select")
        targets = [self.target()]
        while self.peek("op", ","):
            self.take()
            targets.append(self.target())
        self.take("kw", "from")
        if self.peek("op", "("):
            self.take()
            table = ("subquery", self.select())
            self.take("op", ")")
        elif self.peek("str"):
            table = ("name", self.take())
        else:
            table = ("name", self.take("id"))
        if self.peek("id"):                                   # alias
            self.take()
        q = {"targets": targets, "table": table, "where": None, "group_by": None, "order_by": None, "limit": None}
        if self.peek("kw", "where"):
            self.take()
            q["where"] = self.expr()
        if self.peek("kw", "group"):
            self.take()
            self.take("kw", "by")
            q["group_by"] = self.take("id")
        if self.peek("kw", "order"):
            self.take()
            self.take("kw", "by")
            col = self.take("id")
            asc = True
            if self.peek("kw", "asc"):
                self.take()
            elif self.peek("kw", "desc"):
                self.take()
                asc = False
            q["order_by"] = (col, asc)
        if self.peek("kw", "limit"):
            self.take()
            q["limit"] = int(self.take("num"))
        return q

    def target(self):
        if self.peek("op", "*"):
            self.take()
            return ("star",)
        name = self.take("id")
        if self.peek("op", "("):
            self.take()
            arg = "*" if self.peek("op", "*") else None
            if arg:
                self.take()
            else:
                arg = self.take("id")
            self.take("op", ")")
            if self.peek("kw", "as"):
                self.take()
                self.take("id")
            return ("func", name.lower(), arg)
        if self.peek("kw", "as"):
            self.take()
            self.take("id")
        return ("col", name)

    def expr(self):
        e = self.conj()
        while self.peek("kw", "or"):
            self.take()
            e = ("or", [e, self.conj()])
        return e

    def conj(self):
        e = self.unary()
        while self.peek("kw", "and"):
            self.take()
            e = ("and", [e, self.unary()])
        return e

    def unary(self):
        if self.peek("kw", "not"):
            self.take()
            return ("not", self.unary())
        if self.peek("op", "("):
            self.take()
            e = self.expr()
            self.take("op", ")")
            return e
        left = self.operand()
        if self.peek("kw", "between"):
            self.take()
            lo = self.operand()
            self.take("kw", "and")
            hi = self.operand()
            return ("between", left, lo, hi)
        op = self.take("op")
        if op not in (">", ">=", "<", "<=", "=", "<>", "!="):
            raise ValueError("SQL syntax error: comparison operator expected, got %r" % op)
        return ("cmp", left, "<>" if op == "!=" else op, self.operand())

    def operand(self):
        neg = False
        while self.peek("op", "-") or self.peek("op", "+"):
            neg ^= self.take() == "-"
        if self.peek("num"):
            v = self.take()
            return ("const", -v if neg else v)
        if self.peek("str"):
            return ("const", self.take())
        return ("col", self.take("id"))


def parse(sql: str) -> dict:
    """SELECT statement -> query dict (targets, table, where, group_by,
    order_by, limit)."""
    p = _Parser(sql)
    if not p.peek("kw", "select"):
        first = p.t[0][1] if p.t else ""
        raise NotImplementedError("Statement type %s not supported" % str(first).upper())
    q = p.select()
    if p.i != len(p.t):
        raise ValueError("SQL syntax error near %r" % (p.t[p.i][1],))
    return q


# ---------------------------------------------------------------------------
# Reference plan (compat)
# ---------------------------------------------------------------------------
def load_csv(path, column=0):
    """First column (or ``column``) of a CSV file with a header row, as f64
    (voxel_sql.py:334-347)."""
    rows = []
    with open(path, "r") as f:
        reader = _csv.reader(f)
        next(reader, None)
        for row in reader:
            if row and column < len(row):
                try:
                    rows.append(float(row[column]))
                except ValueError:
                    pass
    return np.array(rows, dtype=np.float64)


class SqlCompiler:
    """Compiles and runs SELECT statements (voxel_sql.py:20-82)."""

    def __init__(self, data_sources=None, compat=None):
        self._data = data_sources or {}
        self._compat = compat

    def execute(self, sql: str):
        return self._run(parse(sql))

    # -- table resolution (voxel_sql.py:177-198) --
    def _table(self, table):
        kind, name = table
        if kind == "subquery":
            return self._run(name)
        if name in self._data:
            return self._data[name]
        path = name.strip("'\"")
        if _os.path.exists(path):
            return load_csv(path)
        raise ValueError("Cannot resolve table: %s" % name)

    def _run(self, q):
        table = self._table(q["table"])
        compat = self._compat
        if compat is None:
            compat = not isinstance(table, dict)
        if compat:
            if isinstance(table, dict):
                raise ValueError("the reference plan takes a one-column table")
            return self._run_reference(q, np.asarray(table).astype(np.float64))
        return _run_lowered(q, table, self._data)

    def _run_reference(self, q, table):
        targets = q["targets"]
        agg = targets[0][1] if targets and targets[0][0] == "func" else None
        if q["group_by"] is not None:
            return self._group_by_reference(q, table, agg)
        qq = _vq.from_numpy(table)
        if q["where"] is not None:
            self._where_reference(qq, q["where"])
        for t in targets:                                    # voxel_sql.py:231-251
            if t[0] != "func":
                continue
            if t[1] == "sum":
                qq.sum()
            elif t[1] == "count":
                qq.count()
            elif t[1] == "avg":
                qq.sum()
                qq.count()
            elif t[1] == "min":
                qq.min()
            elif t[1] == "max":
                qq.max()
        if q["order_by"] is not None:
            qq.sort(ascending=q["order_by"][1])
        if q["limit"]:
            qq.topk(int(q["limit"]), largest=True)
        return qq.run()

    def _where_reference(self, qq, e):
        # every comparison under AND / OR / NOT is applied as a filter, in
        # order (voxel_sql.py:208-244); only the ops the reference knows
        kind = e[0]
        if kind in ("and", "or"):
            for a in e[1]:
                self._where_reference(qq, a)
        elif kind == "not":
            self._where_reference(qq, e[1])
        elif kind == "cmp":
            _, left, op, right = e
            if right[0] != "const":
                right_val = None
            else:
                right_val = right[1]
            fn = {">": qq.filter_gt, ">=": qq.filter_ge, "<": qq.filter_lt, "<=": qq.filter_le,
                  "=": qq.filter_eq}.get(op)
            if fn is not None:
                fn(right_val)

    def _group_by_reference(self, q, table, agg):
        mask = np.ones(len(table), bool)
        if q["where"] is not None:
            mask = _where_numpy_reference(table, q["where"])
        filtered = table[mask]
        key = q["group_by"]
        if key in self._data:
            keys = np.array(np.asarray(self._data[key])[np.nonzero(mask)[0]], dtype=np.uint32)
        else:
            keys = np.arange(len(filtered), dtype=np.uint32)
        arena = _vx.Arena()
        h = _vx.HashAggregator(arena)
        h.init(max(128, int(len(keys) / 10) + 1))
        h.accumulate(keys, filtered)
        return {"groups": int(h.group_count()), "function": agg or "sum"}


def _where_numpy_reference(table, e):
    """voxel_sql.py:117-145: AND of every comparison, '=' within 1e-10."""
    kind = e[0]
    if kind in ("and", "or"):
        r = np.ones(len(table), bool)
        for a in e[1]:
            r &= _where_numpy_reference(table, a)
        return r
    if kind == "cmp":
        _, _, op, right = e
        v = right[1]
        if op == ">":
            return table > v
        if op == ">=":
            return table >= v
        if op == "<":
            return table < v
        if op == "<=":
            return table <= v
        if op == "=":
            return np.abs(table - v) < 1e-10
    return np.ones(len(table), bool)


# ---------------------------------------------------------------------------
# Lowered plan (SQL semantics over named columns)
# ---------------------------------------------------------------------------
_CMP = {"=": 0, "<>": 1, "<": 2, "<=": 3, ">": 4, ">=": 5}
_FLIP = {"<": ">", "<=": ">=", ">": "<", ">=": "<=", "=": "=", "<>": "<>"}


def _device_column(v):
    """A caller's HBM-resident f64 column (a DeviceArray, or any object with
    data_ptr() / numel() on a contiguous f64 device buffer, such as a torch
    CUDA tensor) wrapped in place; None for host data."""
    from .devarray import DeviceArray
    if isinstance(v, DeviceArray):
        if v.dtype != np.float64:
            raise TypeError("device columns must be f64")
        return v
    if hasattr(v, "data_ptr") and getattr(v, "is_cuda", False):
        if str(getattr(v, "dtype", "")) not in ("torch.float64", "float64") or not v.is_contiguous():
            raise TypeError("device columns must be contiguous f64")
        return DeviceArray.borrow(v.data_ptr(), v.numel(), np.float64, owner=v)
    return None


class _Cols:
    """Device copies of the table's columns (the library's allocator), made
    on first use; HBM-resident f64 columns are used in place."""

    def __init__(self, table):
        self.host = table if isinstance(table, dict) else {"": table}
        self.dev = {}

    def name(self, col):
        if col in self.host:
            return col
        if "" in self.host and len(self.host) == 1:
            return ""
        raise ValueError("unknown column %s" % col)

    def get(self, col):
        from .devarray import DeviceArray
        c = self.name(col)
        if c not in self.dev:
            v = self.host[c]
            d = _device_column(v)
            if d is None:
                d = DeviceArray.from_numpy(np.ascontiguousarray(np.asarray(v), dtype=np.float64))
            self.dev[c] = d
        return self.dev[c]

    def rows(self):
        v = next(iter(self.host.values()))
        return v.numel() if hasattr(v, "numel") else len(v)


def _comparisons(e):
    """The conjunction's comparisons (column, op, value), or None when the
    condition is not a plain conjunction."""
    kind = e[0]
    if kind == "and":
        out = []
        for a in e[1]:
            r = _comparisons(a)
            if r is None:
                return None
            out += r
        return out
    if kind == "cmp":
        _, left, op, right = e
        if left[0] == "col" and right[0] == "const":
            return [(left[1], op, float(right[1]))]
        if left[0] == "const" and right[0] == "col":
            return [(right[1], _FLIP[op], float(left[1]))]
        return None
    if kind == "between":
        _, c, lo, hi = e
        if c[0] == "col" and lo[0] == "const" and hi[0] == "const":
            return [(c[1], ">=", float(lo[1])), (c[1], "<=", float(hi[1]))]
    return None


def _conjuncts(e):
    """The items of a top-level AND (nested ANDs flattened), or [e]."""
    if e[0] != "and":
        return [e]
    out = []
    for x in e[1]:
        out += _conjuncts(x)
    return out


def _bits(v: float) -> int:
    return int(np.array([v], np.float64).view(np.uint64)[0])


def _bitmap(cols, e):
    """Selection words of condition e (DeviceArray of (n + 63) // 64 u64)."""
    from ._lib import check, lib
    from .devarray import DeviceArray
    n = cols.rows()
    nw = max((n + 63) // 64, 1)

    def alloc():
        # the bitmap kernel writes every word; with no rows the one word is zero
        return DeviceArray(nw, np.uint64) if n else DeviceArray.zeros(nw, np.uint64)
    w = alloc()
    cnt = C.c_uint64(0)
    kind = e[0]
    if kind in ("cmp", "between"):
        comps = _comparisons(e)
        if comps is None:
            raise ValueError("comparison between two columns or two constants is not supported")
        first = True
        for col, op, v in comps:
            d = cols.get(col)
            t = w if first else alloc()
            if lib.voxel_gpu_filter_bitmap(0, C.c_void_p(d.data_ptr()), n, _CMP[op], _bits(v), 0,
                                           C.c_void_p(t.data_ptr()), C.byref(cnt), None) != 0:
                check("voxel_gpu_filter_bitmap")
            if not first and lib.voxel_gpu_bitmap_combine(C.c_void_p(w.data_ptr()), C.c_void_p(t.data_ptr()), n, 0,
                                                          C.byref(cnt), None) != 0:
                check("voxel_gpu_bitmap_combine")
            first = False
        return w
    if kind in ("and", "or"):
        w = _bitmap(cols, e[1][0])
        for a in e[1][1:]:
            o = _bitmap(cols, a)
            if lib.voxel_gpu_bitmap_combine(C.c_void_p(w.data_ptr()), C.c_void_p(o.data_ptr()), n,
                                            0 if kind == "and" else 1, C.byref(cnt), None) != 0:
                check("voxel_gpu_bitmap_combine")
        return w
    if kind == "not":
        inner = _bitmap(cols, e[1])
        ones = alloc()                                     # every row: the always-true predicate's words
        if n and lib.voxel_gpu_filter_bitmap(0, C.c_void_p(cols.get(next(iter(cols.host))).data_ptr()), n, 6, 0, 0,
                                             C.c_void_p(ones.data_ptr()), C.byref(cnt), None) != 0:
            check("voxel_gpu_filter_bitmap")
        if lib.voxel_gpu_bitmap_combine(C.c_void_p(ones.data_ptr()), C.c_void_p(inner.data_ptr()), n, 2,
                                        C.byref(cnt), None) != 0:
            check("voxel_gpu_bitmap_combine")
        return ones
    raise ValueError("unsupported condition %r" % (kind,))


def _aggregate(cols, col, where):
    """COUNT/SUM/MIN/MAX of column ``col`` (None: COUNT(*) over the first
    column) over the rows where ``where`` holds."""
    from . import scan as vs
    agg_col = col if col is not None else next(iter(cols.host))
    a = cols.get(agg_col)
    if where is None:
        return vs.scan(a, vs.TRUE, 0)
    comps = _comparisons(where)
    if comps is not None and len(comps) <= 2:
        name = cols.name(agg_col)
        on_a = [c for c in comps if cols.name(c[0]) == name]
        others = [c for c in comps if cols.name(c[0]) != name]
        if len(others) == 0:                                 # both predicates on the aggregated column
            if len(comps) == 1:
                return vs.scan(a, _CMP[comps[0][1]], comps[0][2])
            return vs.scan(a, _CMP[comps[0][1]], comps[0][2], a, _CMP[comps[1][1]], comps[1][2])
        if len(others) == 1 and len(on_a) <= 1:              # one fused two-column pass
            op_a, t_a = (_CMP[on_a[0][1]], on_a[0][2]) if on_a else (vs.TRUE, 0.0)
            b = cols.get(others[0][0])
            return vs.scan(a, op_a, t_a, b, _CMP[others[0][1]], others[0][2])
    # a conjunction: a fused pass over the aggregated column and one other
    # (a comparison on each), its rows also limited by the selection words of
    # the remaining conjuncts (comparisons, OR / NOT groups)
    comps, groups = [], []
    for item in _conjuncts(where):
        c = _comparisons(item)
        if c is None:
            groups.append(item)
        else:
            comps += c
    if comps:
        name = cols.name(agg_col)
        on_a = [c for c in comps if cols.name(c[0]) == name]
        others = [c for c in comps if cols.name(c[0]) != name]
        op_a, t_a = (_CMP[on_a[0][1]], on_a[0][2]) if on_a else (vs.TRUE, 0.0)
        rest = [("cmp", ("col", c), op, ("const", v)) for c, op, v in on_a[1:] + others[1:]] + groups
        w = _bitmap(cols, ("and", rest)) if rest else None
        if others:
            b = cols.get(others[0][0])
            return vs.scan(a, op_a, t_a, b, _CMP[others[0][1]], others[0][2], valid_a=w)
        return vs.scan(a, op_a, t_a, valid_a=w)
    w = _bitmap(cols, where)
    return vs.scan(a, vs.TRUE, 0, valid_a=w)


def _result(func, r):
    if func == "count":
        return int(r["count"])
    if r["count"] == 0:
        return None if func != "sum" else 0.0
    if func == "sum":
        return float(r["sum"])
    if func == "avg":
        return float(r["sum"]) / r["count"]
    return float(r["min"] if func == "min" else r["max"])


def _run_lowered(q, table, data):
    cols = _Cols(table if isinstance(table, dict) else {"": np.asarray(table)})
    targets = q["targets"]
    where = q["where"]
    if q["group_by"] is not None:
        return _group_by_lowered(cols, q)
    funcs = [t for t in targets if t[0] == "func"]
    if funcs:
        if len(funcs) != len(targets):
            raise ValueError("mixing aggregates and plain columns needs GROUP BY")
        for _, fn, _arg in funcs:
            if fn not in _AGGS:
                raise NotImplementedError("aggregate %s not supported" % fn)
        cache = {}                                           # one scan per aggregated column
        for _, fn, arg in funcs:
            if arg != "*" and cols.name(arg) not in cache:
                cache[cols.name(arg)] = _aggregate(cols, arg, where)
        if not cache:                                        # COUNT(*) alone
            cache[None] = _aggregate(cols, None, where)
        any_scan = next(iter(cache.values()))                # COUNT(*): any scan's row count
        out = [_result(fn, any_scan if arg == "*" else cache[cols.name(arg)]) for _, fn, arg in funcs]
        return out[0] if len(out) == 1 else tuple(out)
    return _rows_lowered(cols, q)


def _selected(cols, col, where):
    """Values of ``col`` in the selected rows, row order (device tensor)."""
    from . import scan as vs
    from .devarray import DeviceArray
    a = cols.get(col)
    if where is None:
        return a
    w = _bitmap(cols, where)
    out = DeviceArray(a.n, np.float64)
    m = vs.nullable_compact(a, w, out)
    return out.view(m)


def _rows_lowered(cols, q):
    from ._lib import check, lib
    from .devarray import DeviceArray
    t = q["targets"][0]
    col = next(iter(cols.host)) if t[0] == "star" else t[1]
    if t[0] == "star" and len(cols.host) > 1:
        raise ValueError("SELECT * needs a one-column table")
    v = _selected(cols, col, q["where"])
    if q["order_by"] is not None:
        ocol, asc = q["order_by"]
        if cols.name(ocol) != cols.name(col):
            raise ValueError("ORDER BY another column than the selected one is not supported")
        n = v.numel()
        idx = DeviceArray(n, np.uint32)
        if n and lib.voxel_gpu_argsort_f64(C.c_void_p(v.data_ptr()), n, 0 if asc else 1, C.c_void_p(idx.ptr),
                                           None) != 0:
            check("voxel_gpu_argsort_f64")
        k = n if q["limit"] is None else min(n, q["limit"])
        out = DeviceArray(k, np.float64)
        if k and lib.voxel_gpu_gather_f64(C.c_void_p(v.data_ptr()), C.c_void_p(idx.ptr), k, C.c_void_p(out.ptr),
                                          None) != 0:
            check("voxel_gpu_gather_f64")
        return out.numpy()
    if q["limit"] is not None:
        v = v.view(min(v.numel(), q["limit"]))
    return v.numpy()


def _group_by_lowered(cols, q):
    from ._lib import check, lib
    from .devarray import DeviceArray
    key = q["group_by"]
    funcs = [t for t in q["targets"] if t[0] == "func"]
    if not funcs:
        raise ValueError("GROUP BY needs an aggregate")
    val_cols = {f[2] for f in funcs if f[2] != "*"}
    if len(val_cols) > 1:
        raise ValueError("GROUP BY over more than one aggregated column is not supported")
    vcol = val_cols.pop() if val_cols else key
    keys_f = _selected(cols, key, q["where"])
    vals = _selected(cols, vcol, q["where"])
    n = keys_f.numel()
    k64 = DeviceArray(n, np.int64)                         # keys as integers, toward zero
    if n and lib.voxel_gpu_f64_to_i64(C.c_void_p(keys_f.data_ptr()), n, C.c_void_p(k64.ptr), None) != 0:
        check("voxel_gpu_f64_to_i64")
    G = max(n, 1)
    res = [DeviceArray(G, np.int64 if i < 2 else np.float64) for i in range(6)]
    ng = C.c_uint64(0)
    if lib.voxel_gpu_groupby_f64(2, C.c_void_p(k64.data_ptr()), C.c_void_p(vals.data_ptr()), n, G,
                                 *(C.c_void_p(r.data_ptr()) for r in res), C.byref(ng), None) != 0:
        check("voxel_gpu_groupby_f64")
    g = int(ng.value)
    out = {key: res[0].view(g).numpy()}
    names = {"count": 1, "sum": 2, "avg": 3, "min": 4, "max": 5}
    for _, fn, arg in funcs:
        out["%s(%s)" % (fn, arg)] = res[names[fn]].view(g).numpy()
    return out


# ---------------------------------------------------------------------------
# Public API (voxel_sql.py:315-347)
# ---------------------------------------------------------------------------
def execute(sql: str, data=None, compat=None):
    """Runs a SELECT statement.  ``data`` maps table names to a one-column
    array (the reference's form) or to a dict of named columns; anything else
    is the table ``t``.  ``compat`` picks the plan (None: the reference's plan
    for a one-column table, the lowered plan for a dict of columns)."""
    compiler = SqlCompiler(data if isinstance(data, dict) else {"t": data}, compat=compat)
    return compiler.execute(sql)

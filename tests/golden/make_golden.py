"""Writes tests/golden/reference_vectors.json from the reference's own C++.

Runs only where the reference checkout exists (it builds oracle/_ref with
oracle/Makefile).  Every value comes from the reference's classes driven by
oracle/ref_driver.cpp: Engine<T>::Run (fast path and Tier-0), RunParallel,
the x86-64 JIT, RadixSort, SortOperator, HashJoin, HashAggregator,
VectorFilter/ColumnFilter/SelectionVector, TopKSelect, plus the program words
of the reference's Instruction constructors.  Inputs are the reference
examples' own: std::mt19937_64 streams with the seeds of validate.cpp,
jit_test.cpp and demo.cpp, and the literal arrays of test_suite.cpp/c_test.cpp.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)

from oracle import oracle as o  # noqa: E402

OUT = os.path.join(os.path.dirname(__file__), "reference_vectors.json")


def enc(op, rd=0, ra=0, rb=0, imm=0):
    return int(o.R.vr_encode(op, rd, ra, rb, imm & 0xFFF))


def vload(vd, ra, seg, cnt): return enc(0x70, vd, ra, 0, ((seg & 0xF) << 8) | (cnt & 0xFF))


def programs():
    fast = [vload(0, 1, 0, 0), enc(0xC5, 1, 0, 3), enc(0xD0, 5, 1), enc(0x2A, 0, 0, 5),
            enc(0x20, 1, 1, 4), enc(0x40, 0, 1, 2), enc(0x52, imm=-6), enc(0x01)]
    mov = enc(0x10, 4, 0, 0, 4)
    vq_sum = [mov] + fast
    vq_count = [mov, vload(0, 1, 0, 0), enc(0xC5, 1, 0, 3), enc(0xD7, 5, 1), enc(0x20, 0, 0, 5),
                enc(0x20, 1, 1, 4), enc(0x40, 0, 1, 2), enc(0x52, imm=-6), enc(0x01)]
    def sel(red, br):
        return [mov, vload(0, 1, 0, 0), enc(0xC5, 1, 0, 3), enc(red, 5, 1), enc(0x40, 0, 0, 5),
                enc(br, imm=2), enc(0x11, 0, 5), enc(0x20, 1, 1, 4), enc(0x40, 0, 1, 2), enc(0x52, imm=-8),
                enc(0x01)]
    vq_min = sel(0xD5, 0x5A)
    vq_max = sel(0xD6, 0x5C)
    vq_sum_lt = [mov, vload(0, 1, 0, 0), enc(0xC0, 1, 0, 3, 2), enc(0xD0, 5, 1), enc(0x2A, 0, 0, 5),
                 enc(0x20, 1, 1, 4), enc(0x40, 0, 1, 2), enc(0x52, imm=-6), enc(0x01)]
    i64_add = [mov, vload(0, 1, 0, 0), enc(0xC5, 1, 0, 3), enc(0xD0, 5, 1), enc(0x20, 0, 0, 5),
               enc(0x20, 1, 1, 4), enc(0x40, 0, 1, 2), enc(0x52, imm=-6), enc(0x01)]
    return dict(fast=fast, vq_sum=vq_sum, vq_count=vq_count, vq_min=vq_min, vq_max=vq_max,
                vq_sum_lt=vq_sum_lt, i64_add=i64_add)


def run_engine(kind, data, code, regs):
    r = np.zeros(64, np.uint64)
    for k, v in regs.items():
        r[k] = v & 0xFFFFFFFFFFFFFFFF
    c = np.array(code, np.uint32)
    getattr(o.R, "vr_engine_run_" + kind)(o.ptr(data), data.size, o.ptr(c), c.size, o.ptr(r))
    return [int(x) for x in r]


def main():
    if o.R is None:
        o.build()
        import importlib
        importlib.reload(o)
    assert o.R is not None, "reference build (oracle/_ref) unavailable"
    P = programs()
    g = {"programs": {k: ["%08x" % w for w in v] for k, v in P.items()}}

    # validate.cpp / jit_test.cpp data: mt19937_64(42), uniform [0,1000), 1e6
    n = 1_000_000
    x = np.empty(n, np.float64)
    o.R.vr_mt_uniform_f64(42, n, 0.0, 1000.0, o.ptr(x))
    g["data_f64_seed42"] = {"seed": 42, "n": n, "lo": 0.0, "hi": 1000.0,
                            "first": ["%016x" % int(b) for b in x[:3].view(np.uint64)]}
    thr = o.f64_bits(500.0)
    base = {0: 0, 1: 0, 2: n, 3: thr, 4: 4}
    res = {}
    regs = run_engine("f64", x, P["fast"], base)
    res["fast"] = {"R0": "%016x" % regs[0], "R1": regs[1]}
    for name in ("vq_sum", "vq_count", "vq_min", "vq_max"):
        regs = run_engine("f64", x, P[name], {0: 0, 1: 0, 2: n, 3: thr})
        res[name] = {"R0": "%016x" % regs[0], "R1": regs[1], "R4": regs[4], "R5": "%016x" % regs[5]}
    regs = run_engine("f64", x, P["vq_sum_lt"], {0: 0, 1: 0, 2: n, 3: thr})
    res["vq_sum_lt"] = {"R0": "%016x" % regs[0], "R5": "%016x" % regs[5]}
    # JIT: voxel_jit_run semantics (R2 = data_len)
    c = np.array(P["fast"], np.uint32)
    size = np.zeros(1, np.uint64)
    fn = o.R.vr_jit_compile(o.ptr(c), c.size, o.ptr(size))
    res["jit"] = {"result": "%016x" % o.f64_bits(o.R.vr_jit_entry(fn, o.ptr(x), n, n, 500.0)),
                  "code_size": int(size[0])}
    # JIT on a row count that is not a multiple of 8 reads past the end in the
    # reference; record a size that is a multiple of 8 instead.
    res["jit_999992"] = {"result": "%016x" % o.f64_bits(o.R.vr_jit_entry(fn, o.ptr(x), 999992, 999992, 500.0))}
    o.R.vr_jit_release(fn)
    # RunParallel(8), result register 0
    r = np.zeros(64, np.uint64)
    r[2], r[3] = n, thr
    o.R.vr_engine_run_parallel_f64(o.ptr(x), n, o.ptr(c), c.size, o.ptr(r), 8, 0)
    res["run_parallel_8"] = {"R0": "%016x" % int(r[0])}
    r = np.zeros(64, np.uint64)
    r[2], r[3] = n, thr
    o.R.vr_engine_run_parallel_f64(o.ptr(x), n, o.ptr(c), c.size, o.ptr(r), 7, 0)
    res["run_parallel_7"] = {"R0": "%016x" % int(r[0])}
    # fast path tail drop: first 1003 rows, threshold 0 (every row passes)
    regs = run_engine("f64", x[:1003].copy(), P["fast"], {2: 1003, 3: o.f64_bits(0.0)})
    res["tail_1003"] = {"R0": "%016x" % regs[0], "R1": regs[1]}
    g["seed42"] = res

    # demo.cpp #2 style: i64 mt19937_64(7) uniform_int(0,1000), Tier-0 ADD variant, > 500
    xi = np.empty(n, np.int64)
    o.R.vr_mt_uniform_i64(7, n, 0, 1000, o.ptr(xi))
    regs = run_engine("i64", xi, P["i64_add"], {0: 0, 1: 0, 2: n, 3: 500})
    g["i64_seed7"] = {"seed": 7, "n": n, "lo": 0, "hi": 1000, "thr": 500, "R0": regs[0], "R1": regs[1],
                      "first": [int(v) for v in xi[:3]]}
    # i32 mt19937_64(9) uniform_int(0, 2^30), fast path > 2^29 (wrapping sum)
    x32 = np.empty(n, np.int32)
    o.R.vr_mt_uniform_i32(9, n, 0, 1 << 30, o.ptr(x32))
    regs = run_engine("i32", x32, P["fast"], {2: n, 3: 1 << 29})
    g["i32_seed9"] = {"seed": 9, "n": n, "lo": 0, "hi": 1 << 30, "thr": 1 << 29, "R0": "%016x" % regs[0],
                      "R1": regs[1], "first": [int(v) for v in x32[:3]]}

    # test_suite.cpp:254-321 / c_test.cpp literal inputs
    d8 = np.arange(1, 9, dtype=np.float64)
    regs = run_engine("f64", d8, [enc(0x10, 4, 0, 0, 4)] + P["fast"], {2: 8, 3: o.f64_bits(4.0)})
    regs_c = run_engine("f64", d8, P["fast"], {2: 8, 3: o.f64_bits(4.0)})
    di = np.array([10, 20, 30, 40], np.int64)
    regs_i = run_engine("i64", di, P["i64_add"], {2: 4, 3: 15})
    c8 = np.array(P["fast"], np.uint32)
    fn = o.R.vr_jit_compile(o.ptr(c8), c8.size, None)
    jit8 = o.R.vr_jit_entry(fn, o.ptr(d8), 8, 8, 4.0)
    o.R.vr_jit_release(fn)
    g["literals"] = {"engine_f64_tier0": o.bits_f64(regs[0]), "engine_f64_fast": o.bits_f64(regs_c[0]),
                     "engine_i64": regs_i[0], "jit_f64": jit8}

    # VectorFilter<f64>::ApplyGT on test_suite.cpp:410-427 data
    vf = np.array([1, 2, 5, 6, 3, 4, 7, 8], np.float64)
    words = np.zeros(1, np.uint64)
    pc = o.R.vr_bitmap_f64(o.ptr(vf), vf.size, 4, 4.0, 0.0, o.ptr(words))
    g["bitmap_gt4"] = {"data": vf.tolist(), "word": int(words[0]), "pass": int(pc)}

    # RadixSort<u64> with payload
    keys = np.array([5, 3, 5, 1, 3, 2**64 - 1, 0], np.uint64)
    idx = np.arange(keys.size, dtype=np.uint32)
    o.R.vr_radix_sort_u64(o.ptr(keys), o.ptr(idx), keys.size)
    g["radix_u64"] = {"in": [5, 3, 5, 1, 3, 2**64 - 1, 0], "keys": [int(k) for k in keys], "idx": idx.tolist()}
    # SortOperator asc/desc (c_test.cpp Test 5) + ties and signed zeros
    for name, arr in (("sort_c_test", [3, 1, 4, 1, 5]),
                      ("sort_ties", [2.0, -0.0, 1.0, 0.0, 2.0, -1.5, 0.0, -0.0, 1.0])):
        a = np.array(arr, np.float64)
        ia, idd = np.zeros(a.size, np.uint32), np.zeros(a.size, np.uint32)
        o.R.vr_sort_f64(o.ptr(a), a.size, 0, o.ptr(ia))
        o.R.vr_sort_f64(o.ptr(a), a.size, 1, o.ptr(idd))
        g[name] = {"data": arr, "asc": ia.tolist(), "desc": idd.tolist()}
    # TopK (c_test.cpp Test 6)
    tk = np.array([3, 1, 4, 1, 5, 9, 2, 6], np.float64)
    out = np.zeros(3, np.float64)
    o.R.vr_topk_f64(o.ptr(tk), tk.size, 3, o.ptr(out), 1)
    g["topk3"] = {"data": tk.tolist(), "largest": out.tolist()}
    # HashJoin<i64>: unique keys and duplicate build keys
    for name, b, p in (("join_small", [10, 20, 30], [20, 40, 10, 20]),
                       ("join_dups", [7, 3, 7, 9, 3, 7], [3, 7, 8, 9, 7, 1])):
        bb, pp = np.array(b, np.int64), np.array(p, np.int64)
        pay = np.array([100 * (i + 1) for i in range(len(b))], np.int64)
        lo, ro = np.zeros(len(p), np.uint32), np.zeros(len(p), np.uint32)
        s = np.zeros(1, np.int64)
        m = o.R.vr_hash_join_i64(o.ptr(bb), bb.size, o.ptr(pay), o.ptr(pp), pp.size, o.ptr(lo), o.ptr(ro), o.ptr(s))
        g[name] = {"build": b, "probe": p, "payload": pay.tolist(), "matches": int(m),
                   "left": lo[:m].tolist(), "right": ro[:m].tolist(), "payload_sum": int(s[0])}
    # HashAggregator<i32,f64> (test_suite.cpp:432-452 style)
    for name, k, v in (("agg_small", [3, 1, 3, 2], [1.5, 2.0, 2.5, 4.0]),
                       ("agg_c_test", [0, 1, 0, 1, 0], [10, 20, 30, 40, 50])):
        kk, vv = np.array(k, np.int32), np.array(v, np.float64)
        G = 16
        ok, oc = np.zeros(G, np.int32), np.zeros(G, np.uint64)
        osum, omin, omax, omean, om2 = (np.zeros(G, np.float64) for _ in range(5))
        ng = o.R.vr_hash_agg_i32_f64(o.ptr(kk), o.ptr(vv), kk.size, 10, G, o.ptr(ok), o.ptr(oc), o.ptr(osum),
                                     o.ptr(omin), o.ptr(omax), o.ptr(omean), o.ptr(om2))
        g[name] = {"keys": k, "vals": [float(t) for t in v], "groups": [
            {"key": int(ok[i]), "count": int(oc[i]), "sum": float(osum[i]), "min": float(omin[i]),
             "max": float(omax[i]), "mean": float(omean[i]), "m2": float(om2[i])} for i in range(ng)]}

    with open(OUT, "w") as f:
        json.dump(g, f, indent=1, sort_keys=True)
    print("wrote", OUT)


if __name__ == "__main__":
    main()

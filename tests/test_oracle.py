"""Pins the CPU oracle (oracle/voxel_oracle.c) before it is trusted as the checker.

Each restated function is checked against (a) the golden vectors produced by the
reference's own C++ (tests/golden/reference_vectors.json, made by
tests/golden/make_golden.py from oracle/_ref), (b) the reference build itself
where it is present, and (c) what mathematics fixes: exact rational sums,
correctly rounded fsum, library stable sorts, brute-force joins and groupings.
CPU only.
"""
from __future__ import annotations

import ctypes as C

import math
import random
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as o

B = lambda h: int(h, 16)  # noqa: E731


def test_mt19937_64_matches_reference_examples(golden, x42):
    g = golden["data_f64_seed42"]
    assert [int(b) for b in x42[:3].view(np.uint64)] == [B(h) for h in g["first"]]
    assert x42[0] == 755.15553295453901 and x42[1] == 639.0313938546974
    xi = o.mt_uniform_i64(7, 5, 0, 1000)
    assert xi[:3].tolist() == golden["i64_seed7"]["first"]
    x32 = o.mt_uniform_i32(9, 5, 0, 1 << 30)
    assert x32[:3].tolist() == golden["i32_seed9"]["first"]


@pytest.mark.skipif(o.R is None, reason="reference build absent")
def test_mt_generators_match_libstdcxx_stream():
    for seed in (1, 42, 12345):
        a = o.mt_uniform_f64(seed, 20000, -3.0, 7.5)
        b = np.empty_like(a)
        o.R.vr_mt_uniform_f64(seed, a.size, -3.0, 7.5, o.ptr(b))
        assert np.array_equal(a, b)
        ai = o.mt_uniform_i64(seed, 20000, -5, 1 << 40)
        bi = np.empty_like(ai)
        o.R.vr_mt_uniform_i64(seed, ai.size, -5, 1 << 40, o.ptr(bi))
        assert np.array_equal(ai, bi)
        a3 = o.mt_uniform_i32(seed, 20000, -7, 12345)
        b3 = np.empty_like(a3)
        o.R.vr_mt_uniform_i32(seed, a3.size, -7, 12345, o.ptr(b3))
        assert np.array_equal(a3, b3)


def test_exact_accumulator_against_fractions():
    rng = random.Random(5)
    for trial in range(30):
        vals = []
        for _ in range(rng.randint(1, 60)):
            e = rng.choice([rng.randint(-1074, -1000), rng.randint(-60, 60), rng.randint(900, 1023)])
            m = rng.random() * rng.choice([-1, 1])
            vals.append(math.ldexp(m, e))
        vals += [5e-324, -5e-324, 1.7976931348623157e308, -1.7976931348623157e308]
        x = np.array(vals, np.float64)
        exact = sum((Fraction(v) for v in vals), Fraction(0))
        assert o.exact_sum(x) == exact
    # catastrophic cancellation
    assert o.exact_sum(np.array([1e300, 1.0, -1e300, 2.0**-1074])) == 1 + Fraction(1, 2**1074)


def test_exact_accumulator_rounds_like_fsum():
    rng = np.random.default_rng(3)
    x = rng.standard_normal(200_000) * np.exp(rng.uniform(-30, 30, 200_000))
    assert float(o.exact_sum(x)) == math.fsum(x.tolist())


def test_filter_count_and_exact_sum_golden(x42):
    r = o.relational(x42, o.CMP["GT"], 500.0)
    assert r["count"] == 500_720                          # docs/Benchmarks.md:7
    assert r["count"] == int(np.count_nonzero(x42 > 500.0))
    assert abs(float(r["sum"]) - 375_538_740.218) < 1e-3
    assert float(r["sum"]) == math.fsum(x42[x42 > 500.0].tolist())
    assert r["min"] == x42[x42 > 500].min() and r["max"] == 999.99852628798419


def test_relational_compound_against_numpy():
    rng = np.random.default_rng(11)
    for dt in (np.float64, np.float32, np.int64, np.int32):
        if np.dtype(dt).kind == "f":
            a = (rng.random(50_000) * 1000).astype(dt)
            b = (rng.random(50_000) * 1000).astype(dt)
            a[::97] = np.nan
        else:
            a = rng.integers(0, 2**31, 50_000).astype(dt)
            b = rng.integers(0, 2**31, 50_000).astype(dt)
        ta, tb = dt(a[7]) if not np.isnan(a[7]) else dt(300), dt(b[9])
        for opa, fa in ((4, np.greater), (1, np.not_equal), (3, np.less_equal)):
            m = fa(a, ta) & (b < tb)
            r = o.relational(a, opa, ta, b, 2, tb)
            assert r["count"] == int(m.sum())
            sel = a[m]
            if np.dtype(dt).kind == "f":
                if not np.isnan(sel).any():
                    assert r["sum"] == sum((Fraction(float(v)) for v in sel), Fraction(0))
                fin = sel[~np.isnan(sel)]
                assert (r["min"] is None) == (fin.size == 0)
                if fin.size:
                    assert r["min"] == fin.min() and r["max"] == fin.max()
            else:
                assert r["sum"] == int(sel.astype(np.int64).sum())
                if sel.size:
                    assert r["min"] == sel.min() and r["max"] == sel.max()


def _fast(x, total, off=0, thr=500.0, acc0=0.0, fop=0xC5):
    acc = np.zeros(1, np.float64)
    end = o.O.vo_fastpath_f64(o.ptr(x), x.size, total, off, thr, acc0, fop, o.ptr(acc))
    return int(acc.view(np.uint64)[0]), int(end)


def test_fastpath_restatement_is_bit_exact(golden, x42):
    g = golden["seed42"]
    bits, end = _fast(x42, x42.size)
    assert bits == B(g["fast"]["R0"]) and end == g["fast"]["R1"]
    # a plain sequential sum differs in the last bits: the golden value pins the order
    assert o.f64_bits(float(np.cumsum(np.where(x42 > 500, x42, 0.0))[-1])) != bits
    # tail of 1003 rows, every row passes, 3 rows dropped
    t = x42[:1003].copy()
    bits, end = _fast(t, 1003, thr=0.0)
    assert bits == B(g["tail_1003"]["R0"]) and end == 1000
    # literal programs (test_suite.cpp:254-285, c_test.cpp:8-36)
    d8 = np.arange(1, 9, dtype=np.float64)
    assert o.bits_f64(_fast(d8, 8, thr=4.0)[0]) == 26.0 == golden["literals"]["engine_f64_fast"]


def test_fastpath_i32_wraps(golden):
    g = golden["i32_seed9"]
    x = o.mt_uniform_i32(9, g["n"], 0, 1 << 30)
    acc = np.zeros(1, np.int32)
    end = o.O.vo_fastpath_i32(o.ptr(x), x.size, x.size, 0, 1 << 29, 0, 0xC5, o.ptr(acc))
    assert (int(acc[0]) & 0xFFFFFFFFFFFFFFFF) == B(g["R0"]) and end == g["R1"]
    exact = int(x[x > (1 << 29)].astype(np.int64).sum())
    assert ((exact + 2**31) % 2**32) - 2**31 == int(acc[0])          # sum mod 2^32, sign-extended


def _tier0(x, red, accop, filt=4, thr=500.0, total=None, off=0):
    total = x.size if total is None else total
    part = np.zeros(1, np.uint64)
    fn = o.O.vo_tier0_f64 if x.dtype == np.float64 else o.O.vo_tier0_i64
    r = fn(o.ptr(x), x.size, off, total, filt, thr, 0, red, accop, 0, o.ptr(part))
    return int(r), int(part[0])


def test_tier0_voxel_query_shapes(golden, x42):
    g = golden["seed42"]
    assert _tier0(x42, 0, 0) == (B(g["vq_sum"]["R0"]), B(g["vq_sum"]["R5"]))
    assert _tier0(x42, 1, 1) == (B(g["vq_count"]["R0"]), B(g["vq_count"]["R5"]))
    assert B(g["vq_count"]["R0"]) == 500_720
    assert _tier0(x42, 2, 2) == (B(g["vq_min"]["R0"]), B(g["vq_min"]["R5"]))
    assert _tier0(x42, 3, 3) == (B(g["vq_max"]["R0"]), B(g["vq_max"]["R5"]))
    assert _tier0(x42, 0, 0, filt=2) == (B(g["vq_sum_lt"]["R0"]), B(g["vq_sum_lt"]["R5"]))
    # Tier-0 sum equals the fast path's bits (same group order)
    assert B(g["vq_sum"]["R0"]) == B(g["fast"]["R0"])


def test_tier0_i64_add_variant(golden):
    g = golden["i64_seed7"]
    x = o.mt_uniform_i64(7, g["n"], 0, 1000)
    assert _tier0(x, 0, 1, thr=500)[0] == g["R0"] == int(x[x > 500].sum())
    d = np.array([10, 20, 30, 40], np.int64)
    assert _tier0(d, 0, 1, thr=15)[0] == 90 == golden["literals"]["engine_i64"]


def test_tier0_minmax_fold_is_a_bit_compare():
    # negative values: the i64 compare of f64 bits is not the f64 order (Q10)
    x = np.array([3.0, -1.0, 2.0, 5.0, -4.0, 7.0, 1.0, 0.5], np.float64)
    r_max, _ = _tier0(x, 3, 3, filt=6)
    r_min, _ = _tier0(x, 2, 2, filt=6)
    # group maxima 5.0, 7.0 (bits positive), acc starts 0.0 -> 7.0
    assert o.bits_f64(r_max) == 7.0
    # group minima -1.0, -4.0: from 0.0, CMP 0 - bits(-1.0) > 0 -> take -1.0;
    # then bits(-1.0) - bits(-4.0) < 0 (as i64) -> keep -1.0
    assert o.bits_f64(r_min) == -1.0


@pytest.mark.skipif(o.R is None, reason="reference build absent")
def test_tier0_against_reference_random_programs():
    rng = np.random.default_rng(21)
    progs = {}
    import json, os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))
    for k in ("vq_sum", "vq_count", "vq_min", "vq_max"):
        progs[k] = np.array([int(w, 16) for w in g["programs"][k]], np.uint32)
    kinds = {"vq_sum": (0, 0), "vq_count": (1, 1), "vq_min": (2, 2), "vq_max": (3, 3)}
    for trial in range(12):
        n = int(rng.integers(1, 400)) * 4
        x = rng.standard_normal(n) * 100
        x[rng.integers(0, n, 5)] = 0.0
        if trial % 3 == 0:
            x[rng.integers(0, n, 3)] = np.nan
        thr = float(rng.standard_normal() * 50)
        for k, (red, accop) in kinds.items():
            regs = np.zeros(64, np.uint64)
            regs[2], regs[3] = n, o.f64_bits(thr)
            o.R.vr_engine_run_f64(o.ptr(x), n, o.ptr(progs[k]), progs[k].size, o.ptr(regs))
            r, part = _tier0(x, red, accop, thr=thr)
            assert (r, part) == (int(regs[0]), int(regs[5])), (trial, k)


@pytest.mark.skipif(o.R is None, reason="reference build absent")
@pytest.mark.parametrize("is_f64", [True, False])
def test_tier0_multi_restatement_against_reference_engine(is_f64):
    """vo_tier0_multi_* (several filters, chained or not, and up to three
    reductions per loop) against the reference engine running the same
    words: every accumulator and partial register bit-identical."""
    from tests import tier0_gen as tg
    rng = np.random.default_rng(31 if is_f64 else 32)
    for trial in range(40):
        n = int(rng.integers(1, 300)) * 4
        if is_f64:
            x = np.round(rng.standard_normal(n) * 100, 1)
            x[rng.integers(0, n, 5)] = 0.0
            if trial % 4 == 0:
                x[rng.integers(0, n, 3)] = np.nan
            thr = [float(np.round(rng.standard_normal() * 60, 1)) for _ in range(2)]
            tbits = [o.f64_bits(t) for t in thr]
        else:
            x = rng.integers(-1000, 1000, n).astype(np.int64)
            x[rng.integers(0, n, 5)] = 0
            thr = [int(rng.integers(-600, 600)) for _ in range(2)]
            tbits = [t & (2**64 - 1) for t in thr]
        code, spec = tg.make_program(rng, is_f64, allow_ne=True)
        regs = np.zeros(64, np.uint64)
        regs[2], regs[3], regs[6] = n, tbits[0], tbits[1]
        run = o.R.vr_engine_run_f64 if is_f64 else o.R.vr_engine_run_i64
        run(o.ptr(x), n, o.ptr(code), code.size, o.ptr(regs))
        acc, part = tg.oracle_run(o, x, n, spec, thr, is_f64)
        for b in range(len(spec["red"])):
            assert int(regs[tg.ACCS[b]]) == acc[b], (trial, b, spec)
            assert int(regs[tg.PARTS[b]]) == part[b], (trial, b, spec)
        assert int(regs[1]) == n


@pytest.mark.skipif(o.R is None, reason="reference build absent")
def test_scalar_words_restatement_against_reference_engine():
    """vo_scalar_run against the reference engine on straight-line programs
    of integer, f64, bitwise and conversion words (including DIV/MOD/DIVF by
    zero, which stop the program, and TRUNC of NaN or out-of-range values):
    all 16 scalar registers bit-identical."""
    from tests import tier0_gen as tg
    rng = np.random.default_rng(41)
    x = np.zeros(4, np.float64)
    checked = stopped = 0
    while checked < 300:
        regs = np.zeros(64, np.uint64)
        for r in range(16):
            kind = rng.integers(0, 4)
            regs[r] = (int(rng.integers(-50, 50)) & (2**64 - 1) if kind == 0 else
                       o.f64_bits(float(rng.standard_normal() * 1e3)) if kind == 1 else
                       int(rng.integers(0, 2**63)) if kind == 2 else 0)
        words = np.array(tg.scalar_words(rng, int(rng.integers(1, 12)), range(16), range(16)) + [tg.w(0x01)],
                         np.uint32)
        mine = regs.copy()
        rc = o.O.vo_scalar_run(o.ptr(words), words.size - 1, o.ptr(mine))
        if rc < 0:
            continue                                   # INT64_MIN / -1: the reference traps
        ref = regs.copy()
        o.R.vr_engine_run_f64(o.ptr(x), x.size, o.ptr(words), words.size, o.ptr(ref))
        assert np.array_equal(mine[:16], ref[:16]), [hex(int(v)) for v in words]
        checked += 1
        stopped += rc == 0
    assert stopped > 0


@pytest.mark.skipif(o.R is None, reason="reference build absent")
@pytest.mark.parametrize("is_f64", [True, False])
def test_segment_aggregates_restatement_against_reference_engine(is_f64):
    """vo_straight_run (AGG_COUNT..AGG_COUNT_DISTINCT, AGG_MEDIAN,
    AGG_PERCENTILE with the engine's persistent slots, mixed with scalar
    words) against the reference engine running the same words, several
    programs in a row on one engine state: all 16 scalar registers
    bit-identical."""
    from tests import tier0_gen as tg
    rng = np.random.default_rng(51 if is_f64 else 52)
    done = 0
    for trial in range(30):
        n = int(rng.integers(1, 3000))
        if is_f64:
            x = np.round(rng.standard_normal(n) * 1e3, 3)
            x[rng.integers(0, n, 4)] = 0.0
            if trial % 5 == 0:
                x[rng.integers(0, n, 2)] = np.inf
        else:
            x = rng.integers(-10**6, 10**6, n).astype(np.int64)
            x[rng.integers(0, n, 4)] = 0
        st = o.O.vo_agg_create(1 if is_f64 else 0)
        regs = np.zeros(64, np.uint64)
        ref_regs = np.zeros(64, np.uint64)
        progs = [tg.straight_program(rng, n) for _ in range(3)]
        # the reference keeps its AggState slots across runs of one engine:
        # run the three programs back to back on one engine as one program
        whole = np.concatenate([p[:-1] for p in progs] + [progs[-1][-1:]]).astype(np.uint32)
        rc = o.O.vo_straight_run(st, o.ptr(x.view(np.uint64)), n, o.ptr(whole), whole.size, o.ptr(regs))
        o.O.vo_agg_destroy(st)
        if rc < 0:
            continue
        run = o.R.vr_engine_run_f64 if is_f64 else o.R.vr_engine_run_i64
        run(o.ptr(x), n, o.ptr(whole), whole.size, o.ptr(ref_regs))
        assert np.array_equal(regs[:16], ref_regs[:16]), (trial, [hex(int(v)) for v in whole])
        done += 1
    assert done >= 20


@pytest.mark.skipif(o.R is None, reason="reference build absent")
@pytest.mark.parametrize("is_f64", [True, False])
def test_tier0_expr_restatement_against_reference_engine(is_f64):
    """vo_tier0_expr_* (loops with VADD..VRCP, VSADD..VSMAX, VCMPxx, the
    logical and shift words and filters on any computed vector, up to three
    reductions) against the reference engine running the same words: every
    accumulator and partial bit-identical.  After the random programs, each
    binary and unary word is the last word of two programs, reduced first."""
    from tests import tier0_gen as tg
    rng = np.random.default_rng(91 if is_f64 else 92)
    forced = tg.VEC_BIN + tg.VEC_UN + [tg.VSPLAT]
    for trial in range(60 + 2 * len(forced)):
        last = forced[(trial - 60) // 2] if trial >= 60 else None
        n = int(rng.integers(1, 300)) * 4
        regs = np.zeros(64, np.uint64)
        if is_f64:
            x = np.round(rng.standard_normal(n) * 100, 2)
            x[rng.integers(0, n, 6)] = 0.0
            if trial % 4 == 0:
                x[rng.integers(0, n, 2)] = np.nan
            for r in tg.SCAL_REGS:
                regs[r] = o.f64_bits(float(np.round(rng.standard_normal() * 10, 1)) if rng.random() < 0.9 else 0.0)
        else:
            x = rng.integers(-1000, 1000, n).astype(np.int64)
            for r in tg.SCAL_REGS:
                regs[r] = int(rng.integers(-50, 50)) & (2**64 - 1)
        regs[2] = n
        code, spec = tg.make_expr_program(rng, is_f64, nops=int(rng.integers(1, 7)), last=last)
        if not is_f64 and tg.expr_traps(o, x, n, spec, regs):
            continue                       # the reference's host would take SIGFPE
        acc, part = tg.oracle_expr(o, x, n, spec, regs, is_f64)
        ref = regs.copy()
        run = o.R.vr_engine_run_f64 if is_f64 else o.R.vr_engine_run_i64
        run(o.ptr(x), n, o.ptr(code), code.size, o.ptr(ref))
        for b in range(len(spec["red"])):
            assert int(ref[tg.ACCS[b]]) == acc[b], (trial, b, spec)
            assert int(ref[tg.PARTS[b]]) == part[b], (trial, b, spec)
            if spec["red"][b] == 1:     # the lanes of the reduced register (vo_tier0_expr_col_*)
                col = tg.oracle_expr_col(o, x, n, spec, regs, spec["bvreg"][b], is_f64)
                live = col[~np.isnan(col)] if is_f64 else col          # VCOUNT counts ordered != 0
                assert int(np.count_nonzero(live)) == int(ref[tg.ACCS[b]]), (trial, b, spec)


def test_jit_restatement(golden, x42):
    g = golden["seed42"]
    r = o.O.vo_jit_f64(o.ptr(x42), x42.size, x42.size, 0, 0xC5, 500.0)
    assert o.f64_bits(r) == B(g["jit"]["result"])
    r = o.O.vo_jit_f64(o.ptr(x42), 999992, 999992, 0, 0xC5, 500.0)
    assert o.f64_bits(r) == B(g["jit_999992"]["result"])


def test_run_parallel_restatement(golden, x42):
    g = golden["seed42"]
    for t in (8, 7):
        r = o.O.vo_run_parallel_fast_f64(o.ptr(x42), x42.size, t, x42.size, 0, 500.0, 0.0, 0xC5)
        assert o.f64_bits(r) == B(g["run_parallel_%d" % t]["R0"])


@pytest.mark.skipif(o.R is None, reason="reference build absent")
def test_run_parallel_tier0_restatement_against_reference():
    """The per-chunk restatement above equals the reference engine's own
    RunParallel (oracle/_ref, threads included) bit for bit on generated
    Tier-0 loops: sums, counts, minima and maxima, 1 to 9 workers."""
    from tests import tier0_gen as tg
    rng = np.random.default_rng(303)
    checked = 0
    for trial in range(40):
        n = int(rng.integers(1, 600)) * 4
        x = rng.standard_normal(n) * 100
        x[rng.integers(0, n, 5)] = 0.0
        thr = [float(rng.standard_normal() * 60) for _ in range(2)]
        code, spec = tg.make_program(rng, True)
        if not tg.mergeable(spec):
            continue
        threads = int(rng.integers(1, 10))
        b = int(rng.integers(0, len(spec["red"])))
        use_acc = bool(rng.integers(0, 2))
        rr = tg.ACCS[b] if use_acc else tg.PARTS[b]
        regs = np.zeros(64, np.uint64)
        regs[2], regs[3], regs[6] = n, o.f64_bits(thr[0]), o.f64_bits(thr[1])
        c = np.array(code, np.uint32)
        o.R.vr_engine_run_parallel_f64(o.ptr(x), n, o.ptr(c), c.size, o.ptr(regs), threads, rr)
        assert int(regs[rr]) == tg.run_parallel_expected(o, x, n, spec, thr, threads, b, use_acc), (trial, spec)
        checked += 1
    assert checked >= 20


def test_bitmap_and_selection(golden):
    g = golden["bitmap_gt4"]
    x = np.array(g["data"], np.float64)
    w = np.zeros(1, np.uint64)
    assert o.O.vo_bitmap_f64(o.ptr(x), x.size, 4, 4.0, 0.0, o.ptr(w)) == g["pass"] == 4
    assert int(w[0]) == g["word"] == (1 << 2) | (1 << 3) | (1 << 6) | (1 << 7)
    rng = np.random.default_rng(2)
    a = rng.integers(0, 100, 10_001).astype(np.int32)
    nw = (a.size + 63) // 64
    wa, wb = np.zeros(nw, np.uint64), np.zeros(nw, np.uint64)
    o.O.vo_bitmap_i32(o.ptr(a), a.size, 2, 30, 0, o.ptr(wa))
    o.O.vo_bitmap_i32(o.ptr(a), a.size, 5, 10, 0, o.ptr(wb))
    bits = np.unpackbits(wa.view(np.uint8), bitorder="little")[:a.size].astype(bool)
    assert np.array_equal(bits, a < 30)
    cnt = o.O.vo_bitmap_combine(o.ptr(wa), o.ptr(wb), a.size, 0)
    m = (a < 30) & (a >= 10)
    assert cnt == m.sum()
    idx = np.zeros(a.size, np.uint32)
    k = o.O.vo_select_from_bitmap(o.ptr(wa), a.size, o.ptr(idx))
    assert np.array_equal(idx[:k], np.nonzero(m)[0])
    if o.R is not None:
        wr = np.zeros(nw, np.uint64)
        assert o.R.vr_bitmap_i32(o.ptr(a), a.size, 2, 30, 0, o.ptr(wr)) == (a < 30).sum()
        wr2 = np.zeros(nw, np.uint64)
        o.R.vr_bitmap_i32(o.ptr(a), a.size, 5, 10, 0, o.ptr(wr2))
        assert o.R.vr_column_filter(o.ptr(wr), o.ptr(wr2), a.size, 0) == cnt
        assert np.array_equal(wr, wa)


def test_radix_sort_restatement(golden):
    g = golden["radix_u64"]
    k = np.array(g["in"], np.uint64)
    i = np.arange(k.size, dtype=np.uint32)
    o.O.vo_radix_sort_u64(o.ptr(k), o.ptr(i), k.size)
    assert k.tolist() == g["keys"] and i.tolist() == g["idx"]
    rng = np.random.default_rng(4)
    for gen in (lambda n: rng.integers(0, 2**64, n, dtype=np.uint64),
                lambda n: rng.integers(0, 50, n).astype(np.uint64)):
        keys = gen(20_000)
        kk, ii = keys.copy(), np.arange(keys.size, dtype=np.uint32)
        o.O.vo_radix_sort_u64(o.ptr(kk), o.ptr(ii), kk.size)
        ref = np.argsort(keys, kind="stable")
        assert np.array_equal(ii, ref) and np.array_equal(kk, keys[ref])
        if o.R is not None:
            k2, i2 = keys.copy(), np.arange(keys.size, dtype=np.uint32)
            o.R.vr_radix_sort_u64(o.ptr(k2), o.ptr(i2), k2.size)
            assert np.array_equal(i2, ii)


def test_mergesort_restatement(golden):
    for name in ("sort_c_test", "sort_ties"):
        g = golden[name]
        x = np.array(g["data"], np.float64)
        for desc, key in ((0, "asc"), (1, "desc")):
            idx = np.zeros(x.size, np.uint32)
            o.O.vo_mergesort_f64(o.ptr(x), x.size, desc, o.ptr(idx))
            assert idx.tolist() == g[key], (name, key)
    rng = np.random.default_rng(8)
    x = np.round(rng.standard_normal(30_000), 2)
    x[::50] = -0.0
    idx = np.zeros(x.size, np.uint32)
    o.O.vo_mergesort_f64(o.ptr(x), x.size, 0, o.ptr(idx))
    assert np.array_equal(idx, np.argsort(x, kind="stable"))
    o.O.vo_mergesort_f64(o.ptr(x), x.size, 1, o.ptr(idx))
    assert np.array_equal(idx, np.argsort(-x, kind="stable"))


def test_hash_join_restatement(golden):
    for name in ("join_small", "join_dups"):
        g = golden[name]
        b, p, pay = (np.array(g[k], np.int64) for k in ("build", "probe", "payload"))
        lo, ro = np.zeros(p.size, np.uint32), np.zeros(p.size, np.uint32)
        s = np.zeros(1, np.int64)
        m = o.O.vo_hashjoin_i64(o.ptr(b), b.size, o.ptr(pay), o.ptr(p), p.size, o.ptr(lo), o.ptr(ro), o.ptr(s))
        assert m == g["matches"] and lo[:m].tolist() == g["left"] and ro[:m].tolist() == g["right"]
        assert int(s[0]) == g["payload_sum"]
    # brute force: the first equal key on a chain is the highest build row
    rng = np.random.default_rng(6)
    b = rng.integers(0, 3000, 5000).astype(np.int64)
    p = rng.integers(0, 4000, 20000).astype(np.int64)
    pay = rng.integers(-2**62, 2**62, b.size).astype(np.int64)
    last = {}
    for i, k in enumerate(b.tolist()):
        last[k] = i
    exp = [(i, last[k]) for i, k in enumerate(p.tolist()) if k in last]
    lo, ro = np.zeros(p.size, np.uint32), np.zeros(p.size, np.uint32)
    s = np.zeros(1, np.int64)
    m = o.O.vo_hashjoin_i64(o.ptr(b), b.size, o.ptr(pay), o.ptr(p), p.size, o.ptr(lo), o.ptr(ro), o.ptr(s))
    assert list(zip(lo[:m].tolist(), ro[:m].tolist())) == exp
    tot = sum(int(pay[r]) for _, r in exp)
    assert int(s[0]) == ((tot + 2**63) % 2**64) - 2**63


def _groupby(keys, vals, signed=True):
    n = keys.size
    G = 4096
    ok, oc = np.zeros(G, np.uint32), np.zeros(G, np.uint64)
    osum, omin, omax, omean, om2 = (np.zeros(G, np.float64) for _ in range(5))
    dig = np.zeros(G * o.NDIG, np.int64)
    g = o.O.vo_groupby_f64(o.ptr(keys.view(np.uint32)), o.ptr(vals), n, 1 if signed else 0, G, o.ptr(ok),
                           o.ptr(oc), o.ptr(osum), o.ptr(omin), o.ptr(omax), o.ptr(omean), o.ptr(om2), o.ptr(dig))
    return g, ok[:g].view(np.int32 if signed else np.uint32), oc[:g], osum[:g], omin[:g], omax[:g], omean[:g], om2[:g], dig


def test_groupby_restatement(golden):
    for name in ("agg_small", "agg_c_test"):
        gg = golden[name]
        keys, vals = np.array(gg["keys"], np.int32), np.array(gg["vals"], np.float64)
        g, k, c, s, mn, mx, mean, m2, _ = _groupby(keys, vals)
        ref = sorted(gg["groups"], key=lambda d: d["key"])
        assert g == len(ref)
        for i, d in enumerate(ref):
            assert (k[i], c[i], s[i], mn[i], mx[i], mean[i], m2[i]) == (
                d["key"], d["count"], d["sum"], d["min"], d["max"], d["mean"], d["m2"])
    rng = np.random.default_rng(10)
    keys = rng.integers(1, 266, 50_000).astype(np.int32)
    vals = rng.lognormal(2.5, 0.6, keys.size)
    g, k, c, s, mn, mx, mean, m2, dig = _groupby(keys, vals)
    for i in range(g):
        sel = vals[keys == k[i]]
        assert c[i] == sel.size and mn[i] == sel.min() and mx[i] == sel.max()
        assert float(o.digits_to_fraction(dig[i * o.NDIG:(i + 1) * o.NDIG])) == math.fsum(sel.tolist())
    if o.R is not None:
        G = 512
        ok, oc = np.zeros(G, np.int32), np.zeros(G, np.uint64)
        osum, omin, omax, omean, om2 = (np.zeros(G, np.float64) for _ in range(5))
        ng = o.R.vr_hash_agg_i32_f64(o.ptr(keys), o.ptr(vals), keys.size, 128, G, o.ptr(ok), o.ptr(oc),
                                     o.ptr(osum), o.ptr(omin), o.ptr(omax), o.ptr(omean), o.ptr(om2))
        order = np.argsort(ok[:ng])
        assert np.array_equal(ok[:ng][order], k)
        assert np.array_equal(oc[:ng][order], c)
        assert np.array_equal(osum[:ng][order], s)         # same row order: bit-exact
        assert np.array_equal(omean[:ng][order], mean) and np.array_equal(om2[:ng][order], m2)


def test_mlp_oracle_pinned_by_finite_differences():
    """The CPU model of the learned-CDF network: its gradient equals central
    finite differences of its loss, and its forward equals a unit-by-unit loop."""
    from oracle import mlp
    rng = np.random.default_rng(3)
    p = rng.standard_normal(mlp.N_PARAMS) * 0.5
    x = rng.random(40)
    y = np.sort(rng.random(40))
    g = mlp.grad(p, x, y)
    h = 1e-6
    for i in list(range(0, 64, 7)) + list(range(64, 1088, 97)) + list(range(1088, mlp.N_PARAMS, 5)):
        e = np.zeros(mlp.N_PARAMS)
        e[i] = h
        fd = (mlp.loss(p + e, x, y) - mlp.loss(p - e, x, y)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
    w1, b1, w2, b2, w3, b3 = mlp.unpack(p)
    for xi, yi in zip(x[:5], mlp.forward(p, x[:5])):
        h1 = [max(w1[j] * xi + b1[j], 0.0) for j in range(32)]
        h2 = [max(sum(w2[k][j] * h1[j] for j in range(32)) + b2[k], 0.0) for k in range(32)]
        assert abs(sum(w3[k] * h2[k] for k in range(32)) + b3 - yi) < 1e-12


WINDOW_KINDS = ("delta", "sum", "mean", "stddev", "variance", "quantile")


def _window(lib_fn, kind, x, out_len, window, q, fill=-7.0):
    out = np.full(out_len, fill)
    n = lib_fn(WINDOW_KINDS.index(kind), o.ptr(x), x.size, o.ptr(out), out_len, window, q)
    return n, out


def test_window_family_closed_forms():
    """vo_window_f64 against numpy on integer-valued data (every window sum
    exact): sums = convolution, delta = diff, mean = sum / w, variance and
    stddev = the two-pass population formulas, quantile = linear
    interpolation between sorted neighbours; the reference's 0-return cases."""
    rng = np.random.default_rng(3)
    x = rng.integers(-1000, 1000, 500).astype(np.float64)
    for w in (1, 2, 5, 20, 128, 200):
        n, out = _window(o.O.vo_window_f64, "sum", x, 600, w, 0.0)
        assert n == x.size - w + 1
        assert np.array_equal(out[:n], np.convolve(x, np.ones(w), "valid"))
        assert np.all(out[n:] == -7.0)                      # outputs past N untouched
        n, out = _window(o.O.vo_window_f64, "mean", x, 600, w, 0.0)
        assert np.array_equal(out[:n], np.convolve(x, np.ones(w), "valid") / w)
    n, out = _window(o.O.vo_window_f64, "delta", x, 600, 0, 0.0)
    assert n == 499 and np.array_equal(out[:n], np.diff(x))
    win = np.lib.stride_tricks.sliding_window_view(x, 20)
    n, out = _window(o.O.vo_window_f64, "variance", x, 600, 20, 0.0)
    assert np.allclose(out[:n], win.var(axis=1), rtol=1e-13, atol=0)
    n, out = _window(o.O.vo_window_f64, "stddev", x, 600, 20, 0.0)
    assert np.allclose(out[:n], win.std(axis=1), rtol=1e-13, atol=0)
    for w, q in ((20, 0.5), (100, 0.05), (200, 0.95), (7, 1.0), (7, 0.0)):
        n, out = _window(o.O.vo_window_f64, "quantile", x, 600, w, q)
        ww = min(w, 128)
        srt = np.sort(np.lib.stride_tricks.sliding_window_view(x, w)[:, :ww], axis=1)
        pos = q * (ww - 1)
        lo = int(pos)
        hi = min(lo + 1, ww - 1)
        want = srt[:, lo] + (srt[:, hi] - srt[:, lo]) * (pos - lo)
        assert n == x.size - w + 1 and np.array_equal(out[:n], want), (w, q)
    # zero-output cases (engine.hpp:198, 246, 262, 279)
    assert _window(o.O.vo_window_f64, "sum", x, 600, 0, 0.0)[0] == 0
    assert _window(o.O.vo_window_f64, "stddev", x, 600, 1, 0.0)[0] == 0
    assert _window(o.O.vo_window_f64, "quantile", x, 600, 5, 1.5)[0] == 0
    assert _window(o.O.vo_window_f64, "sum", x[:3], 600, 4, 0.0)[0] == 0
    assert _window(o.O.vo_window_f64, "delta", x[:1], 600, 0, 0.0)[0] == 0


@pytest.mark.skipif(o.R is None, reason="reference build absent")
def test_window_family_restatement_against_reference():
    """vo_window_f64 / vo_time_window_sum_f64 against Engine<f64>'s static
    window methods: every output bit-identical, same return value."""
    rng = np.random.default_rng(4)
    for trial in range(40):
        n = int(rng.integers(1, 3000))
        x = rng.standard_normal(n) * 10.0 ** int(rng.integers(-3, 6))
        kind = WINDOW_KINDS[trial % 6]
        w = int(rng.choice([0, 1, 2, 3, 20, 100, 128, 129, 255]))
        q = float(rng.choice([0.0, 0.05, 0.5, 0.99, 1.0, -0.5, 2.0]))
        out_len = int(rng.integers(0, n + 10))
        a, oa = _window(o.O.vo_window_f64, kind, x, out_len, w, q)
        b, ob = _window(o.R.vr_window_f64, kind, x, out_len, w, q)
        assert a == b and np.array_equal(oa.view(np.uint64), ob.view(np.uint64)), (trial, kind, w, q)
    for trial in range(20):
        n = int(rng.integers(0, 5000))
        ts = np.sort(rng.integers(-500, 100000, n)).astype(np.int64)
        if trial % 3 == 0:
            rng.shuffle(ts)
        v = rng.standard_normal(n) * 100
        start = int(rng.integers(-1000, 1000))
        dur = int(rng.choice([-5, 0, 1, 7, 3600]))
        out_len = int(rng.integers(0, 200))
        o1 = rng.standard_normal(out_len)
        o2 = o1.copy()
        a = o.O.vo_time_window_sum_f64(o.ptr(ts), o.ptr(v), n, start, dur, o.ptr(o1), out_len)
        b = o.R.vr_time_window_sum_f64(o.ptr(ts), o.ptr(v), n, start, dur, o.ptr(o2), out_len)
        assert a == b and np.array_equal(o1.view(np.uint64), o2.view(np.uint64)), trial


def _random_validity(rng, n, trial):
    nw = max((n + 63) // 64, 1)
    kind = trial % 5
    if kind == 0:
        w = np.full(nw, ~np.uint64(0), np.uint64)                     # all valid
    elif kind == 1:
        w = np.zeros(nw, np.uint64)                                   # all null
    else:
        w = rng.integers(0, 2**63, nw, dtype=np.int64).astype(np.uint64) * np.uint64(2) + \
            rng.integers(0, 2, nw).astype(np.uint64)
        if kind == 3:
            w[rng.integers(0, nw, max(nw // 3, 1))] = 0               # whole null words
    return w                                                          # bits past n may be set


@pytest.mark.skipif(o.R is None, reason="reference build absent")
@pytest.mark.parametrize("is_f64", [True, False])
def test_nullable_restatement_against_reference(is_f64):
    """vo_nullable_* against NullableSegment<T> (SumValid, MinValid,
    MaxValid, ValidCount, NullCountVal, CompactValid) on the same data and
    NullBitmap words: every value bit-identical."""
    rng = np.random.default_rng(8 if is_f64 else 9)
    for trial in range(40):
        n = int(rng.choice([0, 1, 63, 64, 65, 1000, 4097]))
        if is_f64:
            x = rng.standard_normal(n) * 100
            if trial % 7 == 3 and n:
                x[rng.integers(0, n, 3)] = np.nan
                x[0] = np.nan                                         # a NaN first valid value
        else:
            x = rng.integers(-2**40, 2**40, n).astype(np.int64)
        w = _random_validity(rng, n, trial)
        dt = np.float64 if is_f64 else np.int64
        r1, r2 = np.zeros(3, dt), np.zeros(3, dt)
        c1, c2 = np.zeros(3, np.uint64), np.zeros(3, np.uint64)
        k1, k2 = np.zeros(max(n, 1), dt), np.zeros(max(n, 1), dt)
        f_o = o.O.vo_nullable_f64 if is_f64 else o.O.vo_nullable_i64
        f_r = o.R.vr_nullable_f64 if is_f64 else o.R.vr_nullable_i64
        f_o(o.ptr(x), n, o.ptr(w), o.ptr(r1), o.ptr(c1), o.ptr(k1))
        xr = x.copy()
        f_r(o.ptr(xr), n, o.ptr(w), o.ptr(r2), o.ptr(c2), o.ptr(k2))
        assert np.array_equal(c1, c2), (trial, c1, c2)
        assert np.array_equal(r1.view(np.uint64), r2.view(np.uint64)), (trial, r1, r2)
        m = int(c1[2])
        assert np.array_equal(k1[:m].view(np.uint64), k2[:m].view(np.uint64)), trial
        # closed form: the valid rows are exactly the set bits below n
        bits = np.unpackbits(w.view(np.uint8), bitorder="little")[:n].astype(bool)
        assert int(c1[0]) == int(bits.sum()) and np.array_equal(k1[:m].view(np.uint64), x[bits].view(np.uint64))


def _enc_input(rng, kind, n, trial):
    if kind == "f64":
        x = np.round(rng.standard_normal(n) * 10, int(rng.integers(0, 3)))
        if trial % 3 == 0:
            x = np.repeat(x[: (n + 7) // 8], 8)[:n].copy()                 # runs
        if n > 3 and trial % 4 == 1:
            x[1] = -0.0
            x[2] = 0.0
        return x
    x = rng.integers(-50, 50, n).astype(np.int64) * (1 if trial % 2 else 2**40)
    if trial % 3 == 0:
        x = np.repeat(x[: (n + 7) // 8], 8)[:n].copy()
    return x


@pytest.mark.skipif(o.R is None, reason="reference build absent")
@pytest.mark.parametrize("kind", ["f64", "i64"])
def test_dictionary_and_rle_restatement_against_reference(kind):
    """vo_dict_* / vo_rle_* against DictionaryEncoding<T> / RLEEncoding<T>:
    built arrays and DecodeBatch outputs (any offset and count, including
    past the end) bit-identical; decode(build(x)) == x."""
    rng = np.random.default_rng(12 if kind == "f64" else 13)
    dt = np.float64 if kind == "f64" else np.int64
    for trial in range(30):
        n = int(rng.choice([0, 1, 7, 100, 3000]))
        x = _enc_input(rng, kind, n, trial)
        m = max(n, 1)
        d1, d2, i1, i2 = np.zeros(m, dt), np.zeros(m, dt), np.zeros(m, np.uint32), np.zeros(m, np.uint32)
        g1 = getattr(o.O, "vo_dict_build_" + kind)(o.ptr(x), n, o.ptr(d1), o.ptr(i1))
        g2 = getattr(o.R, "vr_dict_build_" + kind)(o.ptr(x), n, o.ptr(d2), o.ptr(i2))
        assert g1 == g2 and np.array_equal(d1[:g1].view(np.uint64), d2[:g2].view(np.uint64)), (trial, n, d1[:g1], d2[:g2], x)
        assert np.array_equal(i1[:n], i2[:n])
        v1, v2, l1, l2 = np.zeros(m, dt), np.zeros(m, dt), np.zeros(m, np.uint32), np.zeros(m, np.uint32)
        r1 = getattr(o.O, "vo_rle_build_" + kind)(o.ptr(x), n, o.ptr(v1), o.ptr(l1))
        r2 = getattr(o.R, "vr_rle_build_" + kind)(o.ptr(x), n, o.ptr(v2), o.ptr(l2))
        assert r1 == r2 and np.array_equal(v1[:r1].view(np.uint64), v2[:r2].view(np.uint64))
        assert np.array_equal(l1[:r1], l2[:r2])
        for off, cnt in ((0, n), (n // 3, n // 2 + 5), (max(n - 2, 0), 9), (n + 4, 3)):
            o1, o2 = np.full(cnt, 7, dt), np.full(cnt, 7, dt)
            getattr(o.O, "vo_dict_decode_" + kind)(o.ptr(d1), o.ptr(i1), n, off, cnt, o.ptr(o1))
            getattr(o.R, "vr_dict_decode_" + kind)(o.ptr(d2), g2, o.ptr(i2), n, off, cnt, o.ptr(o2))
            assert np.array_equal(o1.view(np.uint64), o2.view(np.uint64)), (trial, off, cnt)
            o1, o2 = np.full(cnt, 7, dt), np.full(cnt, 7, dt)
            getattr(o.O, "vo_rle_decode_" + kind)(o.ptr(v1), o.ptr(l1), r1, off, cnt, o.ptr(o1))
            getattr(o.R, "vr_rle_decode_" + kind)(o.ptr(v2), o.ptr(l2), r2, off, cnt, o.ptr(o2))
            assert np.array_equal(o1.view(np.uint64), o2.view(np.uint64)), (trial, off, cnt)
        full = np.zeros(m, dt)
        getattr(o.O, "vo_dict_decode_" + kind)(o.ptr(d1), o.ptr(i1), n, 0, n, o.ptr(full))
        assert np.array_equal(full[:n], x)                       # value equality: -0 may decode as +0
        getattr(o.O, "vo_rle_decode_" + kind)(o.ptr(v1), o.ptr(l1), r1, 0, n, o.ptr(full))
        assert np.array_equal(full[:n], x)
        if kind == "f64" and g1:
            assert np.all(np.diff(d1[:g1]) > 0)                  # sorted, distinct


@pytest.mark.skipif(o.R is None, reason="reference build absent")
def test_delta_bitpacked_for_restatement_against_reference():
    """vo_delta_* / vo_bitpack_* / vo_for_* against DeltaEncoding<T>,
    BitPackedEncoding, FrameOfReferenceEncoding: built arrays and DecodeBatch
    outputs bit-identical; delta and bit packing lossless; FOR within half a
    scale step."""
    rng = np.random.default_rng(14)
    for trial in range(30):
        n = int(rng.choice([0, 1, 33, 1000, 5000]))
        m = max(n, 1)
        for suf, dt, lo, hi in (("i64", np.int64, -2**62, 2**62), ("i32", np.int32, -2**31, 2**31 - 1)):
            x = rng.integers(lo, hi, n).astype(dt)
            a1, a2, b2 = np.zeros(m, dt), np.zeros(m, dt), np.zeros(1, dt)
            getattr(o.O, "vo_delta_build_" + suf)(o.ptr(x), n, o.ptr(a1))
            getattr(o.R, "vr_delta_build_" + suf)(o.ptr(x), n, o.ptr(a2), o.ptr(b2))
            assert b2[0] == 0 and np.array_equal(a1[:n], a2[:n])
            base = int(rng.integers(-1000, 1000))
            for off, cnt in ((0, n), (n // 2, n), (n + 3, 2)):
                o1, o2 = np.full(m + cnt, 5, dt), np.full(m + cnt, 5, dt)
                getattr(o.O, "vo_delta_decode_" + suf)(base, o.ptr(a1), n, off, cnt, o.ptr(o1))
                getattr(o.R, "vr_delta_decode_" + suf)(base, o.ptr(a2), n, off, cnt, o.ptr(o2))
                assert np.array_equal(o1, o2), (suf, trial, off)
            full = np.zeros(m, dt)
            getattr(o.O, "vo_delta_decode_" + suf)(0, o.ptr(a1), n, 0, n, o.ptr(full))
            assert np.array_equal(full[:n], x)
        u = rng.integers(0, 2 ** int(rng.integers(0, 33)), n, dtype=np.uint64).astype(np.uint32)
        w1, w2 = np.zeros(m, np.uint64), np.zeros(m, np.uint64)
        bw1, bw2 = np.zeros(1, np.uint32), np.zeros(1, np.uint32)
        k1 = o.O.vo_bitpack_build(o.ptr(u), n, o.ptr(w1), o.ptr(bw1))
        k2 = o.R.vr_bitpack_build(o.ptr(u), n, o.ptr(w2), o.ptr(bw2))
        assert k1 == k2 and bw1[0] == bw2[0] and np.array_equal(w1[:k1], w2[:k2])
        for off, cnt in ((0, n), (n // 3, n), (n + 1, 4)):
            o1, o2 = np.full(m + cnt, 9, np.uint32), np.full(m + cnt, 9, np.uint32)
            o.O.vo_bitpack_decode(o.ptr(w1), k1, int(bw1[0]), n, off, cnt, o.ptr(o1))
            o.R.vr_bitpack_decode(o.ptr(w2), k2, int(bw2[0]), n, off, cnt, o.ptr(o2))
            assert np.array_equal(o1, o2)
        full = np.zeros(m, np.uint32)
        o.O.vo_bitpack_decode(o.ptr(w1), k1, int(bw1[0]), n, 0, n, o.ptr(full))
        assert np.array_equal(full[:n], u)
        f = rng.standard_normal(n) * 1000 + float(rng.choice([0.0, 5e6]))
        if trial % 5 == 0 and n:
            f[:] = 3.25                                         # range 0: Scale 1
        e1, e2 = np.zeros(m, np.int32), np.zeros(m, np.int32)
        s1, s2 = np.zeros(2), np.zeros(2)
        o.O.vo_for_build(o.ptr(f), n, o.ptr(s1[0:]), o.ptr(s1[1:]), o.ptr(e1))
        o.R.vr_for_build(o.ptr(f), n, o.ptr(s2[0:]), o.ptr(s2[1:]), o.ptr(e2))
        assert np.array_equal(s1.view(np.uint64), s2.view(np.uint64)) and np.array_equal(e1[:n], e2[:n])
        for off, cnt in ((0, n), (n // 4, n), (n + 2, 3)):
            o1, o2 = np.full(m + cnt, 1.5), np.full(m + cnt, 1.5)
            o.O.vo_for_decode(s1[0], s1[1], o.ptr(e1), n, off, cnt, o.ptr(o1))
            o.R.vr_for_decode(s2[0], s2[1], o.ptr(e2), n, off, cnt, o.ptr(o2))
            assert np.array_equal(o1.view(np.uint64), o2.view(np.uint64))
        if n:
            dec = np.zeros(n)
            o.O.vo_for_decode(s1[0], s1[1], o.ptr(e1), n, 0, n, o.ptr(dec))
            assert np.all(np.abs(dec - f) <= s1[1] * 0.5 + 1e-9 * np.abs(f).max())


@pytest.mark.skipif(o.R is None, reason="reference build absent")
def test_tier0_window_restatement_against_reference_engine():
    """vo_tier0_window_f64 (loops loading WINDOW_SUM / WINDOW_MEAN / WSTD /
    WVARIANCE_W / WDELTA / WQUANTILE, vector words, up to three reductions) against the
    reference engine running the same words: every accumulator and partial
    bit-identical."""
    from tests import tier0_gen as tg
    rng = np.random.default_rng(93)
    for trial in range(96):
        wkind = 1 + trial % 6
        n = int(rng.integers(6, 400))
        window = int(rng.choice([1, 2, 3, 7, 20, 64])) if wkind != 5 else 0
        if wkind in (3, 4) and window < 2:
            window = 2
        if wkind == 6:
            window = tg.wquantile_imm(rng)
        x = np.round(rng.standard_normal(n) * 100, 2) + 0.0      # no -0.0: a +-0 tie's order is unspecified
        regs = np.zeros(64, np.uint64)
        for r in tg.SCAL_REGS:
            regs[r] = o.f64_bits(float(np.round(rng.standard_normal() * 10, 1)))
        total = tg.window_total(rng, n, wkind)
        regs[2] = total
        code, spec = tg.make_window_program(rng, wkind, window, nops=int(rng.integers(0, 4)))
        ok, acc, part = tg.oracle_window_loop(o, x, total, spec, regs, wkind, window)
        assert ok, (trial, wkind, n, total)
        ref = regs.copy()
        o.R.vr_engine_run_f64(o.ptr(x), n, o.ptr(code), code.size, o.ptr(ref))
        assert int(ref[1]) == total
        for b in range(len(spec["red"])):
            assert int(ref[tg.ACCS[b]]) == acc[b], (trial, wkind, b, spec)
            assert int(ref[tg.PARTS[b]]) == part[b], (trial, wkind, b, spec)
    # the offset starts at (or past) the end the word steps to: it writes zero
    # lanes and advances by 0, so one pass runs and CMP falls through
    for trial in range(24):
        wkind = 1 + trial % 6
        n = int(rng.integers(6, 100))
        window = {5: 0, 6: tg.wquantile_imm(rng)}.get(wkind, int(rng.choice([2, 3, 7])))
        x = np.round(rng.standard_normal(n) * 100, 2) + 0.0
        regs = np.zeros(64, np.uint64)
        for r in tg.SCAL_REGS:
            regs[r] = o.f64_bits(float(np.round(rng.standard_normal() * 10, 1)))
        start = (n - 1 if wkind == 5 else n) + int(rng.integers(0, 3))
        regs[1] = regs[2] = start
        code, spec = tg.make_window_program(rng, wkind, window, nops=int(rng.integers(0, 4)))
        acc0 = [int(regs[a]) for a in tg.ACCS[:len(spec["red"])]]
        ok, acc, part = tg.oracle_window_loop(o, x, start, spec, regs, wkind, window, acc0=acc0, off=start)
        assert ok, (trial, wkind, n, start)
        ref = regs.copy()
        o.R.vr_engine_run_f64(o.ptr(x), n, o.ptr(code), code.size, o.ptr(ref))
        assert int(ref[1]) == start
        for b in range(len(spec["red"])):
            assert int(ref[tg.ACCS[b]]) == acc[b], (trial, wkind, b, spec)
            assert int(ref[tg.PARTS[b]]) == part[b], (trial, wkind, b, spec)


def _brute_join(kind, l, r):
    """Counts by Python == over floats / ints (NaN never equal, -0.0 == 0.0)."""
    if kind == 0:
        return sum(sum(1 for b in r if a == b) for a in l)
    if kind == 1:
        from collections import Counter
        cl = Counter(float(a) if isinstance(a, float) else a for a in l if a == a)
        cr = Counter(float(b) if isinstance(b, float) else b for b in r if b == b)
        return sum(min(c, cr.get(v, 0)) for v, c in cl.items())
    found = [any(a == b for b in r) for a in l]
    return sum(found) if kind == 2 else len(l) - sum(found)


def test_join_count_and_partial_sort_closed_forms():
    """vo_join_count against brute-force == loops (pairs, merge, semi, anti;
    NaN never equal, -0.0 == +0.0); vo_partial_sort: the first k rows are
    the k largest (smallest) in order, the whole stays a permutation."""
    rng = np.random.default_rng(31)
    for trial in range(60):
        is_f64 = trial % 2 == 0
        nl, nr = int(rng.integers(0, 60)), int(rng.integers(0, 60))
        if is_f64:
            l = np.round(rng.standard_normal(nl) * 3)
            r = np.round(rng.standard_normal(nr) * 3)
            if trial % 4 == 0 and nl and nr:
                l[0], r[-1] = np.nan, -0.0
        else:
            l = rng.integers(-5, 6, nl).astype(np.int64)
            r = rng.integers(-5, 6, nr).astype(np.int64)
        lb, rb = l.view(np.uint64), r.view(np.uint64)
        for kind in range(4):
            got = o.O.vo_join_count(kind, o.ptr(lb), nl, o.ptr(rb), nr, int(is_f64))
            if kind == 1 and is_f64 and (np.isnan(l).any() or np.isnan(r).any()):
                assert got == -1
                continue
            assert got == _brute_join(kind, l.tolist(), r.tolist()), (trial, kind)
        if nl:
            k = int(rng.integers(1, nl + 1))
            for desc in (1, 0):
                a = l.copy()
                if is_f64:
                    a = a[~np.isnan(a)] + 0.0
                if a.size < k:
                    continue
                b = a.view(np.uint64).copy()
                o.O.vo_partial_sort(o.ptr(b), k, b.size, int(is_f64), desc)
                v = b.view(a.dtype)
                want = np.sort(a)[::-1][:k] if desc else np.sort(a)[:k]
                assert np.array_equal(v[:k], want) and np.array_equal(np.sort(v), np.sort(a))


@pytest.mark.skipif(o.R is None, reason="reference build absent")
@pytest.mark.parametrize("is_f64", [True, False])
def test_hash_sort_join_words_restatement_against_reference_engine(is_f64):
    """vo_relw_run (HASH_INIT/BUILD/PROBE/LOOKUP, SORT_ASC/DESC/TOPK/BOTTOMK,
    JOIN_HASH/MERGE/NESTED/ANTI/SEMI, PARTITION_HASH in straight-line
    programs over up to three segments) against the reference engine
    running the same words: every register and every segment row after the
    run bit-identical.  Inputs hold ties, NaN and -0.0 except where the
    reference's result is unspecified (std::sort's order of -0.0 and +0.0,
    and any order with NaN, in SORT_ASC/DESC and JOIN_MERGE); partial sorts
    and hash tables see them all."""
    from tests import tier0_gen as tg
    rng = np.random.default_rng(57 if is_f64 else 58)
    runs = 0
    for trial in range(120):
        nseg = int(rng.integers(1, 4))
        full_order = trial % 3 != 0            # programs with SORT_ASC/DESC / JOIN_MERGE
        segs = []
        for s in range(nseg):
            n = int(rng.integers(1, 300))
            if is_f64:
                x = np.round(rng.standard_normal(n) * 4, int(rng.integers(0, 2)))
                if full_order:
                    x = x + 0.0                      # no -0.0
                else:
                    x[rng.integers(0, n, 2)] = np.nan
            else:
                x = rng.integers(-20, 20, n).astype(np.int64)
                x[rng.integers(0, n, 2)] = rng.integers(-2**62, 2**62, 2)
            segs.append(x)
        ops = tg.REL_OPS if full_order else [op for op in tg.REL_OPS if op not in (0xF2, 0xF3, 0xF7)]
        code = tg.rel_program(rng, nseg, k=int(rng.integers(1, 12)), ops=ops,
                              kmax=min(a.size for a in segs))
        regs = np.zeros(64, np.uint64)
        for r in tg.KEY_REGS:
            src = segs[int(rng.integers(0, nseg))]
            regs[r] = src.view(np.uint64)[int(rng.integers(0, src.size))] if rng.random() < 0.8 else \
                np.uint64(rng.integers(0, 2**63))
        for r in tg.K_REGS:
            regs[r] = int(rng.integers(0, 5000)) if rng.random() < 0.8 else 2**64 - 1
        mine = [a.view(np.uint64).copy() for a in segs]
        p, n = o.seg_ptrs(mine)
        st = o.O.vo_relw_create()
        ro = regs.copy()
        rc = o.O.vo_relw_run(st, o.ptr(p), o.ptr(n), nseg, int(is_f64), o.ptr(code), code.size, o.ptr(ro))
        o.O.vo_relw_destroy(st)
        assert rc == 1, (trial, [hex(int(c)) for c in code])
        theirs = [a.copy() for a in segs]
        pr, nr = o.seg_ptrs(theirs)
        rr = regs.copy()
        (o.R.vr_engine_run_segs_f64 if is_f64 else o.R.vr_engine_run_segs_i64)(
            o.ptr(pr), o.ptr(nr), nseg, o.ptr(code), code.size, o.ptr(rr))
        assert np.array_equal(ro, rr), (trial, [hex(int(c)) for c in code], np.nonzero(ro != rr))
        for a, b in zip(mine, theirs):
            assert np.array_equal(a, b.view(np.uint64)), (trial, [hex(int(c)) for c in code])
        runs += 1
    assert runs == 120


@pytest.mark.skipif(o.R is None, reason="reference build absent")
@pytest.mark.parametrize("is_f64", [True, False])
def test_sum_distinct_and_mode_restatement_against_reference_engine(is_f64):
    """vo_straight_run with AGG_SUM_DISTINCT (sharing the distinct set of
    AGG_COUNT_DISTINCT's slot) and AGG_MODE, against the reference engine on
    the same words: every register bit-identical except an f64
    SUM_DISTINCT, which agrees within 1e-12 of the sum of the distinct |x|
    (the reference adds the set in unordered_map order, the restatement in
    ascending order).  Programs whose AGG_MODE has a tied largest count are
    skipped (the restatement refuses them); data with repeated values keeps
    most counts unique."""
    from tests import tier0_gen as tg
    rng = np.random.default_rng(53 if is_f64 else 54)
    done = modes = sums = 0
    for trial in range(60):
        n = int(rng.integers(1, 3000))
        if is_f64:
            x = np.round(rng.standard_normal(n) * 4, 1)          # repeated values
            x[rng.integers(0, n, 3)] = -0.0
            x[rng.integers(0, n, 3)] = 0.0
        else:
            x = rng.integers(-40, 40, n).astype(np.int64) * 1_000_003
        st = o.O.vo_agg_create(1 if is_f64 else 0)
        regs = np.zeros(64, np.uint64)
        ref_regs = np.zeros(64, np.uint64)
        code = tg.straight_program(rng, n, k=10, ops=[0xE9, 0xEA, 0xEA, 0xEC, 0xEC, 0xE1],
                                   scalar_reads=(1, 2, 3, 13, 14, 15))
        rc = o.O.vo_straight_run(st, o.ptr(x.view(np.uint64)), n, o.ptr(code), code.size, o.ptr(regs))
        o.O.vo_agg_destroy(st)
        if rc < 0:
            continue
        run = o.R.vr_engine_run_f64 if is_f64 else o.R.vr_engine_run_i64
        run(o.ptr(x), n, o.ptr(code), code.size, o.ptr(ref_regs))
        writers = tg.last_agg_writers(code)
        scale = float(np.abs(np.unique(x)).sum()) if is_f64 else 0.0
        for r in range(16):
            if regs[r] == ref_regs[r]:
                continue
            assert is_f64 and writers.get(r) in (0xEA, 0xE1), (trial, r, [hex(int(v)) for v in code])
            assert abs(o.bits_f64(int(regs[r])) - o.bits_f64(int(ref_regs[r]))) <= 1e-12 * max(scale, 1.0)
        modes += sum(1 for c in code if int(c) & 0xFF == 0xEC)
        sums += sum(1 for c in code if int(c) & 0xFF == 0xEA)
        done += 1
    assert done >= 30 and modes >= 30 and sums >= 30, (done, modes, sums)


def test_mode_restatement_closed_forms():
    """AGG_MODE on hand-made columns: the unique most frequent value; -0.0 and
    +0.0 are one key whose value is the first row's; NaN rows are entries of
    their own; a tied largest count is refused; an empty range gives 0."""
    def mode(x, start=0, cnt=0):
        x = np.asarray(x, np.float64)
        st = o.O.vo_agg_create(1)
        regs = np.zeros(64, np.uint64)
        regs[1] = start
        code = np.array([0xEC | (4 << 8) | (0 << 12) | (1 << 16) | (cnt << 20), 1], np.uint32)
        rc = o.O.vo_straight_run(st, o.ptr(x.view(np.uint64)), x.size, o.ptr(code), code.size, o.ptr(regs))
        o.O.vo_agg_destroy(st)
        return rc, int(regs[4])
    assert mode([3.0, 1.0, 3.0, 2.0]) == (1, o.f64_bits(3.0))
    assert mode([-0.0, 5.0, 0.0]) == (1, o.f64_bits(-0.0))
    assert mode([0.0, -0.0, 5.0]) == (1, o.f64_bits(0.0))
    assert mode([np.nan, np.nan, 7.0, 7.0])[1] == o.f64_bits(7.0)
    assert mode([np.nan, 7.0])[0] == -1
    assert mode([1.0, 2.0])[0] == -1
    assert mode([1.0, 2.0, 2.0], start=3) == (1, 0)


@pytest.mark.skipif(o.R is None, reason="reference build absent")
@pytest.mark.parametrize("is_f64", [True, False])
def test_tier0_lane_words_restatement_against_reference_engine(is_f64):
    """vo_tier0_expr_* with the lane words (VINSERT, VPERMUTE, VSHUFFLE,
    VREVERSE, VROTATE, VSLIDE, VBLEND on masks the body's filters and
    compares wrote, VPOW, VGATHER from a second segment or a missing one)
    against the reference engine (two segments) running the same words:
    every accumulator and partial bit-identical.  After the random
    programs, each lane word is the last word of three programs, reduced
    first."""
    from tests import tier0_gen as tg
    rng = np.random.default_rng(95 if is_f64 else 96)
    forced = tg.LANE_OPS
    done = 0
    for trial in range(80 + 3 * len(forced)):
        last = forced[(trial - 80) // 3] if trial >= 80 else None
        n = int(rng.integers(1, 200)) * 4
        m = int(rng.integers(1, 400))
        regs = np.zeros(64, np.uint64)
        if is_f64:
            x = np.round(rng.standard_normal(n) * 100, 2)
            x[rng.integers(0, n, 6)] = 0.0
            g = np.round(rng.standard_normal(m) * 10, 3)
            for r in tg.SCAL_REGS:
                regs[r] = o.f64_bits(float(np.round(rng.standard_normal() * 10, 1)))
        else:
            x = rng.integers(-20, 600, n).astype(np.int64)          # small values: in-range gather indices
            g = rng.integers(-1000, 1000, m).astype(np.int64)
            for r in tg.SCAL_REGS:
                regs[r] = int(rng.integers(-5, 50)) & (2**64 - 1)
        regs[tg.SLIDE_REG] = int(rng.choice([0, 1, 2, 3, 5, -1, -2])) & (2**64 - 1)
        regs[2] = n
        code, spec = tg.make_expr_program(rng, is_f64, nops=int(rng.integers(1, 7)), last=last, lane_words=True)
        if not is_f64 and tg.expr_traps(o, x, n, spec, regs):
            continue
        o.O.vo_expr_reset_masks()
        o.O.vo_expr_set_gather(1, o.ptr(g), m)
        acc, part = tg.oracle_expr(o, x, n, spec, regs, is_f64)
        ref = regs.copy()
        ptrs = (C.c_void_p * 2)(x.ctypes.data, g.ctypes.data)
        ns = (C.c_size_t * 2)(n, m)
        run = o.R.vr_engine_run_segs_f64 if is_f64 else o.R.vr_engine_run_segs_i64
        run(ptrs, ns, 2, o.ptr(code), code.size, o.ptr(ref))
        for b in range(len(spec["red"])):
            # VPOW runs the host's libm pow on both sides: bit-identical too
            assert int(ref[tg.ACCS[b]]) == acc[b], (trial, b, spec)
            assert int(ref[tg.PARTS[b]]) == part[b], (trial, b, spec)
        done += 1
    assert done >= 60


@pytest.mark.skipif(o.R is None, reason="reference build absent")
@pytest.mark.parametrize("is_f64", [True, False])
def test_tier0_vstore_restatement_against_reference_engine(is_f64):
    """vo_tier0_expr_* with VSTOREs at the loop offset (into a third segment,
    or into the loaded column itself, lane counts 1..4 and more) after
    vector arithmetic, filters and lane words, with or without reductions,
    against the reference engine on three segments: the stored segments
    and every accumulator and partial bit-identical."""
    from tests import tier0_gen as tg
    rng = np.random.default_rng(97 if is_f64 else 98)
    done = 0
    for trial in range(80):
        n = int(rng.integers(1, 150)) * 4
        m = int(rng.integers(1, 300))
        ny = int(rng.integers(1, n + 9))
        regs = np.zeros(64, np.uint64)
        if is_f64:
            x = np.round(rng.standard_normal(n) * 100, 2)
            g = np.round(rng.standard_normal(m) * 10, 3)
            y = np.round(rng.standard_normal(ny), 1)
            for r in tg.SCAL_REGS:
                regs[r] = o.f64_bits(float(np.round(rng.standard_normal() * 10, 1)))
        else:
            x = rng.integers(-20, 600, n).astype(np.int64)
            g = rng.integers(-1000, 1000, m).astype(np.int64)
            y = rng.integers(-9, 9, ny).astype(np.int64)
            for r in tg.SCAL_REGS:
                regs[r] = int(rng.integers(-5, 50)) & (2**64 - 1)
        regs[tg.SLIDE_REG] = int(rng.choice([0, 1, 2, -1])) & (2**64 - 1)
        regs[2] = n
        segs = (2,) if trial % 3 else (2, 0)
        code, spec = tg.make_expr_program(rng, is_f64, nops=int(rng.integers(1, 6)), lane_words=trial % 2 == 0,
                                          store_segs=segs, lane_ops=[op for op in tg.LANE_OPS if op != 0x8F])
        if not is_f64 and tg.expr_traps(o, x, n, spec, regs):
            continue
        xo, yo = x.copy(), y.copy()
        o.O.vo_expr_reset_masks()
        o.O.vo_expr_set_gather(1, o.ptr(g), m)
        acc = part = []
        store_ids = {op[4] >> 8 for op in spec["ops"] if op[0] == 0x71}
        if len(store_ids) > 1:
            continue                                           # the restatement holds one store segment
        sid = store_ids.pop()
        o.O.vo_expr_set_store(sid, o.ptr(xo if sid == 0 else yo), n if sid == 0 else ny)
        if spec["red"]:
            acc, part = tg.oracle_expr(o, xo, n, spec, regs, is_f64)
        else:
            tg.oracle_expr(o, xo, n, spec, regs, is_f64)
        xr, yr = x.copy(), y.copy()
        ref = regs.copy()
        ptrs = (C.c_void_p * 3)(xr.ctypes.data, g.ctypes.data, yr.ctypes.data)
        ns = (C.c_size_t * 3)(n, m, ny)
        run = o.R.vr_engine_run_segs_f64 if is_f64 else o.R.vr_engine_run_segs_i64
        run(ptrs, ns, 3, o.ptr(code), code.size, o.ptr(ref))
        assert np.array_equal(yo.view(np.uint64), yr.view(np.uint64)), (trial, spec)
        assert np.array_equal(xo.view(np.uint64), xr.view(np.uint64)), (trial, spec)
        for b in range(len(spec["red"])):
            assert int(ref[tg.ACCS[b]]) == acc[b], (trial, b, spec)
            assert int(ref[tg.PARTS[b]]) == part[b], (trial, b, spec)
        done += 1
    assert done >= 50, done


@pytest.mark.skipif(o.R is None, reason="reference build absent")
def test_reference_compound_path_matches_relational_restatement():
    """The reference's compound-predicate operators (VectorFilter ApplyGT /
    ApplyLT, ColumnFilter AndWith, SelectionVector, then a row-order loop;
    the CPU baseline of the two-column scan) give the oracle's relational
    count, min and max exactly and its exact sum within the row-order sum's
    rounding."""
    rng = np.random.default_rng(12)
    for n in (0, 1, 63, 64, 65, 10_007):
        a = rng.random(n) * 1000.0
        b = rng.random(n) * 1000.0
        out = np.zeros(4)
        o.R.vr_compound_f64(o.ptr(a), o.ptr(b), n, 292.8932188, 707.1067812, o.ptr(out))
        r = o.relational(a, 4, 292.8932188, b, 2, 707.1067812)
        assert int(out[0]) == r["count"]
        if r["count"]:
            assert out[2] == r["min"] and out[3] == r["max"]
            assert abs(out[1] - float(r["sum"])) <= 1e-12 * float(r["sum"])

"""GPU parity of the scan family (vx_scan.hip) through the C ABI.

Bars (stated per test):
* counts, integer sums, bit-pattern min/max, register effects: bit-exact with
  the oracle / the reference's golden values;
* f64 sums: |gpu - exact| <= 2 ulp(exact) against the oracle's exact long
  accumulator, and within the reference's own 1e-12 relative criterion
  (examples/validate.cpp:68) of its result;
* f32 sums: |gpu - exact| <= 4 ulp_f32(exact) (compensated in f32).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as o

pytestmark = pytest.mark.gpu
B = lambda h: int(h, 16)  # noqa: E731


def close_f64(got: float, exact, ulps: float = 2.0) -> bool:
    e = float(exact)
    return abs(got - e) <= ulps * o.ulp(e) + 0.0


def words(golden, name):
    return [int(w, 16) for w in golden["programs"][name]]


def engine_f64(x, code, regs):
    from voxel_amd import voxel_py as vx
    e = vx.EngineF64()
    e.add_segment(x)
    for k, v in regs.items():
        e.set_scalar(k, v)
    e.load_program(code)
    return e


# ---------------------------------------------------------------- fast path
def test_fast_path_f64_golden(gpu, golden, x42):
    g = golden["seed42"]
    e = engine_f64(x42, words(golden, "fast"), {2: x42.size, 3: o.f64_bits(500.0)})
    e.run()
    exact = o.relational(x42, 4, 500.0)["sum"]
    got = e.get_scalar_f64(0)
    ref = o.bits_f64(B(g["fast"]["R0"]))
    assert close_f64(got, exact), (got, float(exact))
    assert abs(got - ref) <= 1e-12 * abs(ref)
    assert e.get_scalar(1) == g["fast"]["R1"] == 1_000_000
    print("fast path f64: gpu err %.3g ulp, reference err %.3g ulp" % (
        abs(got - float(exact)) / o.ulp(float(exact)), abs(ref - float(exact)) / o.ulp(float(exact))))


def test_fast_path_tail_and_offsets(gpu, golden, x42):
    t = x42[:1003].copy()
    e = engine_f64(t, words(golden, "fast"), {2: 1003, 3: o.f64_bits(0.0)})
    e.run()
    assert e.get_scalar(1) == 1000                                   # tail of 3 rows dropped
    assert close_f64(e.get_scalar_f64(0), o.exact_sum(t[:1000]))
    # nonzero start offset, initial accumulator, R[total] beyond Count
    e = engine_f64(t, words(golden, "fast"), {0: o.f64_bits(1.5), 1: 5, 2: 10**9, 3: o.f64_bits(500.0)})
    e.run()
    acc = np.zeros(1, np.float64)
    end = o.O.vo_fastpath_f64(o.ptr(t), t.size, 10**9, 5, 500.0, 1.5, 0xC5, o.ptr(acc))
    assert e.get_scalar(1) == end == 1001
    sel = t[5:1001]
    assert close_f64(e.get_scalar_f64(0), o.exact_sum(np.concatenate([[1.5], sel[sel > 500]])))
    # fewer than kL rows: no group runs, registers keep their values
    e = engine_f64(t[:3].copy(), words(golden, "fast"), {0: o.f64_bits(-0.0), 2: 3, 3: o.f64_bits(0.0)})
    e.run()
    assert e.get_scalar(0) == o.f64_bits(-0.0) and e.get_scalar(1) == 0


def test_fast_path_every_filter_opcode(gpu, golden):
    rng = np.random.default_rng(1)
    x = np.round(rng.standard_normal(4096) * 10)                     # many exact ties
    x[::37] = np.nan
    base = words(golden, "fast")
    for fop in range(0xC1, 0xC7):
        code = list(base)
        code[1] = (int(code[1]) & 0xFFFFFF00) | fop
        e = engine_f64(x, code, {2: x.size, 3: o.f64_bits(3.0)})
        e.run()
        acc = np.zeros(1, np.float64)
        o.O.vo_fastpath_f64(o.ptr(x), x.size, x.size, 0, 3.0, 0.0, fop, o.ptr(acc))
        # oracle sums in the reference order; both are NaN for NE (NaN != 3 passes)
        got, ref = e.get_scalar_f64(0), float(acc[0])
        if math.isnan(ref):
            assert math.isnan(got)
        else:
            m = {0xC1: x == 3, 0xC2: x != 3, 0xC3: x < 3, 0xC4: x <= 3, 0xC5: x > 3, 0xC6: x >= 3}[fop]
            assert got == float(o.exact_sum(x[m]))                   # small integers: exact


def test_fast_path_i32_wraps_bit_exact(gpu, golden):
    from voxel_amd._lib import lib
    g = golden["i32_seed9"]
    x = o.mt_uniform_i32(9, g["n"], 0, 1 << 30)
    e = lib.voxel_gpu_engine_create(3)
    lib.voxel_gpu_engine_add_segment_raw(e, x.ctypes.data, x.size)
    lib.voxel_engine_set_scalar(e, 2, x.size)
    lib.voxel_engine_set_scalar(e, 3, 1 << 29)
    c = np.array(words(golden, "fast"), np.uint32)
    lib.voxel_engine_load_program(e, c.ctypes.data, c.size)
    lib.voxel_engine_run(e)
    assert lib.voxel_last_error() in (b"", None)
    assert lib.voxel_engine_get_scalar(e, 0) == B(g["R0"])
    lib.voxel_engine_destroy(e)


def test_literal_programs(gpu, golden):
    d8 = np.arange(1, 9, dtype=np.float64)
    e = engine_f64(d8, words(golden, "fast"), {2: 8, 3: o.f64_bits(4.0)})      # c_test.cpp Test 1
    e.run()
    assert e.get_scalar_f64(0) == 26.0
    e = engine_f64(d8, [0x00400410] + words(golden, "fast"), {2: 8, 3: o.f64_bits(4.0)})
    e.run()                                                           # test_suite.cpp:254-285
    assert e.get_scalar_f64(0) == 26.0 and e.get_scalar(4) == 4
    from voxel_amd import voxel_py as vx
    ei = vx.EngineI64()
    ei.add_segment(np.array([10, 20, 30, 40], np.int64))
    ei.set_scalar(2, 4)
    ei.set_scalar(3, 15)
    ei.load_program(words(golden, "i64_add"))
    ei.run()
    assert ei.get_scalar_i64(0) == 90                                 # test_suite.cpp:290-321


# ---------------------------------------------------------------- Tier-0
def test_voxel_query_shapes_golden(gpu, golden, x42):
    g = golden["seed42"]
    from voxel_amd import voxel_query as vq
    exact = o.relational(x42, 4, 500.0)["sum"]
    s = vq.from_numpy(x42).filter_gt(500.0).sum().run()
    assert close_f64(s, exact)
    assert abs(s - o.bits_f64(B(g["vq_sum"]["R0"]))) <= 1e-12 * s
    c = vq.from_numpy(x42).filter_gt(500.0).count().run()
    assert o.f64_bits(c) == B(g["vq_count"]["R0"]) == 500_720         # count bits read as f64 (Q9)
    mn = vq.from_numpy(x42).filter_gt(500.0).min().run()
    assert o.f64_bits(mn) == B(g["vq_min"]["R0"]) == 0                # Q10
    mx = vq.from_numpy(x42).filter_gt(500.0).max().run()
    assert o.f64_bits(mx) == B(g["vq_max"]["R0"])
    lt = vq.from_numpy(x42).filter_lt(500.0).sum().run()
    assert close_f64(lt, o.relational(x42, 2, 500.0)["sum"])
    assert abs(lt - o.bits_f64(B(g["vq_sum_lt"]["R0"]))) <= 1e-12 * lt


def test_tier0_register_effects(gpu, golden, x42):
    g = golden["seed42"]
    for name in ("vq_sum", "vq_count", "vq_min", "vq_max"):
        e = engine_f64(x42, words(golden, name), {2: x42.size, 3: o.f64_bits(500.0)})
        e.run()
        assert e.get_scalar(1) == g[name]["R1"]
        assert e.get_scalar(4) == g[name]["R4"]
        assert e.get_scalar(5) == B(g[name]["R5"]), name               # last group's partial, bit-exact
        if name != "vq_sum":
            assert e.get_scalar(0) == B(g[name]["R0"]), name


def test_tier0_i64_add_variant_golden(gpu, golden):
    g = golden["i64_seed7"]
    from voxel_amd import voxel_py as vx
    x = o.mt_uniform_i64(7, g["n"], 0, 1000)
    e = vx.EngineI64()
    e.add_segment(x)
    e.set_scalar(2, x.size)
    e.set_scalar(3, 500)
    e.load_program(words(golden, "i64_add"))
    e.run()
    assert e.get_scalar_i64(0) == g["R0"] == 374_259_792


@pytest.mark.parametrize("seed", range(6))
def test_tier0_random_programs_against_oracle(gpu, golden, seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(1, 3000)) * 4
    x = rng.standard_normal(n) * 100
    x[rng.integers(0, n, 7)] = 0.0
    if seed % 2:
        x[rng.integers(0, n, 4)] = np.nan
    thr = float(rng.standard_normal() * 50)
    shapes = {"vq_count": (1, 1), "vq_min": (2, 2), "vq_max": (3, 3), "vq_sum": (0, 0)}
    for name, (red, accop) in shapes.items():
        e = engine_f64(x, words(golden, name), {2: n, 3: o.f64_bits(thr)})
        e.run()
        part = np.zeros(1, np.uint64)
        r = o.O.vo_tier0_f64(o.ptr(x), n, 0, n, 4, thr, 0, red, accop, 0, o.ptr(part))
        assert e.get_scalar(5) == int(part[0]), name
        if name == "vq_sum":
            got = e.get_scalar_f64(0)
            if math.isnan(o.bits_f64(int(r))):
                assert math.isnan(got)
            else:
                sel = np.where(x > thr, x, 0.0)
                assert close_f64(got, o.exact_sum(sel[~np.isnan(sel)]))
        else:
            assert e.get_scalar(0) == int(r), name


def _block_vector(x, spec, thr, b):
    """The reduced vector of block b: filters of its chain applied in order."""
    ops = {0: np.equal, 1: np.not_equal, 2: np.less, 3: np.less_equal, 4: np.greater, 5: np.greater_equal}
    chain = []
    j = spec["bsrc"][b]
    while j >= 0:
        chain.append(j)
        j = spec["fsrc"][j]
    v = x.copy()
    with np.errstate(invalid="ignore"):
        for j in reversed(chain):
            v = np.where(ops[spec["fcmp"][j]](v, thr[j]), v, 0)
    return v


@pytest.mark.parametrize("is_f64", [True, False])
def test_tier0_filter_chains_and_several_reductions(gpu, is_f64):
    """Loops with up to two chained VFILTERs and up to three reductions
    (generated words): every partial register, count, min and max equal the
    oracle's bits (vo_tier0_multi_*, pinned against the reference engine);
    f64 sums within 2 ulp of the exact sum of the block's vector, i64 sums
    exact.  A chain joining != with another filter is refused."""
    from tests import tier0_gen as tg
    from voxel_amd import voxel_py as vx
    rng = np.random.default_rng(61 if is_f64 else 62)
    refused = 0
    for trial in range(24):
        n = int(rng.integers(1, 2000)) * 4
        if is_f64:
            x = rng.standard_normal(n) * 100
            x[rng.integers(0, n, 7)] = 0.0
            if trial % 3 == 0:
                x[rng.integers(0, n, 3)] = np.nan
            thr = [float(rng.standard_normal() * 60) for _ in range(2)]
            tb = [o.f64_bits(t) for t in thr]
            e = vx.EngineF64()
        else:
            x = rng.integers(-2**40, 2**40, n).astype(np.int64)
            x[rng.integers(0, n, 7)] = 0
            thr = [int(rng.integers(-2**39, 2**39)) for _ in range(2)]
            tb = [t & (2**64 - 1) for t in thr]
            e = vx.EngineI64()
        code, spec = tg.make_program(rng, is_f64, allow_ne=True)
        e.add_segment(x)
        for r, v in ((2, n), (3, tb[0]), (6, tb[1])):
            e.set_scalar(r, v)
        e.load_program([int(c) for c in code])
        if not tg.mergeable(spec):
            with pytest.raises(RuntimeError, match="filter chain"):
                e.run()
            refused += 1
            continue
        e.run()
        acc, part = tg.oracle_run(o, x, n, spec, thr, is_f64)
        assert e.get_scalar(1) == n
        for b, red in enumerate(spec["red"]):
            assert e.get_scalar(tg.PARTS[b]) == part[b], (trial, b, spec)
            got = e.get_scalar(tg.ACCS[b])
            if red == 0 and is_f64:
                g = o.bits_f64(got)
                if math.isnan(o.bits_f64(acc[b])):
                    assert math.isnan(g)
                else:
                    v = _block_vector(x, spec, thr, b)
                    assert close_f64(g, o.exact_sum(v)), (trial, b, g)
            else:
                assert got == acc[b], (trial, b, spec)
    assert refused < 24


@pytest.mark.parametrize("is_f64", [True, False])
def test_tier0_vector_arithmetic(gpu, is_f64, expr_mode):
    """Loops with vector words (VADD..VRCP, VSADD..VSMAX, VCMPxx, VCMPNULL,
    VCMPNOTNULL, VAND..VANDN, VSHL/VSHR/VSAR) and filters on computed
    vectors, up to three reductions (generated words): partials,
    counts, min and max equal the oracle's bits (vo_tier0_expr_*, pinned
    against the reference engine), NaN results compared as NaN; f64 sums
    within 2 ulp of the exact sum of the reduced register's lanes; i64 sums
    exact (division and modulo with the reference's zero-divisor results).
    An INT64_MIN / -1 lane, a trap in the reference, is refused."""
    from tests import tier0_gen as tg
    from voxel_amd import voxel_py as vx
    rng = np.random.default_rng(71 if is_f64 else 72)

    def same(got, want):
        if is_f64 and math.isnan(o.bits_f64(want)):
            return math.isnan(o.bits_f64(got))
        return got == want

    forced = tg.VEC_BIN + tg.VEC_UN + [tg.VSPLAT]                 # each binary and unary word ends two programs
    for trial in range(30 + 2 * len(forced)):
        last = forced[(trial - 30) // 2] if trial >= 30 else None
        n = int(rng.integers(1, 3000)) * 4
        regs = np.zeros(64, np.uint64)
        if is_f64:
            x = np.round(rng.standard_normal(n) * 100, 2)
            x[rng.integers(0, n, 6)] = 0.0
            if trial % 4 == 0:
                x[rng.integers(0, n, 2)] = np.nan
            for r in tg.SCAL_REGS:
                regs[r] = o.f64_bits(float(np.round(rng.standard_normal() * 10, 1)) if rng.random() < 0.9 else 0.0)
            e = vx.EngineF64()
        else:
            x = rng.integers(-2**40, 2**40, n).astype(np.int64)
            x[rng.integers(0, n, 6)] = 0
            for r in tg.SCAL_REGS:
                regs[r] = int(rng.integers(-2**20, 2**20)) & (2**64 - 1)
            e = vx.EngineI64()
        regs[2] = n
        code, spec = tg.make_expr_program(rng, is_f64, nops=int(rng.integers(1, 9)), last=last)
        e.add_segment(x)
        for r in (2,) + tg.SCAL_REGS:
            e.set_scalar(r, int(regs[r]))
        e.load_program([int(c) for c in code])
        if not is_f64 and tg.expr_traps(o, x, n, spec, regs):
            with pytest.raises(RuntimeError, match="INT64_MIN"):
                e.run()
            continue
        e.run()
        acc, part = tg.oracle_expr(o, x, n, spec, regs, is_f64)
        assert e.get_scalar(1) == n
        for b, red in enumerate(spec["red"]):
            assert same(e.get_scalar(tg.PARTS[b]), part[b]), (trial, b, spec)
            got = e.get_scalar(tg.ACCS[b])
            if red == 0 and is_f64:
                g = o.bits_f64(got)
                if math.isnan(o.bits_f64(acc[b])):
                    assert math.isnan(g)
                else:
                    col = tg.oracle_expr_col(o, x, n, spec, regs, spec["bvreg"][b], is_f64)
                    if np.all(np.isfinite(col)):
                        assert close_f64(g, o.exact_sum(col)), (trial, b, g)
                    else:                                   # infinite lanes: the sum is that infinity
                        assert g == o.bits_f64(acc[b]), (trial, b, g)
            else:
                assert same(got, acc[b]), (trial, b, spec)
    if not is_f64:
        # INT64_MIN / -1 traps in the reference's x86 build: refused, registers untouched
        e = vx.EngineI64()
        x = np.arange(1, 65, dtype=np.int64)
        x[17] = np.iinfo(np.int64).min
        e.add_segment(x)
        e.set_scalar(2, 64)
        e.set_scalar(3, (1 << 64) - 1)                       # R3 = -1
        e.load_program([tg.w(0x10, rd=4, imm=4), tg.w(0x70, rd=0, ra=1), tg.w(0x93, rd=1, ra=0, rb=3),
                        tg.w(0xD0, rd=5, ra=1), tg.w(0x20, rd=8, ra=8, rb=5), tg.w(0x20, rd=1, ra=1, rb=4),
                        tg.w(0x40, ra=1, rb=2), tg.w(0x52, imm=(1 - 8) & 0xFFF), tg.w(0x01)])
        with pytest.raises(RuntimeError, match="INT64_MIN"):
            e.run()
        assert e.get_scalar(8) == 0 and e.get_scalar(1) == 0


@pytest.mark.parametrize("is_f64", [True, False])
def test_tier0_expression_loops_compiled_at_size(gpu, is_f64):
    """Loops of 2^21+ rows with vector words and VSUM / VCOUNT / VRED_MIN /
    VRED_MAX reductions run compiled whole (the default mode compiles loops
    of 2^20 rows or more): partial registers, counts, min and max equal the
    oracle's bits (vo_tier0_expr_*; f64 min / max over both signs take the
    exact in-order fold), i64 sums exact, f64 sums within 2 ulp of the exact
    sum of the reduced register's lanes; the same shape run again with other
    scalar operands reuses the compiled kernel and still matches."""
    from tests import tier0_gen as tg
    from voxel_amd import voxel_py as vx
    assert gpu.voxel_gpu_set_expr_jit(1) in (0, 1, 2)
    rng = np.random.default_rng(75 if is_f64 else 76)
    for trial in range(6):
        n = (1 << 21) + 4 * int(rng.integers(0, 1000))
        code, spec = tg.make_expr_program(rng, is_f64, nops=int(rng.integers(1, 7)),
                                          reds=(0, 1) if trial < 2 else (0, 1, 2, 3), lane_words=trial % 3 == 2)
        for rep in range(2):
            regs = np.zeros(64, np.uint64)
            if is_f64:
                x = np.round(rng.standard_normal(n) * 100, 2)
                for r in tg.SCAL_REGS:
                    regs[r] = o.f64_bits(float(np.round(rng.standard_normal() * 10, 1)) or 1.0)
                e = vx.EngineF64()
            else:
                x = rng.integers(-2**40, 2**40, n).astype(np.int64)
                for r in tg.SCAL_REGS:
                    regs[r] = int(rng.integers(1, 2**20)) & (2**64 - 1)
                e = vx.EngineI64()
            regs[2] = n
            regs[tg.SLIDE_REG] = int(rng.choice([0, 1, 2, -1])) & (2**64 - 1)
            g = x[: n // 2].copy()                            # VGATHER's segment
            e.add_segment(x)
            e.add_segment(g)
            for r in (2, tg.SLIDE_REG) + tg.SCAL_REGS:
                e.set_scalar(r, int(regs[r]))
            e.load_program([int(c) for c in code])
            o.O.vo_expr_reset_masks()
            o.O.vo_expr_set_gather(1, o.ptr(g), g.size)
            if not is_f64 and tg.expr_traps(o, x, n, spec, regs):
                continue
            e.run()
            acc, part = tg.oracle_expr(o, x, n, spec, regs, is_f64)
            assert e.get_scalar(1) == n
            for b, red in enumerate(spec["red"]):
                got_p, got = e.get_scalar(tg.PARTS[b]), e.get_scalar(tg.ACCS[b])
                if is_f64 and math.isnan(o.bits_f64(part[b])):
                    assert math.isnan(o.bits_f64(got_p))
                else:
                    assert got_p == part[b], (trial, rep, b, spec)
                if red == 0 and is_f64:
                    col = tg.oracle_expr_col(o, x, n, spec, regs, spec["bvreg"][b], is_f64)
                    g = o.bits_f64(got)
                    if np.all(np.isfinite(col)):
                        assert close_f64(g, o.exact_sum(col)), (trial, rep, b, g)
                    else:
                        assert math.isnan(g) == math.isnan(o.bits_f64(acc[b]))
                else:
                    assert got == acc[b], (trial, rep, b, spec)


def test_tier0_window_word_loops(gpu, expr_mode):
    """Loops loading WINDOW_SUM / WINDOW_MEAN / WSTD / WVARIANCE_W / WDELTA /
    WQUANTILE (the word advancing the offset), vector words and up to three
    reductions: partials, counts, min and max equal the oracle's bits
    (vo_tier0_window_f64, pinned against the reference engine), NaN results
    compared as NaN; sums within 2 ulp of the exact sum of the reduced lanes
    (computed by the oracle's vector evaluation on the window values).
    Loop bounds the reference never reaches are refused."""
    from tests import tier0_gen as tg
    from voxel_amd import voxel_py as vx
    rng = np.random.default_rng(73)

    def same(got, want):
        if math.isnan(o.bits_f64(want)):
            return math.isnan(o.bits_f64(got))
        return got == want

    for trial in range(48):
        wkind = 1 + trial % 6
        n = int(rng.integers(8, 20000))
        window = int(rng.choice([1, 2, 5, 20, 255])) if wkind != 5 else 0
        if wkind in (3, 4) and window < 2:
            window = 2
        if wkind == 6:
            window = tg.wquantile_imm(rng)                 # the WQUANTILE word's imm
        x = np.round(rng.standard_normal(n) * 100, 2) + 0.0      # no -0.0: a +-0 tie's order is unspecified
        regs = np.zeros(64, np.uint64)
        for r in tg.SCAL_REGS:
            regs[r] = o.f64_bits(float(np.round(rng.standard_normal() * 10, 1)))
        total = tg.window_total(rng, n, wkind)
        regs[2] = total
        code, spec = tg.make_window_program(rng, wkind, window, nops=int(rng.integers(0, 4)))
        ok, acc, part = tg.oracle_window_loop(o, x, total, spec, regs, wkind, window)
        assert ok
        e = vx.EngineF64()
        e.add_segment(x)
        for r in (2,) + tg.SCAL_REGS:
            e.set_scalar(r, int(regs[r]))
        e.load_program([int(c) for c in code])
        e.run()
        assert e.get_scalar(1) == total
        for b, red in enumerate(spec["red"]):
            assert same(e.get_scalar(tg.PARTS[b]), part[b]), (trial, wkind, b, spec)
            got = e.get_scalar(tg.ACCS[b])
            if red == 0:
                g, want = o.bits_f64(got), o.bits_f64(acc[b])
                if math.isnan(want):
                    assert math.isnan(g)
                else:
                    # the reduced lanes: window values row by row, then the body's words
                    wv = np.zeros(max(n, 4 * ((total + 3) // 4)) + 4)
                    kind = {1: "sum", 2: None, 3: "stddev", 4: "variance", 5: "delta", 6: None}[wkind]
                    if wkind == 6:                        # WQUANTILE: one order statistic per window
                        q = tg.wquantile_values(x, window)
                        wv[:q.size] = q
                    elif wkind == 2:                      # WINDOW_MEAN multiplies by 1 / w
                        m = max(n - window + 1, 0)
                        sums = np.zeros(m)
                        o.O.vo_window_f64(1, o.ptr(x), n, o.ptr(sums), m, window, 0.0)
                        wv[:m] = sums * (1.0 / window)
                    else:
                        m = o.O.vo_window_f64(("delta", "sum", "mean", "stddev", "variance").index(kind),
                                              o.ptr(x), n, o.ptr(wv), wv.size, window, 0.0)
                        wv[m:] = 0.0
                    lanes = tg.oracle_expr_col(o, wv[:n], 4 * ((total + 3) // 4), spec, regs, spec["bvreg"][b], True)
                    lanes = lanes[: 4 * ((total + 3) // 4)]
                    if np.all(np.isfinite(lanes)):
                        assert close_f64(g, o.exact_sum(lanes)), (trial, wkind, b)
                    else:                                   # infinite lanes: the sum is that infinity
                        assert g == want, (trial, wkind, b)
            else:
                assert same(got, acc[b]), (trial, wkind, b, spec)
    # the offset starts at (or past) the end the word steps to: one pass over
    # zero lanes, then CMP falls through (the oracle's restatement of it is
    # pinned against the reference engine)
    for trial in range(12):
        wkind = 1 + trial % 6
        n = int(rng.integers(8, 300))
        window = {5: 0, 6: tg.wquantile_imm(rng)}.get(wkind, int(rng.choice([2, 3, 7])))
        x = np.round(rng.standard_normal(n) * 100, 2) + 0.0
        regs = np.zeros(64, np.uint64)
        for r in tg.SCAL_REGS:
            regs[r] = o.f64_bits(float(np.round(rng.standard_normal() * 10, 1)))
        start = (n - 1 if wkind == 5 else n) + int(rng.integers(0, 3))
        code, spec = tg.make_window_program(rng, wkind, window, nops=int(rng.integers(1, 4)))
        ok, acc, part = tg.oracle_window_loop(o, x, start, spec, regs, wkind, window, off=start)
        assert ok
        e = vx.EngineF64()
        e.add_segment(x)
        for r in tg.SCAL_REGS:
            e.set_scalar(r, int(regs[r]))
        e.set_scalar(1, start)
        e.set_scalar(2, start)
        e.load_program([int(c) for c in code])
        e.run()
        assert e.get_scalar(1) == start
        for b in range(len(spec["red"])):
            assert same(e.get_scalar(tg.PARTS[b]), part[b]), (trial, wkind, b, spec)
            assert same(e.get_scalar(tg.ACCS[b]), acc[b]), (trial, wkind, b, spec)
    # a bound the window word never steps onto: refused, registers untouched
    e = vx.EngineF64()
    x = np.arange(10, dtype=np.float64)
    e.add_segment(x)
    e.set_scalar(2, 7)
    code, _ = tg.make_window_program(np.random.default_rng(1), 1, 3, nops=0)
    e.load_program([int(c) for c in code])
    with pytest.raises(RuntimeError, match="would not terminate"):
        e.run()
    assert e.get_scalar(1) == 0


@pytest.mark.parametrize("is_f64", [True, False])
def test_tier0_scalar_prefix_and_epilogue(gpu, is_f64):
    """Scalar words (integer, f64 and bitwise arithmetic, DIV/MOD/DIVF by
    zero stopping the program) before and after a multi-reduction loop: all
    16 scalar registers equal the oracle's (vo_scalar_run, pinned against the
    reference engine, around vo_tier0_multi_*).  f64 engines use the exact
    reductions (VCOUNT, VRED_MIN/MAX) so every register compares bit for bit."""
    from tests import tier0_gen as tg
    from voxel_amd import voxel_py as vx
    rng = np.random.default_rng(71 if is_f64 else 72)
    done = stopped = 0
    while done < 20:
        n = int(rng.integers(1, 1500)) * 4
        if is_f64:
            x = np.round(rng.standard_normal(n) * 100, 2)
            x[rng.integers(0, n, 7)] = 0.0
            thr = [float(np.round(rng.standard_normal() * 60, 2)) for _ in range(2)]
            tb = [o.f64_bits(t) for t in thr]
        else:
            x = rng.integers(-2**40, 2**40, n).astype(np.int64)
            thr = [int(rng.integers(-2**39, 2**39)) for _ in range(2)]
            tb = [t & (2**64 - 1) for t in thr]
        regs0 = np.zeros(64, np.uint64)
        for r in (0, 5, 7, 8, 9, 10, 11, 12, 13, 14, 15):
            regs0[r] = int(rng.integers(-40, 40)) & (2**64 - 1) if rng.random() < 0.5 else o.f64_bits(float(rng.standard_normal()))
        regs0[1], regs0[2], regs0[3], regs0[6] = 0, n, tb[0], tb[1]
        pre = tg.scalar_words(rng, int(rng.integers(0, 6)), range(16), tg.SCRATCH + tg.ACCS)
        epi = tg.scalar_words(rng, int(rng.integers(0, 6)), range(16), range(16))
        code, spec = tg.make_program(rng, is_f64, reds=(1, 2, 3) if is_f64 else (0, 1, 2, 3), prefix=pre, epilogue=epi)
        # expected registers
        exp = regs0.copy()
        pw = np.array(pre + [0], np.uint32)
        rc = o.O.vo_scalar_run(o.ptr(pw), len(pre), o.ptr(exp))
        if rc < 0:
            continue
        if rc == 1:
            exp[4] = 4
            nb = len(spec["red"])
            acc0 = [int(exp[tg.ACCS[b]]) for b in range(nb)]
            t_now = [o.bits_f64(int(exp[3])), o.bits_f64(int(exp[6]))] if is_f64 else \
                    [int(np.int64(np.uint64(exp[3]))), int(np.int64(np.uint64(exp[6])))]
            acc, part = tg.oracle_run(o, x, n, spec, t_now, is_f64, acc0=acc0)
            for b in range(nb):
                exp[tg.ACCS[b]], exp[tg.PARTS[b]] = acc[b], part[b]
            exp[1] = n
            ew = np.array(epi + [0], np.uint32)
            if o.O.vo_scalar_run(o.ptr(ew), len(epi), o.ptr(exp)) < 0:
                continue
        else:
            stopped += 1
        e = vx.EngineF64() if is_f64 else vx.EngineI64()
        e.add_segment(x)
        for r in range(16):
            e.set_scalar(r, int(regs0[r]))
        e.load_program([int(c) for c in code])
        e.run()
        got = [e.get_scalar(r) for r in range(16)]
        assert got == [int(v) for v in exp[:16]], (done, [hex(int(c)) for c in code])
        done += 1


@pytest.mark.parametrize("is_f64", [True, False])
def test_tier0_vstore(gpu, is_f64):
    """Loops that VSTORE computed vectors at the loop offset (into a third
    segment over caller memory, or into the loaded column itself; lane
    counts 1..4 and more; rows past the segment's end dropped), with or
    without reductions, after vector arithmetic, filters and lane words:
    the caller's arrays hold exactly what the oracle (vo_tier0_expr_* with
    its store segment, pinned against the reference engine) wrote, and the
    registers match as in the arithmetic tests."""
    from tests import tier0_gen as tg
    from voxel_amd import voxel_py as vx
    rng = np.random.default_rng(77 if is_f64 else 78)
    checked = 0
    for trial in range(40):
        n = int(rng.integers(1, 2000)) * 4
        m = int(rng.integers(1, 3000))
        ny = int(rng.integers(1, n + 9))
        regs = np.zeros(64, np.uint64)
        if is_f64:
            x = np.round(rng.standard_normal(n) * 100, 2)
            g = np.round(rng.standard_normal(m) * 10, 3)
            y = np.round(rng.standard_normal(ny), 1)
            for r in tg.SCAL_REGS:
                regs[r] = o.f64_bits(float(np.round(rng.standard_normal() * 10, 1)))
            e = vx.EngineF64()
        else:
            x = rng.integers(-20, 6000, n).astype(np.int64)
            g = rng.integers(-1000, 1000, m).astype(np.int64)
            y = rng.integers(-9, 9, ny).astype(np.int64)
            for r in tg.SCAL_REGS:
                regs[r] = int(rng.integers(-5, 50)) & (2**64 - 1)
            e = vx.EngineI64()
        regs[tg.SLIDE_REG] = int(rng.choice([0, 1, 2, -1])) & (2**64 - 1)
        regs[2] = n
        code, spec = tg.make_expr_program(rng, is_f64, nops=int(rng.integers(1, 7)), lane_words=trial % 2 == 0,
                                          store_segs=(2,) if trial % 3 else (0,),
                                          lane_ops=[op for op in tg.LANE_OPS if op != 0x8F])
        if not is_f64 and tg.expr_traps(o, x, n, spec, regs):
            continue
        sid = 2 if trial % 3 else 0
        xo, yo = x.copy(), y.copy()
        o.O.vo_expr_reset_masks()
        o.O.vo_expr_set_gather(1, o.ptr(g), m)
        o.O.vo_expr_set_store(sid, o.ptr(xo if sid == 0 else yo), n if sid == 0 else ny)
        acc, part = tg.oracle_expr(o, xo, n, spec, regs, is_f64) if spec["red"] else ([], [])
        if not spec["red"]:
            tg.oracle_expr(o, xo, n, spec, regs, is_f64)
        for a in (x, g, y):
            e.add_segment(a)
        for r in (2, tg.SLIDE_REG) + tg.SCAL_REGS:
            e.set_scalar(r, int(regs[r]))
        e.load_program([int(c) for c in code])
        e.run()
        assert np.array_equal(x.view(np.uint64), xo.view(np.uint64)), (trial, spec)
        assert np.array_equal(y.view(np.uint64), yo.view(np.uint64)), (trial, spec)
        assert e.get_scalar(1) == n
        for b, red in enumerate(spec["red"]):
            if red == 0 and is_f64:
                continue                              # f64 sums: the arithmetic tests' tolerance; here the stores
            got_p, got = e.get_scalar(tg.PARTS[b]), e.get_scalar(tg.ACCS[b])
            if is_f64 and math.isnan(o.bits_f64(part[b])):
                assert math.isnan(o.bits_f64(got_p))
            else:
                assert got_p == part[b], (trial, b, spec)
        checked += 1
    assert checked >= 30


@pytest.mark.parametrize("is_f64", [True, False])
def test_tier0_lane_words(gpu, is_f64, expr_mode):
    """Loops with the lane words (VINSERT, VPERMUTE, VSHUFFLE, VREVERSE,
    VROTATE, VSLIDE by a register, VBLEND on masks that filters and compares
    of the same iteration wrote, VPOW, VGATHER from a second segment or a
    missing one) mixed with vector arithmetic and filters (generated words):
    partials, counts, min and max equal the oracle's bits (vo_tier0_expr_*,
    pinned against the reference engine), NaN compared as NaN; sums of
    reduced registers within 2 ulp of the exact sum of the oracle's lanes
    (i64 exact).  An f64 VPOW lane is the device pow: programs with one are
    compared within 1e-12 of the sum of |lanes| (the reference's pow is the
    host libm's; both are faithfully rounded)."""
    from tests import tier0_gen as tg
    from voxel_amd import voxel_py as vx
    rng = np.random.default_rng(75 if is_f64 else 76)

    def same(got, want):
        if is_f64 and math.isnan(o.bits_f64(want)):
            return math.isnan(o.bits_f64(got))
        return got == want

    forced = tg.LANE_OPS
    checked = 0
    for trial in range(40 + 3 * len(forced)):
        last = forced[(trial - 40) // 3] if trial >= 40 else None
        n = int(rng.integers(1, 3000)) * 4
        m = int(rng.integers(1, 5000))
        regs = np.zeros(64, np.uint64)
        if is_f64:
            x = np.round(rng.standard_normal(n) * 100, 2)
            x[rng.integers(0, n, 6)] = 0.0
            g = np.round(rng.standard_normal(m) * 10, 3)
            for r in tg.SCAL_REGS:
                regs[r] = o.f64_bits(float(np.round(rng.standard_normal() * 10, 1)))
            e = vx.EngineF64()
        else:
            x = rng.integers(-20, 6000, n).astype(np.int64)
            g = rng.integers(-1000, 1000, m).astype(np.int64)
            for r in tg.SCAL_REGS:
                regs[r] = int(rng.integers(-5, 50)) & (2**64 - 1)
            e = vx.EngineI64()
        regs[tg.SLIDE_REG] = int(rng.choice([0, 1, 2, 3, 5, -1, -2])) & (2**64 - 1)
        regs[2] = n
        code, spec = tg.make_expr_program(rng, is_f64, nops=int(rng.integers(1, 9)), last=last, lane_words=True)
        e.add_segment(x)
        e.add_segment(g)
        for r in (2, tg.SLIDE_REG) + tg.SCAL_REGS:
            e.set_scalar(r, int(regs[r]))
        e.load_program([int(c) for c in code])
        if not is_f64 and tg.expr_traps(o, x, n, spec, regs):
            continue
        e.run()
        o.O.vo_expr_reset_masks()
        o.O.vo_expr_set_gather(1, o.ptr(g), m)
        acc, part = tg.oracle_expr(o, x, n, spec, regs, is_f64)
        has_pow = is_f64 and any(op[0] == 0x8F for op in spec["ops"])
        assert e.get_scalar(1) == n
        for b, red in enumerate(spec["red"]):
            got_p, got = e.get_scalar(tg.PARTS[b]), e.get_scalar(tg.ACCS[b])
            col = tg.oracle_expr_col(o, x, n, spec, regs, spec["bvreg"][b], is_f64)
            if has_pow and red != 1:
                for gv, wv, lanes in ((got_p, part[b], col[-4:]), (got, acc[b], col)):
                    gf, wf = o.bits_f64(gv), o.bits_f64(wv)
                    if math.isnan(wf) or math.isinf(wf):
                        assert math.isnan(gf) == math.isnan(wf) and (math.isnan(wf) or gf == wf), (trial, b)
                    else:
                        fin = np.abs(lanes[np.isfinite(lanes)])
                        scale = max(float(fin.sum()) if fin.size else 0.0, abs(wf))
                        assert abs(gf - wf) <= 1e-12 * scale + 1e-300, (trial, b, gf, wf)
                continue
            assert same(got_p, part[b]), (trial, b, spec)
            if red == 0 and is_f64:
                gf = o.bits_f64(got)
                if math.isnan(o.bits_f64(acc[b])):
                    assert math.isnan(gf)
                elif np.all(np.isfinite(col)):
                    assert close_f64(gf, o.exact_sum(col)), (trial, b, gf)
                else:
                    assert gf == o.bits_f64(acc[b]), (trial, b, gf)
            else:
                assert same(got, acc[b]), (trial, b, spec)
        checked += 1
    assert checked >= 50
    # refused shapes: VBLEND on a mask the iteration has not written, a partial VGATHER
    e = vx.EngineF64()
    e.add_segment(np.arange(64, dtype=np.float64))
    e.set_scalar(2, 64)
    body_blend = [tg.w(0x10, rd=4, imm=4), tg.w(0x70, rd=0, ra=1), tg.w(0xC7, rd=1, ra=0, rb=0, imm=3),
                  tg.w(0xD0, rd=5, ra=1), tg.w(0x2A, rd=8, ra=8, rb=5), tg.w(0x20, rd=1, ra=1, rb=4),
                  tg.w(0x40, ra=1, rb=2), tg.w(0x52, imm=(1 - 8) & 0xFFF), tg.w(0x01)]
    e.load_program(body_blend)
    with pytest.raises(RuntimeError, match="VBLEND"):
        e.run()
    body_gather = list(body_blend)
    body_gather[2] = tg.w(0x72, rd=1, ra=0, imm=(0 << 8) | 2)
    e.load_program(body_gather)
    with pytest.raises(RuntimeError, match="VGATHER"):
        e.run()


@pytest.mark.parametrize("is_f64", [True, False])
def test_segment_aggregates_against_oracle(gpu, is_f64):
    """Straight-line programs of AGG_* words (persistent slots across three
    runs of one engine) and scalar words: counts, MIN/MAX, FIRST/LAST,
    distinct counts, MEDIAN and PERCENTILE bit-identical with the oracle
    (vo_straight_run, pinned against the reference engine); f64 SUM/AVG
    within 1e-12 of the sum of |x| (the reference adds in row order, the GPU
    a compensated sum), STDDEV/VARIANCE within 1e-9 relative (Welford in row
    order vs merged workgroup moments; the i64 engine's truncated integer
    within 1)."""
    from tests import tier0_gen as tg
    from voxel_amd import voxel_py as vx
    rng = np.random.default_rng(81 if is_f64 else 82)
    inexact = {0xE1, 0xE2, 0xE7, 0xE8} if is_f64 else {0xE7, 0xE8}
    runs = 0
    for trial in range(12):
        writers = {}
        n = int(rng.integers(1, 20_000))
        if is_f64:
            x = np.round(rng.standard_normal(n) * 1e3, 3)
            x[rng.integers(0, n, 4)] = 0.0
            if trial % 4 == 0:
                x[rng.integers(0, n, 2)] = np.inf
        else:
            x = rng.integers(-10**6, 10**6, n).astype(np.int64)
            x[rng.integers(0, n, 4)] = 0
        scale = float(np.abs(x[np.isfinite(x)]).sum()) if is_f64 else 0.0
        e = vx.EngineF64() if is_f64 else vx.EngineI64()
        e.add_segment(x)
        st = o.O.vo_agg_create(1 if is_f64 else 0)
        regs = np.zeros(64, np.uint64)
        try:
            for k in range(3):
                code = tg.straight_program(rng, n, ops=tg.AGG_OPS, scalar_reads=(1, 2, 3, 13, 14, 15))
                rc = o.O.vo_straight_run(st, o.ptr(x.view(np.uint64)), n, o.ptr(code), code.size, o.ptr(regs))
                assert rc >= 0
                e.load_program([int(c) for c in code])
                e.run()
                writers.update(tg.last_agg_writers(code))
                for r in range(16):
                    got, want = e.get_scalar(r), int(regs[r])
                    if got == want:
                        continue
                    op = writers.get(r)
                    assert op in inexact, (trial, k, r, hex(op or 0), got, want)
                    if is_f64:
                        g, w = o.bits_f64(got), o.bits_f64(want)
                        if op in (0xE1, 0xE2):
                            assert abs(g - w) <= 1e-12 * scale, (trial, k, r, g, w)
                        else:
                            assert abs(g - w) <= 1e-9 * abs(w), (trial, k, r, g, w)
                    else:
                        assert abs(int(np.int64(np.uint64(got))) - int(np.int64(np.uint64(want)))) <= 1
                runs += 1
        finally:
            o.O.vo_agg_destroy(st)
    assert runs == 36


def test_segment_aggregate_programs_are_lowered_or_refused(gpu):
    """AGG_MODE runs where the largest count is unique (NaN rows are entries
    of their own) and is refused at run time on a tie, with the registers
    untouched; NaN keys in AGG_COUNT_DISTINCT / AGG_SUM_DISTINCT are refused
    at run time with the registers untouched."""
    from tests import tier0_gen as tg
    from voxel_amd import voxel_py as vx
    from voxel_amd._lib import lib
    x = np.array([1.0, np.nan, 2.0, 2.0])
    e = vx.EngineF64()
    e.add_segment(x)
    e.load_program([tg.w(0xEC, rd=4, ra=0, rb=1), 1])
    e.run()
    assert o.bits_f64(e.get_scalar(4)) == 2.0
    e.load_program([tg.w(0xEC, rd=5, ra=0, rb=1, imm=2), 1])             # 1.0 and NaN: one row each
    with pytest.raises(RuntimeError, match="AGG_MODE"):
        e.run()
    assert e.get_scalar(5) == 0
    e.load_program([tg.w(0xEA, rd=9, ra=0, rb=1), 1])
    with pytest.raises(RuntimeError, match="NaN"):
        e.run()
    assert e.get_scalar(9) == 0
    e.load_program([tg.w(0xE9, rd=4, ra=0, rb=1), 1])
    with pytest.raises(RuntimeError, match="NaN"):
        e.run()
    assert o.bits_f64(e.get_scalar(4)) == 2.0                             # the MODE run's result, untouched
    e.load_program([tg.w(0xE0, rd=4, ra=0, rb=1), tg.w(0xE3, rd=5, ra=0, rb=1), 1])
    e.run()
    assert e.get_scalar(4) == 4 and o.bits_f64(e.get_scalar(5)) == 1.0     # NaN counts as != 0; MIN skips it


@pytest.mark.parametrize("is_f64", [True, False])
def test_sum_distinct_and_mode_against_oracle(gpu, is_f64):
    """Straight-line programs with AGG_SUM_DISTINCT (sharing the distinct set
    of AGG_COUNT_DISTINCT's slot across three runs of one engine) and
    AGG_MODE, against the oracle (vo_straight_run, pinned against the
    reference engine): registers bit-identical, except an f64 SUM_DISTINCT
    or SUM within 1e-12 of the sum of |x| over the distinct values / rows.
    A program whose AGG_MODE has a tied largest count fails on the GPU with
    the registers untouched (the oracle refuses it too)."""
    from tests import tier0_gen as tg
    from voxel_amd import voxel_py as vx
    rng = np.random.default_rng(83 if is_f64 else 84)
    runs = refused = 0
    for trial in range(10):
        n = int(rng.integers(1, 30_000))
        if is_f64:
            x = np.round(rng.standard_normal(n) * 8, 1)
            x[rng.integers(0, n, 3)] = -0.0
        else:
            x = rng.integers(-60, 60, n).astype(np.int64) * 1_000_003
        scale = float(np.abs(x).sum()) if is_f64 else 0.0
        e = vx.EngineF64() if is_f64 else vx.EngineI64()
        e.add_segment(x)
        st = o.O.vo_agg_create(1 if is_f64 else 0)
        regs = np.zeros(64, np.uint64)
        writers = {}
        try:
            for k in range(3):
                code = tg.straight_program(rng, n, k=8, ops=[0xE9, 0xEA, 0xEA, 0xEC, 0xE1],
                                           scalar_reads=(1, 2, 3, 13, 14, 15))
                rc = o.O.vo_straight_run(st, o.ptr(x.view(np.uint64)), n, o.ptr(code), code.size, o.ptr(regs))
                e.load_program([int(c) for c in code])
                if rc < 0:
                    # the oracle refused a tied AGG_MODE: so does the GPU, and the
                    # engine is left as it was (the oracle's slots may have moved:
                    # start a fresh pair)
                    snap = [e.get_scalar(r) for r in range(16)]
                    with pytest.raises(RuntimeError, match="AGG_MODE"):
                        e.run()
                    assert [e.get_scalar(r) for r in range(16)] == snap
                    refused += 1
                    break
                e.run()
                # every AGG word that writes a register (a program may stop at a
                # division by zero before its last writer runs)
                for x_ in code:
                    if 0xE0 <= (int(x_) & 0xFF) <= 0xED and ((int(x_) >> 12) & 0xF) == 0:
                        writers.setdefault((int(x_) >> 8) & 0xF, set()).add(int(x_) & 0xFF)
                for r in range(16):
                    got, want = e.get_scalar(r), int(regs[r])
                    if got == want:
                        continue
                    assert is_f64 and writers.get(r, set()) & {0xEA, 0xE1}, (trial, k, r, writers.get(r))
                    assert abs(o.bits_f64(got) - o.bits_f64(want)) <= 1e-12 * max(scale, 1.0), (trial, k, r)
                runs += 1
        finally:
            o.O.vo_agg_destroy(st)
    assert runs >= 15, (runs, refused)


def test_failed_aggregate_program_leaves_slots_untouched(gpu):
    """A straight-line program that fails part-way leaves the registers and
    the persistent AGG slots as they were: after AGG_SUM, AGG_COUNT_DISTINCT
    and AGG_COUNT updated their slots, an AGG_MEDIAN over a NaN segment (or
    an AGG_COUNT_DISTINCT on a slot AGG_SUM started) fails; the same words
    run again without the failing one then give bit for bit what one clean
    run on a fresh engine gives (no row counted twice)."""
    from tests import tier0_gen as tg
    from voxel_amd import voxel_py as vx
    rng = np.random.default_rng(23)
    clean = np.round(rng.standard_normal(5000) * 100, 1)
    bad = clean.copy()
    bad[17] = np.nan
    good = [tg.w(0xE1, rd=4, ra=0, rb=1), tg.w(0xE9, rd=5, ra=0, rb=1), tg.w(0xE0, rd=6, ra=0, rb=1)]
    failing = [good + [tg.w(0xEB, rd=7, ra=1, rb=1), 1],            # MEDIAN over NaN
               good + [tg.w(0xE9, rd=4, ra=0, rb=1), 1]]            # COUNT_DISTINCT on SUM's slot
    fresh = vx.EngineF64()
    fresh.add_segment(clean)
    fresh.add_segment(bad)
    fresh.load_program(good + [1])
    fresh.run()
    want = [fresh.get_scalar(r) for r in (4, 5, 6)]
    assert want[1] == np.unique(clean).size and want[2] == np.count_nonzero(clean)
    for prog in failing:
        e = vx.EngineF64()
        e.add_segment(clean)
        e.add_segment(bad)
        e.load_program(prog)
        with pytest.raises(RuntimeError):
            e.run()
        assert [e.get_scalar(r) for r in (4, 5, 6, 7)] == [0, 0, 0, 0]
        e.load_program(good + [1])
        e.run()
        assert [e.get_scalar(r) for r in (4, 5, 6)] == want
        # a second clean run accumulates on the committed slots only
        e.run()
        fresh2 = vx.EngineF64()
        fresh2.add_segment(clean)
        fresh2.load_program(good + [1])
        fresh2.run()
        fresh2.run()
        assert [e.get_scalar(r) for r in (4, 5, 6)] == [fresh2.get_scalar(r) for r in (4, 5, 6)]


def test_tier0_minmax_mixed_signs_uses_exact_fold(gpu, golden):
    # values whose bit patterns straddle both wrap points: the GPU falls back
    # to the in-order fold; result must equal the oracle's bits exactly
    rng = np.random.default_rng(9)
    for trial in range(4):
        x = rng.standard_normal(4 * 1500) * 1e3
        x[::11] = rng.choice([np.inf, -np.inf, 1e300, -1e300, 5e-324], x[::11].size)
        for name, (red, accop) in (("vq_min", (2, 2)), ("vq_max", (3, 3))):
            thr = -5e2
            e = engine_f64(x, words(golden, name), {2: x.size, 3: o.f64_bits(thr)})
            e.run()
            part = np.zeros(1, np.uint64)
            r = o.O.vo_tier0_f64(o.ptr(x), x.size, 0, x.size, 4, thr, 0, red, accop, 0, o.ptr(part))
            assert e.get_scalar(0) == int(r) and e.get_scalar(5) == int(part[0]), (trial, name)


def test_tier0_zero_padding_past_count(gpu, golden):
    # R[total] past Count: VLOAD reads zeros there (engine.hpp:1959-1962)
    x = np.array([5.0, -3.0, 7.0, 1.0, 2.0, 9.0], np.float64)
    for name, (red, accop) in (("vq_count", (1, 1)), ("vq_max", (3, 3)), ("vq_min", (2, 2))):
        e = engine_f64(x, words(golden, name), {2: 12, 3: o.f64_bits(-10.0)})
        e.run()
        part = np.zeros(1, np.uint64)
        r = o.O.vo_tier0_f64(o.ptr(x), x.size, 0, 12, 4, -10.0, 0, red, accop, 0, o.ptr(part))
        assert (e.get_scalar(0), e.get_scalar(5)) == (int(r), int(part[0])), name


def test_tier0_nonterminating_loop_is_refused(gpu, golden):
    x = np.arange(10, dtype=np.float64)
    e = engine_f64(x, words(golden, "vq_sum"), {2: 10, 3: o.f64_bits(1.0)})   # 10 rows: 10 % 4 != 0
    with pytest.raises(RuntimeError, match="would not terminate"):
        e.run()
    assert e.get_scalar(0) == 0 and e.get_scalar(1) == 0


# ---------------------------------------------------------------- JIT
def test_jit_run_golden(gpu, golden, x42):
    from voxel_amd import voxel_py as vx
    from voxel_amd._lib import lib
    g = golden["seed42"]
    c = np.array(words(golden, "fast"), np.uint32)
    r = lib.voxel_jit_run(c.ctypes.data, c.size, x42.ctypes.data, x42.size, 500.0)
    exact = o.relational(x42, 4, 500.0)["sum"]
    assert close_f64(r, exact)
    ref = o.bits_f64(B(g["jit"]["result"]))
    assert abs(r - ref) <= 1e-12 * ref
    assert close_f64(vx.jit_run(words(golden, "vq_sum"), x42, 500.0, x42.size), exact)
    d8 = np.arange(1, 9, dtype=np.float64)
    assert lib.voxel_jit_run(c.ctypes.data, c.size, d8.ctypes.data, 8, 4.0) == 26.0
    # not fusable: 0.0 and an error, as the reference's C entry returns 0.0
    cc = np.array(words(golden, "vq_count"), np.uint32)
    assert lib.voxel_jit_run(cc.ctypes.data, cc.size, x42.ctypes.data, x42.size, 500.0) == 0.0
    assert b"JIT lowering" in lib.voxel_last_error()
    with pytest.raises(RuntimeError):
        vx.jit_run(words(golden, "vq_count"), x42, 500.0, x42.size)


def test_jit_ordered_compares_and_over_read(gpu, golden):
    from voxel_amd._lib import lib
    x = np.array([1.0, np.nan, 3.0, 4.0, np.nan, 6.0, 7.0, 8.0, 9.0, 10.0], np.float64)
    base = words(golden, "fast")
    for fop in range(0xC1, 0xC7):
        code = np.array(base, np.uint32)
        code[1] = (int(code[1]) & 0xFFFFFF00) | fop
        r = lib.voxel_jit_run(code.ctypes.data, code.size, x.ctypes.data, x.size, 4.0)
        ref = o.O.vo_jit_f64(o.ptr(x), x.size, x.size, 0, fop, 4.0)
        assert r == ref, hex(fop)                                    # NE fails NaN in the JIT


# ---------------------------------------------------------------- run_parallel
def test_run_parallel_golden(gpu, golden, x42):
    g = golden["seed42"]
    for t in (8, 7):
        e = engine_f64(x42, words(golden, "fast"), {2: x42.size, 3: o.f64_bits(500.0)})
        e.run_parallel(0, t, 0)
        ref = o.bits_f64(B(g["run_parallel_%d" % t]["R0"]))
        got = e.get_scalar_f64(0)
        # each worker drops its own tail: compare with the exact sum of the rows kept
        chunk = -(-x42.size // t)
        kept = []
        for w in range(t):
            s = w * chunk
            cnt = min(chunk, x42.size - s)
            seg = x42[s:s + (cnt // 4) * 4]
            kept.append(seg[seg > 500])
        exact = o.exact_sum(np.concatenate(kept))
        assert close_f64(got, exact, ulps=4) and abs(got - ref) <= 1e-12 * ref


# ---------------------------------------------------------------- relational scans
TORCH_DT = {np.float64: "float64", np.float32: "float32", np.int64: "int64", np.int32: "int32"}


def dev(x):
    import torch
    return torch.from_numpy(x).cuda()


def gen(dt, n, rng):
    if np.dtype(dt).kind == "f":
        return (rng.random(n) * 1000).astype(dt)
    return rng.integers(0, 2**31, n).astype(dt)


def test_run_parallel_tier0_programs(gpu):
    """RunParallel over generated Tier-0 loops (engine.hpp:143-175; the
    reference exposes it on Engine<f64> only) through the engine ABI, against
    the per-chunk restatement tier0_gen.run_parallel_expected (pinned bit for
    bit against the reference's own RunParallel in test_oracle): counts,
    minima and maxima bit-exact; sums within 1e-12 of the sum of |x| (the
    reference adds lanes in row order, the GPU a compensated sum).  1 to 16
    workers, chunks of ragged length."""
    from tests import tier0_gen as tg
    from voxel_amd import voxel_py as vx
    is_f64 = True
    rng = np.random.default_rng(91)
    checked = 0
    for trial in range(24):
        n = int(rng.integers(1, 1500)) * 4
        if is_f64:
            x = rng.standard_normal(n) * 100
            x[rng.integers(0, n, 5)] = 0.0
            thr = [float(rng.standard_normal() * 60) for _ in range(2)]
            tb = [o.f64_bits(t) for t in thr]
            e = vx.EngineF64()
        else:
            x = rng.integers(-2**40, 2**40, n).astype(np.int64)
            thr = [int(rng.integers(-2**39, 2**39)) for _ in range(2)]
            tb = [t & (2**64 - 1) for t in thr]
            e = vx.EngineI64()
        code, spec = tg.make_program(rng, is_f64)
        if not tg.mergeable(spec):
            continue
        threads = int(rng.choice([1, 2, 3, 7, 8, 16]))
        b = int(rng.integers(0, len(spec["red"])))
        use_acc = bool(rng.integers(0, 2))
        rr = tg.ACCS[b] if use_acc else tg.PARTS[b]
        e.add_segment(x)
        for r, v in ((2, n), (3, tb[0]), (6, tb[1])):
            e.set_scalar(r, v)
        e.load_program([int(c) for c in code])
        e.run_parallel(0, threads, rr)
        want = tg.run_parallel_expected(o, x, n, spec, thr, threads, b, use_acc, is_f64)
        got = e.get_scalar(rr)
        if is_f64 and use_acc and spec["red"][b] == 0:
            assert abs(o.bits_f64(got) - o.bits_f64(want)) <= 1e-12 * float(np.abs(x).sum()), (trial, spec)
        else:
            assert got == want, (trial, threads, spec)
        assert e.get_scalar(2) == n                       # the caller's registers other than R[result] stay
        checked += 1
    assert checked >= 12


def test_run_parallel_straight_programs(gpu):
    """RunParallel of straight-line AGG_* programs: each worker is a fresh
    engine (fresh aggregate slots) over its chunk; R[result] = the sum in f64
    of the workers' R[result] (register bits read as f64, as ScalarAs<T>
    does), against vo_straight_run per chunk (pinned against the reference
    engine), on registers whose last writer is exact (not SUM/AVG/STDDEV/
    VARIANCE).  A sort word sorts each chunk of the caller's rows in place,
    as the reference's workers do."""
    from tests import tier0_gen as tg
    from voxel_amd import voxel_py as vx
    rng = np.random.default_rng(93)
    checked = 0
    for trial in range(14):
        n = int(rng.integers(1, 5000))
        x = np.round(rng.standard_normal(n) * 1e3, 3)
        code = tg.straight_program(rng, n, ops=tg.AGG_OPS, scalar_reads=(1, 2, 3, 13, 14, 15))
        writers = tg.last_agg_writers(code)
        exact = [r for r, op in writers.items() if op not in (0xE1, 0xE2, 0xE7, 0xE8)]
        if not exact:
            continue
        rr = int(rng.choice(exact))
        threads = int(rng.choice([1, 2, 5, 8]))
        e = vx.EngineF64()
        e.add_segment(x)
        e.load_program([int(c) for c in code])
        e.run_parallel(0, threads, rr)
        chunk = -(-n // threads)
        want = 0.0
        for t in range(threads):
            s0 = t * chunk
            if s0 >= n:
                continue
            xc = np.ascontiguousarray(x[s0:s0 + min(chunk, n - s0)])
            st = o.O.vo_agg_create(1)
            try:
                regs = np.zeros(64, np.uint64)
                assert o.O.vo_straight_run(st, o.ptr(xc.view(np.uint64)), xc.size, o.ptr(code), code.size, o.ptr(regs)) >= 0
            finally:
                o.O.vo_agg_destroy(st)
            want = want + o.bits_f64(int(regs[rr]))
        assert e.get_scalar(rr) == o.f64_bits(want), (trial, threads, rr, hex(writers[rr]))
        checked += 1
    assert checked >= 8
    # SORT_ASC over each worker's chunk: the caller's rows come back sorted chunk by chunk
    x = np.round(rng.standard_normal(103) * 1e3, 3)
    keep = x.copy()
    e = vx.EngineF64()
    e.add_segment(x)
    e.load_program([tg.w(0xF2, rd=4, ra=0), 1])
    e.run_parallel(0, 4, 4)
    chunk = -(-x.size // 4)
    want = np.concatenate([np.sort(keep[i:i + chunk]) for i in range(0, x.size, chunk)])
    assert np.array_equal(x, want)
    # each worker's R4 = its chunk's row count (integer bits), summed as f64
    sizes = [min(chunk, x.size - i) for i in range(0, x.size, chunk)]
    assert e.get_scalar(4) == o.f64_bits(sum((o.bits_f64(c) for c in sizes), 0.0))


@pytest.mark.parametrize("dt", [np.float64, np.float32, np.int64, np.int32])
@pytest.mark.parametrize("n", [0, 1, 5, 1023, 65_537, 2_000_003])
def test_relational_compound_scan(gpu, dt, n):
    from voxel_amd import scan
    rng = np.random.default_rng(n + 7)
    a, b = gen(dt, n, rng), gen(dt, n, rng)
    if np.dtype(dt).kind == "f":
        ta, tb = 292.8932188, 707.1067812
    else:
        ta, tb = int((1 - math.sqrt(0.5)) * 2**31), int(math.ceil(math.sqrt(0.5) * 2**31))
    ta, tb = dt(ta), dt(tb)
    r = scan.scan(dev(a), "gt", ta, dev(b), "lt", tb)
    ref = o.relational(a, 4, ta, b, 2, tb)
    assert r["count"] == ref["count"]
    if np.dtype(dt).kind == "f":
        if dt == np.float64:
            assert close_f64(r["sum"], ref["sum"])
        else:
            e = float(ref["sum"])
            assert abs(r["sum"] - e) <= 4 * float(np.spacing(np.float32(abs(e)))) + 0.0
        assert (r["min"], r["max"]) == ((ref["min"], ref["max"]) if ref["min"] is not None else (None, None))
    else:
        assert r["sum"] == ref["sum"]
        assert (r["min"], r["max"]) == ((int(ref["min"]), int(ref["max"])) if ref["count"] else (None, None))


def test_relational_special_values_and_unaligned(gpu):
    import torch
    from voxel_amd import scan
    x = np.array([np.nan, -0.0, 0.0, np.inf, -np.inf, 1.0, -1.0, 5e-324, np.nan, 2.0, 3.0], np.float64)
    for op in ("eq", "ne", "lt", "le", "gt", "ge", "true"):
        for t in (0.0, -0.0, 1.0, np.inf, np.nan):
            opc = scan.OPS[op]
            r = scan.scan(dev(x), op, t)
            ref = o.relational(x, opc, t)
            assert r["count"] == ref["count"], (op, t)
    # unaligned start: a column view one element in (pointer not 16-byte aligned)
    big = np.arange(1, 100_001, dtype=np.float64)
    t = torch.from_numpy(big).cuda()[1:]
    r = scan.scan(t, "gt", 10.0)
    sel = big[1:][big[1:] > 10]
    assert r["count"] == sel.size and r["sum"] == float(sel.sum()) and r["min"] == 11.0
    # b shifted against a (different misalignment): scalar path
    r = scan.scan(torch.from_numpy(big).cuda()[1:50001], "gt", 10.0, torch.from_numpy(big).cuda()[2:50002], "lt", 30000.0)
    aa, bb = big[1:50001], big[2:50002]
    assert r["count"] == int(((aa > 10) & (bb < 30000)).sum())


def test_scan_is_deterministic(gpu):
    from voxel_amd import scan
    rng = np.random.default_rng(3)
    a = rng.standard_normal(3_000_001) * 1e6
    d = dev(a)
    r1 = scan.scan(d, "gt", -1.0)
    r2 = scan.scan(d, "gt", -1.0)
    assert (r1["sum_hi"], r1["sum_lo"], r1["count"]) == (r2["sum_hi"], r2["sum_lo"], r2["count"])
    assert close_f64(r1["sum"], o.relational(a, 4, -1.0)["sum"])


def test_full_size_compound_properties(gpu):
    """2^28-row f64 columns: count of (a>t1 AND b<t2) plus count of (a>t1 AND
    b>=t2) equals count(a>t1); the sum splits the same way (within the stated
    ulp bound); counts match torch's exact integer counts."""
    import torch
    from voxel_amd import scan
    n = 1 << 28
    g = torch.Generator(device="cuda").manual_seed(1)
    a = torch.rand(n, dtype=torch.float64, device="cuda", generator=g) * 1000
    b = torch.rand(n, dtype=torch.float64, device="cuda", generator=g) * 1000
    t1, t2 = 292.8932188, 707.1067812
    r = scan.scan(a, "gt", t1, b, "lt", t2)
    r2 = scan.scan(a, "gt", t1, b, "ge", t2)
    r1 = scan.scan(a, "gt", t1)
    assert r["count"] + r2["count"] == r1["count"]
    assert r["count"] == int(((a > t1) & (b < t2)).sum().item())
    assert abs((r["sum"] + r2["sum"]) - r1["sum"]) <= 4 * o.ulp(r1["sum"])
    assert r["min"] == float(a[(a > t1) & (b < t2)].min().item())
    del a, b
    torch.cuda.empty_cache()

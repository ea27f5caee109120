"""GPU tests of the voxel_py operator surface (voxel_amd/voxel_py.py, the
reference's voxel_bindings.cpp:224-300) and of voxel_query's group_by /
sort / topk paths (bindings/voxel_query.py:150-193), through the C ABI.

Bars: bitmaps, counts, indices, join pairs and group counts bit-exact with
the oracle (vo_bitmap_f64, vo_mergesort_f64, vo_hashjoin_i64) or numpy.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as o

pytestmark = pytest.mark.gpu


def test_vector_filter_apply_gt(gpu):
    from voxel_amd import voxel_py as vx
    rng = np.random.default_rng(1)
    for n in (0, 1, 63, 64, 65, 100_003):
        x = np.round(rng.standard_normal(n) * 4)                      # ties at the threshold
        words, cnt = vx.VectorFilterF64().apply_gt(x, 1.0)
        ref = np.zeros((n + 63) // 64, np.uint64)
        rc = o.O.vo_bitmap_f64(o.ptr(x), n, 4, 1.0, 0.0, o.ptr(ref))
        assert cnt == rc == int((x > 1.0).sum())
        assert np.array_equal(words, ref)


def test_hash_aggregator_group_count(gpu):
    from voxel_amd import voxel_py as vx
    rng = np.random.default_rng(2)
    agg = vx.HashAggregator(vx.Arena())
    agg.init(128)
    seen = set()
    for _ in range(4):
        k = rng.integers(0, 700, 20_000).astype(np.uint32)
        v = rng.standard_normal(20_000)
        agg.accumulate(k, v)
        seen |= set(k.tolist())
    assert agg.group_count() == len(seen)
    agg.init(16)
    assert agg.group_count() == 0


def test_sort_ascending_and_topk(gpu):
    from voxel_amd import voxel_py as vx
    rng = np.random.default_rng(3)
    x = np.round(rng.standard_normal(50_001) * 100) / 4
    x[::97] = -0.0
    x[::89] = 0.0
    idx = vx.sort_ascending(x)
    ref = np.zeros(x.size, np.uint32)
    o.O.vo_mergesort_f64(o.ptr(x), x.size, 0, o.ptr(ref))
    assert np.array_equal(idx, ref)
    for k, largest in ((1, True), (10, False), (1000, True)):
        top = vx.topk_select(x, k, largest)
        want = np.sort(x)[::-1][:k] if largest else np.sort(x)[:k]
        assert np.array_equal(top, want)


def test_hash_join_u32_pairs_and_contains(gpu):
    from voxel_amd import voxel_py as vx
    rng = np.random.default_rng(4)
    build = rng.integers(0, 5000, 3000).astype(np.uint32)            # duplicate build keys
    probe = rng.integers(0, 8000, 20_000).astype(np.uint32)
    j = vx.HashJoin(vx.Arena())
    j.build(build)
    left, right, m = j.probe(probe)
    b64, p64 = build.astype(np.int64), probe.astype(np.int64)
    lo, ro = np.zeros(probe.size, np.uint32), np.zeros(probe.size, np.uint32)
    mref = o.O.vo_hashjoin_i64(o.ptr(b64), b64.size, None, o.ptr(p64), p64.size, o.ptr(lo), o.ptr(ro), None)
    assert m == mref
    assert np.array_equal(left, lo[:m]) and np.array_equal(right, ro[:m])
    assert j.contains(int(build[7])) and not j.contains(4_000_000_000)


def test_voxel_query_group_sort_topk(gpu):
    from voxel_amd import voxel_query as vq
    rng = np.random.default_rng(5)
    x = rng.random(10_000) * 1000
    keys = rng.integers(0, 37, x.size).astype(np.uint32)
    assert vq.from_numpy(x).group_by(keys).run() == 37
    idx = vq.from_numpy(x).sort().run()
    assert np.array_equal(idx, np.argsort(x, kind="stable").astype(np.uint32))
    top = vq.from_numpy(x).topk(5).run()
    assert np.array_equal(top, np.sort(x)[::-1][:5])


def test_import_voxel_py_as_the_reference_callers_do(gpu, golden, x42):
    """`import voxel_py as _vx` (bindings/voxel_query.py:8) resolves to this
    backend: a program built with _vx.Instruction (the words voxel_query's
    sum() emits, bindings/voxel_query.py:74-110) runs on _vx.EngineF64 and
    gives the reference's bits; _vx.jit_run gives the JIT's golden result
    within the reference's own 1e-12 criterion (validate.cpp:68)."""
    import voxel_py as _vx
    I = _vx.Instruction
    code = [I.mov(4, 4).raw, I.vload(0, 1, 0, 0).raw, I.vfilter_gt(1, 0, 3).raw, I.vsum(5, 1).raw,
            I.addf(0, 0, 5).raw, I.add(1, 1, 4).raw, I.cmp(1, 2).raw, I.jnz(-6).raw, I.halt().raw]
    assert code == [int(w, 16) for w in golden["programs"]["vq_sum"]]
    e = _vx.EngineF64()
    e.add_segment(x42)
    e.set_scalar(2, x42.size)
    e.set_scalar_f64(3, 500.0)
    e.load_program(code)
    e.run()
    exact = float(o.relational(x42, 4, 500.0)["sum"])
    got = e.get_scalar_f64(0)
    assert abs(got - exact) <= 2 * o.ulp(exact)
    assert e.get_scalar(1) == x42.size
    ref = o.bits_f64(int(golden["seed42"]["vq_sum"]["R0"], 16))
    assert abs(got - ref) <= 1e-12 * abs(ref)
    fast = [int(w, 16) for w in golden["programs"]["fast"]]
    j = _vx.jit_run(fast, x42, 500.0, x42.size)
    ref = o.bits_f64(int(golden["seed42"]["jit"]["result"], 16))
    assert abs(j - ref) <= 1e-12 * abs(ref) and abs(j - exact) <= 2 * o.ulp(exact)

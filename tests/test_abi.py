"""C-ABI surface checks that need no GPU: the library loads, exports every symbol
include/voxel_c.h declares (the 64 of the reference's bindings/voxel_c.h among
them), encodes instructions like the reference, lowers the reference's programs
to the right plan kinds, and the Python surfaces emit the reference's words.
"""
from __future__ import annotations

import ctypes as C
import re

import numpy as np
import pytest

from oracle import oracle as o
from voxel_amd import _lib
from voxel_amd._lib import lib

REF_ABI = """voxel_engine_create_f64 voxel_engine_destroy voxel_engine_add_segment voxel_engine_load_program
voxel_engine_run voxel_engine_run_parallel voxel_engine_set_scalar voxel_engine_get_scalar
voxel_engine_set_scalar_f64 voxel_engine_get_scalar_f64 voxel_engine_k_lanes voxel_engine_create_i64
voxel_engine_destroy_i64 voxel_engine_add_segment_i64 voxel_engine_load_program_i64 voxel_engine_run_i64
voxel_engine_set_scalar_i64 voxel_engine_get_scalar_i64 voxel_jit_run voxel_instr_vload voxel_instr_vfilter_gt
voxel_instr_vfilter_ge voxel_instr_vfilter_lt voxel_instr_vfilter_le voxel_instr_vsum voxel_instr_vred_min
voxel_instr_vred_max voxel_instr_vcount voxel_instr_addf voxel_instr_add voxel_instr_sub voxel_instr_mul
voxel_instr_mov voxel_instr_movr voxel_instr_cmp voxel_instr_jnz voxel_instr_jz voxel_instr_jmp voxel_instr_halt
voxel_null_create voxel_null_destroy voxel_null_is_null voxel_null_set_null voxel_null_set_valid
voxel_null_null_count voxel_null_valid_count voxel_agg_create voxel_agg_destroy voxel_agg_init
voxel_agg_accumulate voxel_agg_group_count voxel_sort_ascending voxel_sort_descending voxel_topk
voxel_dict_create voxel_dict_destroy voxel_dict_build voxel_dict_cardinality voxel_dict_decode_batch
voxel_pool_create voxel_pool_destroy voxel_pool_wait_all voxel_pool_thread_count voxel_last_error""".split()


def test_reference_abi_is_64_symbols_and_all_exported():
    assert len(REF_ABI) == 64
    declared = _lib.header_symbols()
    for s in REF_ABI:
        assert s in declared, s
    for s in declared:
        assert hasattr(lib, s), "declared but not exported: " + s


def words(golden, name):
    return [int(w, 16) for w in golden["programs"][name]]


def test_instruction_encoders_match_reference_words(golden):
    fast = words(golden, "fast")
    got = [lib.voxel_instr_vload(0, 1, 0, 0), lib.voxel_instr_vfilter_gt(1, 0, 3), lib.voxel_instr_vsum(5, 1),
           lib.voxel_instr_addf(0, 0, 5), lib.voxel_instr_add(1, 1, 4), lib.voxel_instr_cmp(1, 2),
           lib.voxel_instr_jnz(-6), lib.voxel_instr_halt()]
    assert got == fast
    assert lib.voxel_instr_mov(4, 4) == words(golden, "vq_sum")[0]
    assert lib.voxel_instr_vcount(5, 1) == words(golden, "vq_count")[3]
    assert lib.voxel_instr_vred_min(5, 1) == words(golden, "vq_min")[3]
    assert lib.voxel_instr_vred_max(5, 1) == words(golden, "vq_max")[3]
    assert lib.voxel_instr_movr(0, 5) == words(golden, "vq_min")[6]
    if o.R is not None:
        rng = np.random.default_rng(0)
        for _ in range(200):
            rd, ra, rb = (int(v) for v in rng.integers(0, 16, 3))
            imm = int(rng.integers(-2048, 2048))
            assert lib.voxel_instr_sub(rd, ra, rb) == o.R.vr_encode(0x21, rd, ra, rb, 0)
            assert lib.voxel_instr_mul(rd, ra, rb) == o.R.vr_encode(0x22, rd, ra, rb, 0)
            assert lib.voxel_instr_mov(rd, imm) == o.R.vr_encode(0x10, rd, 0, 0, imm & 0xFFF)
            assert lib.voxel_instr_jz(imm) == o.R.vr_encode(0x51, 0, 0, 0, imm & 0xFFF)
            assert lib.voxel_instr_jmp(imm) == o.R.vr_encode(0x50, 0, 0, 0, imm & 0xFFF)
            assert lib.voxel_instr_vfilter_lt(rd, ra, rb) == o.R.vr_encode(0xC3, rd, ra, rb, 0)
            assert lib.voxel_instr_vfilter_le(rd, ra, rb) == o.R.vr_encode(0xC4, rd, ra, rb, 0)
            assert lib.voxel_instr_vfilter_ge(rd, ra, rb) == o.R.vr_encode(0xC6, rd, ra, rb, 0)


def lowered(code, i64=False):
    e = lib.voxel_engine_create_i64() if i64 else lib.voxel_engine_create_f64()
    c = np.array(code, np.uint32)
    (lib.voxel_engine_load_program_i64 if i64 else lib.voxel_engine_load_program)(e, c.ctypes.data, c.size)
    k = lib.voxel_gpu_engine_lowered(e)
    (lib.voxel_engine_destroy_i64 if i64 else lib.voxel_engine_destroy)(e)
    return k


def test_lowering_classifies_reference_programs(golden):
    assert lowered(words(golden, "fast")) == 1                       # DetectFastPath matches
    for k in ("vq_sum", "vq_count", "vq_min", "vq_max", "vq_sum_lt"):
        assert lowered(words(golden, k)) == 2, k                    # MOV prefix -> Tier-0 loop
    assert lowered(words(golden, "i64_add"), i64=True) == 2
    # the fast-path matcher ignores what follows word 6 (engine.hpp:84-107)
    assert lowered(words(golden, "fast")[:7] + [0x00000002]) == 1
    # programs that are not scan loops are refused
    assert lowered([0x00000001]) == 0
    assert lowered(words(golden, "vq_sum")[:-2] + [lib.voxel_instr_jnz(-5), 1]) == 0
    bad = words(golden, "vq_sum")
    bad[4] = lib.voxel_instr_addf(0, 0, 0)                          # accumulate reads acc twice
    assert lowered(bad) == 0


def test_lowering_accepts_hash_sort_join_words():
    """Straight-line programs of HASH_*, SORT_*, JOIN_* and PARTITION_HASH
    words (with scalar and AGG_* words) lower to the straight-line plan;
    SERIALIZE / DESERIALIZE / WINDOW_ROW / WINDOW_RANGE do not."""
    from tests import tier0_gen as tg
    for op in tg.REL_OPS:
        assert lowered([tg.w(0x10, rd=13, imm=5), tg.w(op, rd=2, ra=0, rb=1), tg.w(0xE1, rd=4, ra=0, rb=1), 1]) == 3
        assert lowered([tg.w(op, rd=2, ra=0, rb=13, imm=7), 1], i64=True) == 3
    for op in (0xFB, 0xFC, 0xFE, 0xFF):
        assert lowered([tg.w(op, rd=2, ra=0, rb=1), 1]) == 0


def test_lowering_accepts_filter_chains_and_several_reductions():
    from tests import tier0_gen as tg
    rng = np.random.default_rng(5)
    for trial in range(30):
        code, spec = tg.make_program(rng, trial % 2 == 0, allow_ne=True)
        assert lowered(list(code), i64=trial % 2 == 1) == 2, spec
    W = tg.w
    base = [W(0x10, rd=4, imm=4), W(0x70, rd=0, ra=1)]
    tail = [W(0x20, rd=1, ra=1, rb=4), W(0x40, ra=1, rb=2)]
    def loop(body):
        words = base + body + tail
        return words + [W(0x52, imm=(1 - len(words)) & 0xFFF), W(0x01)]
    f1 = W(0xC5, rd=1, ra=0, rb=3)
    f2 = W(0xC3, rd=2, ra=1, rb=6)
    f3 = W(0xC6, rd=3, ra=2, rb=3)
    red = [W(0xD0, rd=5, ra=2), W(0x2A, rd=0, ra=0, rb=5)]
    assert lowered(loop([f1, f2] + red)) == 2
    assert lowered(loop([f1, f2, f3, W(0xD0, rd=5, ra=3), W(0x2A, rd=0, ra=0, rb=5)])) == 0   # three deep
    assert lowered(loop([f1, W(0xD0, rd=5, ra=7), W(0x2A, rd=0, ra=0, rb=5)])) == 0          # V7 undefined
    assert lowered(loop([f1, W(0xC3, rd=1, ra=0, rb=6)] + red)) == 0                          # V1 redefined
    two = [W(0xD0, rd=5, ra=1), W(0x2A, rd=0, ra=0, rb=5), W(0xD7, rd=7, ra=1), W(0x20, rd=0, ra=0, rb=7)]
    assert lowered(loop([f1] + two)) == 0                                                     # shared accumulator
    two[3] = W(0x20, rd=8, ra=8, rb=7)
    assert lowered(loop([f1] + two)) == 2
    assert lowered(loop([W(0xC5, rd=1, ra=0, rb=5), W(0xD0, rd=5, ra=1), W(0x2A, rd=0, ra=0, rb=5)])) == 0  # threshold written


def test_lowering_accepts_vector_arithmetic():
    from tests import tier0_gen as tg
    rng = np.random.default_rng(6)
    for trial in range(30):
        is_f64 = trial % 2 == 0
        code, spec = tg.make_expr_program(rng, is_f64, nops=int(rng.integers(1, 9)))
        assert lowered(list(code), i64=not is_f64) == 2, spec
    W = tg.w
    base = [W(0x10, rd=4, imm=4), W(0x70, rd=0, ra=1)]
    tail = [W(0x20, rd=1, ra=1, rb=4), W(0x40, ra=1, rb=2)]
    def loop(body):
        words = base + body + tail
        return words + [W(0x52, imm=(1 - len(words)) & 0xFFF), W(0x01)]
    red = [W(0xD0, rd=5, ra=2), W(0x2A, rd=0, ra=0, rb=5)]
    mul = W(0x92, rd=1, ra=0, rb=3)                                  # V1 = V0 * R3
    assert lowered(loop([mul, W(0xC5, rd=2, ra=1, rb=6)] + red)) == 2
    # filters three deep behind arithmetic are evaluated per row
    chain = [mul, W(0xC5, rd=2, ra=1, rb=6), W(0xC3, rd=3, ra=2, rb=6), W(0xC6, rd=4, ra=3, rb=6)]
    assert lowered(loop(chain + [W(0xD0, rd=5, ra=4), W(0x2A, rd=0, ra=0, rb=5)])) == 2
    assert lowered(loop([W(0x80, rd=2, ra=0, rb=7)] + red)) == 0                   # V7 undefined
    assert lowered(loop([mul, W(0x80, rd=1, ra=0, rb=0)] + red)) == 0             # V1 redefined
    assert lowered(loop([W(0x92, rd=2, ra=0, rb=5)] + red)) == 0                  # scalar operand written by the loop
    assert lowered(loop([W(0x8F, rd=2, ra=0, rb=0)] + red)) == 2                  # VPOW
    # lane words: VREVERSE, VROTATE, VSLIDE by R6, VSHUFFLE, VPERMUTE, VINSERT R3 into lane 2, VGATHER
    for lw in (W(0x7D, rd=2, ra=0), W(0x7E, rd=2, ra=0, imm=1), W(0x7F, rd=2, ra=0, rb=6),
               W(0x7C, rd=2, ra=0, rb=0, imm=3), W(0x7B, rd=2, ra=0, rb=0), W(0x7A, rd=2, ra=0, rb=3, imm=2),
               W(0x72, rd=2, ra=0, imm=(1 << 8) | 0)):
        assert lowered(loop([lw] + red)) == 2, hex(lw)
    # VBLEND needs a mask written earlier in the iteration; VGATHER needs every lane
    blend = W(0xC7, rd=2, ra=0, rb=1, imm=3)
    assert lowered(loop([W(0xC5, rd=1, ra=0, rb=6, imm=3), blend] + red)) == 2
    assert lowered(loop([W(0xC5, rd=1, ra=0, rb=6, imm=2), blend] + red)) == 0
    assert lowered(loop([W(0x72, rd=2, ra=0, imm=(1 << 8) | 3)] + red)) == 0
    # VSTORE at the loop offset, lanes 1..4, with or without a reduction
    store = W(0x71, rd=0, ra=1, imm=(2 << 8) | 3)
    v2 = W(0x92, rd=2, ra=0, rb=3)
    assert lowered(loop([v2, store] + red)) == 2 and lowered(loop([v2, store])) == 2
    assert lowered(loop([v2, W(0x71, rd=0, ra=1, imm=(2 << 8) | 5)] + red)) == 0    # past the register
    assert lowered(loop([v2, W(0x71, rd=0, ra=3, imm=2 << 8)] + red)) == 0          # another offset register
    assert lowered(loop([W(0x72, rd=2, ra=0, imm=2 << 8), W(0x71, rd=0, ra=1, imm=2 << 8)] + red)) == 0
    segs = [W(0x72, rd=2, ra=0, imm=1 << 8), W(0x72, rd=3, ra=0, imm=2 << 8), W(0x72, rd=4, ra=0, imm=3 << 8)]
    assert lowered(loop(segs[:2] + [W(0xD0, rd=5, ra=3), W(0x2A, rd=0, ra=0, rb=5)])) == 2
    assert lowered(loop(segs + [W(0xD0, rd=5, ra=4), W(0x2A, rd=0, ra=0, rb=5)])) == 0
    for op in (0x83, 0x84, 0x93, 0x94):                                           # division and modulo
        div = loop([W(op, rd=2, ra=0, rb=3 if op >= 0x90 else 0)] + red)
        assert lowered(div) == 2 and lowered(div, i64=True) == 2


def test_jit_lowering_matches_fusion_detector(golden):
    def ok(code):
        c = np.array(code, np.uint32)
        return lib.voxel_gpu_jit_lowerable(c.ctypes.data, c.size)
    assert ok(words(golden, "fast")) == 1
    assert ok(words(golden, "vq_sum")) == 1                         # MOV prefix still fuses in the JIT
    assert ok(words(golden, "vq_count")) == 0                       # VCOUNT never fuses
    assert "JIT lowering" in _lib.last_error()
    assert ok(words(golden, "vq_sum_lt")) == 0                      # generic VFILTER never fuses


def test_voxel_query_emits_reference_words(golden):
    from voxel_amd import voxel_query as vq
    x = np.arange(8, dtype=np.float64)
    for name, build in (("vq_sum", lambda q: q.filter_gt(1).sum()), ("vq_count", lambda q: q.filter_gt(1).count()),
                        ("vq_min", lambda q: q.filter_gt(1).min()), ("vq_max", lambda q: q.filter_gt(1).max()),
                        ("vq_sum_lt", lambda q: q.filter_lt(1).sum())):
        q = build(vq.from_numpy(x))
        code = [vq._vx.Instruction.mov(4, 4).raw] + q.code + [1]
        assert code == words(golden, name), name


def test_voxel_py_instruction_words(golden):
    from voxel_amd.voxel_py import Instruction as I
    assert I.vfilter(1, 0, 3, 2).raw == words(golden, "vq_sum_lt")[2]
    assert I.jle(2).raw == words(golden, "vq_min")[5] and I.jge(2).raw == words(golden, "vq_max")[5]
    assert I.jnz(-8).raw == words(golden, "vq_min")[9]


def test_host_utilities_match_reference_behaviour():
    nb = lib.voxel_null_create(64)                                  # c_test.cpp Test 3
    for r in (10, 20, 30):
        lib.voxel_null_set_null(nb, r)
    assert lib.voxel_null_null_count(nb) == 3 and lib.voxel_null_valid_count(nb) == 61
    assert lib.voxel_null_is_null(nb, 20) == 1 and lib.voxel_null_is_null(nb, 21) == 0
    lib.voxel_null_set_valid(nb, 20)
    assert lib.voxel_null_null_count(nb) == 2
    lib.voxel_null_destroy(nb)
    nb = lib.voxel_null_create(70)
    assert lib.voxel_null_null_count(nb) == 0 and lib.voxel_null_valid_count(nb) == 70
    lib.voxel_null_destroy(nb)
    d = lib.voxel_dict_create()
    x = np.array([3.0, 1.0, 3.0, -0.0, 2.0, 0.0, 1.0], np.float64)
    lib.voxel_dict_build(d, x.ctypes.data, x.size)
    assert lib.voxel_dict_cardinality(d) == 4
    out = np.zeros(x.size, np.float64)
    lib.voxel_dict_decode_batch(d, out.ctypes.data, 0, x.size)
    assert out.tolist() == [3.0, 1.0, 3.0, -0.0, 2.0, -0.0, 1.0]     # equal values share the first one
    lib.voxel_dict_destroy(d)
    p = lib.voxel_pool_create(3)
    assert lib.voxel_pool_thread_count(p) == 3
    lib.voxel_pool_wait_all(p)
    lib.voxel_pool_destroy(p)
    assert lib.voxel_engine_k_lanes() == 4


def test_engine_registers_are_host_state():
    e = lib.voxel_engine_create_f64()
    lib.voxel_engine_set_scalar_f64(e, 3, 500.0)
    assert lib.voxel_engine_get_scalar(e, 3) == o.f64_bits(500.0)
    lib.voxel_engine_set_scalar(e, 64 + 7, 9)                       # index masked to 0x3F
    assert lib.voxel_engine_get_scalar(e, 7) == 9
    lib.voxel_engine_destroy(e)


def test_no_gpu_means_loud_failure():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from voxel_amd import voxel_py as vx
    e = vx.EngineF64()
    x = np.arange(8, dtype=np.float64)
    e.add_segment(x)
    e.set_scalar(2, 8)
    e.load_program([lib.voxel_instr_vload(0, 1, 0, 0), lib.voxel_instr_vfilter_gt(1, 0, 3),
                    lib.voxel_instr_vsum(5, 1), lib.voxel_instr_addf(0, 0, 5), lib.voxel_instr_add(1, 1, 4),
                    lib.voxel_instr_cmp(1, 2), lib.voxel_instr_jnz(-6), lib.voxel_instr_halt()])
    with pytest.raises(RuntimeError, match="no usable HIP device"):
        e.run()
    assert e.get_scalar(0) == 0 and e.get_scalar(1) == 0            # registers untouched


def test_lowering_accepts_window_word_loops():
    from tests import tier0_gen as tg
    rng = np.random.default_rng(7)
    for wkind in range(1, 6):
        code, spec = tg.make_window_program(rng, wkind, 8 if wkind != 5 else 0, nops=2)
        assert lowered(list(code)) == 2, (wkind, spec)
    W = tg.w
    body = [W(0xD0, rd=5, ra=0), W(0x2A, rd=0, ra=0, rb=5)]
    def loop(load, extra=()):
        words = [load] + body + list(extra) + [W(0x40, ra=1, rb=2)]
        return words + [W(0x52, imm=(0 - len(words)) & 0xFFF), W(0x01)]
    assert lowered(loop(W(0xDE, rd=0, ra=1, imm=20))) == 2
    assert lowered(loop(W(0xDE, rd=0, ra=1, imm=0))) == 0              # window 0: the word is a no-op
    assert lowered(loop(W(0x97, rd=0, ra=1, imm=1))) == 0              # WSTD needs a window of 2 or more
    assert lowered(loop(W(0x99, rd=0, ra=1, imm=5 | (3 << 5)))) == 2   # WQUANTILE, window 5, q = 3
    assert lowered(loop(W(0x99, rd=0, ra=1, imm=1))) == 0              # WQUANTILE window 1: a no-op


def test_import_voxel_py_resolves_to_the_backend():
    """`import voxel_py` (the reference's module name, bindings/voxel_bindings.cpp:25,
    imported as `_vx` by bindings/voxel_query.py:8 and voxel_market.py:13)
    yields voxel_amd.voxel_py, with every name those callers use and the
    pybind module's enums and constants."""
    import voxel_py as _vx
    from voxel_amd import voxel_py as impl
    assert _vx is impl
    for name in ("Instruction", "EngineF64", "EngineI64", "CMP_EQ", "CMP_NE", "CMP_LT", "CMP_LE", "CMP_GT",
                 "CMP_GE", "HashAggregator", "Arena", "sort_ascending", "topk_select", "HashJoin",
                 "VectorFilterF64", "jit_run", "jit_compile", "NullBitmap", "SegmentF64", "cpu_info"):
        assert hasattr(_vx, name), name
    assert (_vx.CMP_EQ, _vx.CMP_NE, _vx.CMP_LT, _vx.CMP_LE, _vx.CMP_GT, _vx.CMP_GE) == (0, 1, 2, 3, 4, 5)
    assert _vx.Float64 == _vx.DataType.Float64 == 11 and _vx.Timestamp == 32 and _vx.VSUM == 0xD0
    assert _vx.EngineF64.k_lanes == 4 and _vx.EngineI64.k_lanes == 4
    seg = _vx.SegmentF64(np.arange(5.0))
    assert len(seg) == seg.size == 5 and seg[3] == 3.0
    nb = _vx.NullBitmap(100)
    nb.set_null(7)
    assert nb.is_null(7) and not nb.is_null(8) and nb.null_count() == 1 and nb.valid_count() == 99
    info = _vx.cpu_info()
    assert set(info) == {"logical_cores", "physical_cores", "l1_cache", "l2_cache", "l3_cache", "has_avx2",
                         "has_avx512", "has_fma"}
    with pytest.raises(AttributeError, match="StreamingEngine"):
        _vx.StreamingEngine

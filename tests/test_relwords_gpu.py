"""GPU parity of the hash / sort / join words of straight-line programs
(vx_relwords.hip; the reference's Engine<T> handlers, exec/engine.hpp:
704-842) through the C ABI (voxel_py.EngineF64 / EngineI64).

Oracle: vo_relw_run, pinned against the reference engine on generated
programs (tests/test_oracle.py).  Bar: every register and every segment row
after the run bit-identical with the oracle -- hash results, sorted rows,
partial-sort rows (libstdc++'s heap order, NaN and -0.0 included) and join
counts -- on programs over up to three segments of up to 2^20 rows; sorts of
2^22 rows checked for order and multiset; refusals leave registers, slots
and rows as they were.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import oracle as o

pytestmark = pytest.mark.gpu


def _segments(rng, nseg, is_f64, nmax, full_order):
    segs = []
    for _ in range(nseg):
        n = int(rng.integers(1, nmax))
        if is_f64:
            x = np.round(rng.standard_normal(n) * 8, int(rng.integers(0, 2)))
            if full_order:
                x = x + 0.0
            else:
                x[rng.integers(0, n, 3)] = np.nan
        else:
            x = rng.integers(-50, 50, n).astype(np.int64)
            x[rng.integers(0, n, 2)] = rng.integers(-2**62, 2**62, 2)
        segs.append(x)
    return segs


def _oracle_run(segs, is_f64, code, regs, state=None):
    mine = [a.view(np.uint64).copy() for a in segs]
    p, n = o.seg_ptrs(mine)
    st = state or o.O.vo_relw_create()
    r = regs.copy()
    rc = o.O.vo_relw_run(st, o.ptr(p), o.ptr(n), len(segs), int(is_f64), o.ptr(code), code.size, o.ptr(r))
    if state is None:
        o.O.vo_relw_destroy(st)
    return rc, r, mine


def _engine(segs, is_f64, regs):
    from voxel_amd import voxel_py as vx
    e = vx.EngineF64() if is_f64 else vx.EngineI64()
    for a in segs:
        e.add_segment(a)
    for r in range(64):
        if regs[r]:
            e.set_scalar(r, int(regs[r]))
    return e


@pytest.mark.parametrize("is_f64", [True, False])
def test_relational_words_random_programs(gpu, is_f64):
    from tests import tier0_gen as tg
    rng = np.random.default_rng(211 if is_f64 else 212)
    for trial in range(40):
        nseg = int(rng.integers(1, 4))
        full_order = trial % 3 != 0
        nmax = [300, 5000, 1 << 16, 1 << 20][trial % 4]
        segs = _segments(rng, nseg, is_f64, nmax, full_order)
        ops = tg.REL_OPS if full_order else [op for op in tg.REL_OPS if op not in (0xF2, 0xF3, 0xF7)]
        code = tg.rel_program(rng, nseg, k=int(rng.integers(1, 10)), ops=ops, kmax=min(a.size for a in segs))
        regs = np.zeros(64, np.uint64)
        for r in tg.KEY_REGS:
            src = segs[int(rng.integers(0, nseg))]
            regs[r] = src.view(np.uint64)[int(rng.integers(0, src.size))] if rng.random() < 0.8 else \
                np.uint64(rng.integers(0, 2**63))
        for r in tg.K_REGS:
            regs[r] = int(rng.integers(0, 5000)) if rng.random() < 0.8 else 2**64 - 1
        rc, want, want_segs = _oracle_run(segs, is_f64, code, regs)
        assert rc == 1
        e = _engine(segs, is_f64, regs)
        e.load_program([int(c) for c in code])
        e.run()
        got = np.array([e.get_scalar(r) for r in range(64)], np.uint64)
        ctx = (trial, [hex(int(c)) for c in code])
        assert np.array_equal(got, want), (ctx, np.nonzero(got != want))
        for a, b in zip(segs, want_segs):                 # the words reorder the caller's rows in place
            assert np.array_equal(a.view(np.uint64), b), ctx


def test_sort_and_partial_sort_at_scale(gpu):
    """2^22 rows: SORT_ASC / SORT_DESC give the sorted multiset; SORT_TOPK /
    SORT_BOTTOMK with k = 4095 equal the oracle's libstdc++ restatement row
    for row; JOIN_HASH / JOIN_MERGE / JOIN_SEMI / JOIN_ANTI of two 2^21-row
    segments equal the oracle's counts."""
    from tests import tier0_gen as tg
    rng = np.random.default_rng(213)
    n = 1 << 22
    x = np.round(rng.standard_normal(n) * 1e4, 1) + 0.0
    for op in (0xF2, 0xF3):
        a = x.copy()
        e = _engine([a], True, np.zeros(64, np.uint64))
        e.load_program([tg.w(op, rd=5, ra=0), 1])
        e.run()
        assert e.get_scalar(5) == n
        want = np.sort(x) if op == 0xF2 else np.sort(x)[::-1]
        assert np.array_equal(a, want)
    for op, is_f64 in ((0xF4, True), (0xF5, True), (0xF4, False), (0xF5, False)):
        a = (x.copy() if is_f64 else (x * 10).astype(np.int64))
        a[::97] = -0.0 if is_f64 else a[::97]
        regs = np.zeros(64, np.uint64)
        regs[13] = 4095
        code = np.array([tg.w(op, rd=5, ra=0, rb=13, imm=4095), 1], np.uint32)
        rc, want, (want_rows,) = _oracle_run([a], is_f64, code, regs)
        e = _engine([a], is_f64, regs)
        e.load_program([int(c) for c in code])
        e.run()
        assert e.get_scalar(5) == 4095 and np.array_equal(a.view(np.uint64), want_rows), (op, is_f64)
    l = rng.integers(0, 1 << 19, 1 << 21).astype(np.int64)
    r = rng.integers(0, 1 << 19, 1 << 21).astype(np.int64)
    for op in (0xF6, 0xF7, 0xF9, 0xFA):
        code = np.array([tg.w(op, rd=6, ra=0, rb=1), 1], np.uint32)
        rc, want, _ = _oracle_run([l, r], False, code, np.zeros(64, np.uint64))
        e = _engine([l, r], False, np.zeros(64, np.uint64))
        e.load_program([int(c) for c in code])
        e.run()
        assert e.get_scalar(6) == want[6], hex(op)


def test_hash_tables_persist_and_failures_change_nothing(gpu):
    """Hash tables live in the engine's 16 slots across runs (HashTables_,
    engine.hpp:704-744); a program that fails part-way (SORT over NaN,
    SORT_TOPK past the segment) leaves registers, hash slots and rows as they
    were."""
    from tests import tier0_gen as tg
    rng = np.random.default_rng(214)
    a = np.round(rng.standard_normal(20000) * 50) + 0.0
    b = np.round(rng.standard_normal(3000) * 5) + 0.0
    bad = a.copy()
    bad[5] = np.nan
    small = b[:10].copy()
    e = _engine([a, b, bad, small], True, np.zeros(64, np.uint64))
    key = o.f64_bits(float(a[1234]))
    e.set_scalar(8, key)
    e.load_program([tg.w(0xEE, rd=1, ra=0), 1])              # slot 1 over segment 0
    e.run()
    nd = e.get_scalar(1)
    assert nd == np.unique(a).size
    e.load_program([tg.w(0xEF, rd=4, ra=1, rb=8), 1])       # next run probes the kept table
    e.run()
    assert e.get_scalar(4) == int(np.argmax(a == a[1234]))
    before = bad.copy()
    for prog in ([tg.w(0xEE, rd=1, ra=1), tg.w(0xF2, rd=6, ra=2), 1],
                 [tg.w(0xEE, rd=1, ra=1), tg.w(0x10, rd=13, imm=2000), tg.w(0xF4, rd=6, ra=3, rb=13, imm=4000), 1]):
        e.load_program(prog)
        with pytest.raises(RuntimeError):
            e.run()
        assert e.get_scalar(1) == nd and e.get_scalar(6) == 0
        assert np.array_equal(bad.view(np.uint64), before.view(np.uint64))
        e.load_program([tg.w(0xEF, rd=4, ra=1, rb=8), 1])   # slot 1 still holds segment 0's table
        e.run()
        assert e.get_scalar(4) == int(np.argmax(a == a[1234]))


def test_sort_words_on_device_and_resident_segments(gpu):
    """A device segment is sorted where it lies in HBM; a resident host
    segment's HBM copy and the caller's rows both end sorted."""
    import torch
    from tests import tier0_gen as tg
    from voxel_amd._lib import lib, last_error
    rng = np.random.default_rng(215)
    x = np.round(rng.standard_normal(100_000) * 1e3) + 0.0
    t = torch.from_numpy(x.copy()).cuda()
    h = np.ascontiguousarray(x.copy())
    e = lib.voxel_engine_create_f64()
    try:
        lib.voxel_gpu_engine_add_device_segment(e, C.c_void_p(t.data_ptr()), t.numel())
        lib.voxel_engine_add_segment(e, h.ctypes.data, h.size)
        assert lib.voxel_gpu_engine_set_resident(e, 1, 1) == 0
        code = np.array([tg.w(0xF3, rd=5, ra=0), tg.w(0xF2, rd=6, ra=1), tg.w(0xE5, rd=7, ra=1, rb=2), 1], np.uint32)
        lib.voxel_engine_load_program(e, code.ctypes.data, code.size)
        lib.voxel_engine_run(e)
        assert last_error() == "", last_error()
        assert np.array_equal(t.cpu().numpy(), np.sort(x)[::-1])
        assert np.array_equal(h, np.sort(x))
        assert o.bits_f64(lib.voxel_engine_get_scalar(e, 7)) == np.min(x)     # AGG_FIRST reads the sorted HBM copy
    finally:
        lib.voxel_engine_destroy(e)

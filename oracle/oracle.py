"""ctypes view of the test oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module.  It exposes

* ``O``   -- the C restatement (oracle/voxel_oracle.c -> oracle/liboracle.so),
* ``R``   -- the reference's own C++ built by oracle/Makefile into
             oracle/_ref/libvoxel_ref.so (None when that build is absent),

plus numpy-level helpers.  The exact long accumulator is returned as digits and
turned into a Python ``Fraction`` here, so float results can be compared with
the exact value and rounded correctly.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(HERE, "liboracle.so")
_REF = os.path.join(HERE, "_ref", "libvoxel_ref.so")

P = C.c_void_p
U64, I64, U32, I32, U8, SZ, F64, F32, INT = (C.c_uint64, C.c_int64, C.c_uint32, C.c_int32,
                                             C.c_uint8, C.c_size_t, C.c_double, C.c_float, C.c_int)


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load(path):
    return C.CDLL(path) if os.path.exists(path) else None


if not os.path.exists(_LIB):
    build()
O = C.CDLL(_LIB)
R = _load(_REF)

_SIGS = {
    "vo_mt64_uniform_f64": (None, [U64, SZ, F64, F64, P]),
    "vo_mt64_uniform_i64": (None, [U64, SZ, I64, I64, P]),
    "vo_mt64_uniform_i32": (None, [U64, SZ, I32, I32, P]),
    "vo_acc_digits": (INT, []),
    "vo_exact_sum_f64": (None, [P, SZ, P]),
    "vo_rel_f64": (None, [P, P, SZ, INT, F64, INT, F64, P, P, P, P, P, P]),
    "vo_rel_f32": (None, [P, P, SZ, INT, F32, INT, F32, P, P, P, P, P, P]),
    "vo_rel_i64": (None, [P, P, SZ, INT, I64, INT, I64, P, P, P, P, P, P]),
    "vo_rel_i32": (None, [P, P, SZ, INT, I32, INT, I32, P, P, P, P, P, P]),
    "vo_fastpath_f64": (U64, [P, U64, U64, U64, F64, F64, U8, P]),
    "vo_fastpath_f32": (U64, [P, U64, U64, U64, F32, F32, U8, P]),
    "vo_fastpath_i64": (U64, [P, U64, U64, U64, I64, I64, U8, P]),
    "vo_fastpath_i32": (U64, [P, U64, U64, U64, I32, I32, U8, P]),
    "vo_tier0_f64": (U64, [P, U64, U64, U64, INT, F64, INT, INT, INT, U64, P]),
    "vo_tier0_i64": (U64, [P, U64, U64, U64, INT, I64, INT, INT, INT, U64, P]),
    "vo_tier0_multi_f64": (None, [P, U64, U64, U64, INT, INT, P, P, P, INT, P, P, P, P, P]),
    "vo_tier0_multi_i64": (None, [P, U64, U64, U64, INT, INT, P, P, P, INT, P, P, P, P, P]),
    "vo_scalar_run": (INT, [P, SZ, P]),
    "vo_tier0_expr_f64": (None, [P, U64, U64, U64, INT, INT, P, P, INT, P, P, P, P, P]),
    "vo_tier0_expr_i64": (None, [P, U64, U64, U64, INT, INT, P, P, INT, P, P, P, P, P]),
    "vo_tier0_expr_col_f64": (None, [P, U64, U64, U64, INT, INT, P, P, INT, P]),
    "vo_tier0_expr_col_i64": (None, [P, U64, U64, U64, INT, INT, P, P, INT, P]),
    "vo_expr_reset_masks": (None, []),
    "vo_expr_set_gather": (None, [INT, P, U64]),
    "vo_expr_set_store": (None, [INT, P, U64]),
    "vo_window_f64": (I64, [INT, P, SZ, P, SZ, INT, F64]),
    "vo_tier0_window_f64": (INT, [P, U64, U64, U64, INT, INT, INT, INT, P, P, INT, P, P, P, P, P]),
    "vo_nullable_f64": (None, [P, SZ, P, P, P, P]),
    "vo_dict_build_f64": (SZ, [P, SZ, P, P]),
    "vo_dict_build_i64": (SZ, [P, SZ, P, P]),
    "vo_dict_decode_f64": (None, [P, P, SZ, SZ, SZ, P]),
    "vo_dict_decode_i64": (None, [P, P, SZ, SZ, SZ, P]),
    "vo_rle_build_f64": (SZ, [P, SZ, P, P]),
    "vo_rle_build_i64": (SZ, [P, SZ, P, P]),
    "vo_rle_decode_f64": (None, [P, P, SZ, SZ, SZ, P]),
    "vo_rle_decode_i64": (None, [P, P, SZ, SZ, SZ, P]),
    "vo_delta_build_i64": (None, [P, SZ, P]),
    "vo_delta_build_i32": (None, [P, SZ, P]),
    "vo_delta_decode_i64": (None, [I64, P, SZ, SZ, SZ, P]),
    "vo_delta_decode_i32": (None, [I32, P, SZ, SZ, SZ, P]),
    "vo_bitpack_build": (SZ, [P, SZ, P, P]),
    "vo_bitpack_decode": (None, [P, SZ, U32, SZ, SZ, SZ, P]),
    "vo_for_build": (None, [P, SZ, P, P, P]),
    "vo_for_decode": (None, [F64, F64, P, SZ, SZ, SZ, P]),
    "vo_nullable_i64": (None, [P, SZ, P, P, P, P]),
    "vo_time_window_sum_f64": (I64, [P, P, SZ, I64, I64, P, SZ]),
    "vo_agg_create": (P, [INT]),
    "vo_agg_destroy": (None, [P]),
    "vo_straight_run": (INT, [P, P, U64, P, SZ, P]),
    "vo_jit_f64": (F64, [P, U64, U64, U64, U8, F64]),
    "vo_run_parallel_fast_f64": (F64, [P, U64, U32, U64, U64, F64, F64, U8]),
    "vo_bitmap_f64": (U64, [P, SZ, INT, F64, F64, P]),
    "vo_bitmap_f32": (U64, [P, SZ, INT, F32, F32, P]),
    "vo_bitmap_i64": (U64, [P, SZ, INT, I64, I64, P]),
    "vo_bitmap_i32": (U64, [P, SZ, INT, I32, I32, P]),
    "vo_bitmap_combine": (U64, [P, P, SZ, INT]),
    "vo_select_from_bitmap": (U64, [P, SZ, P]),
    "vo_compact_f64": (U64, [P, SZ, INT, F64, P]),
    "vo_compact_f32": (U64, [P, SZ, INT, F32, P]),
    "vo_compact_i64": (U64, [P, SZ, INT, I64, P]),
    "vo_compact_i32": (U64, [P, SZ, INT, I32, P]),
    "vo_radix_sort_u64": (None, [P, P, SZ]),
    "vo_mergesort_f64": (None, [P, SZ, INT, P]),
    "vo_hash_i64": (U64, [I64]),
    "vo_hashjoin_i64": (U64, [P, SZ, P, P, SZ, P, P, P]),
    "vo_groupby_f64": (SZ, [P, P, SZ, INT, SZ, P, P, P, P, P, P, P, P]),
    "vo_splitmix64": (U64, [U64]),
    "vo_hash_u64": (U64, [U64]),
    "vo_hash_element": (U64, [U64, INT]),
    "vo_value_key": (U64, [U64, INT, INT]),
    "vo_partial_sort": (None, [P, U64, U64, INT, INT]),
    "vo_join_count": (I64, [INT, P, U64, P, U64, INT]),
    "vo_relw_create": (P, []),
    "vo_relw_destroy": (None, [P]),
    "vo_relw_run": (INT, [P, P, P, INT, INT, P, SZ, P]),
}
for _n, (_r, _a) in _SIGS.items():
    _f = getattr(O, _n)
    _f.restype, _f.argtypes = _r, _a

_RSIGS = {
    "vr_mt_uniform_f64": (None, [U64, SZ, F64, F64, P]),
    "vr_mt_uniform_i64": (None, [U64, SZ, I64, I64, P]),
    "vr_mt_uniform_i32": (None, [U64, SZ, I32, I32, P]),
    "vr_engine_run_f64": (INT, [P, SZ, P, SZ, P]),
    "vr_engine_run_i64": (INT, [P, SZ, P, SZ, P]),
    "vr_engine_run_i32": (INT, [P, SZ, P, SZ, P]),
    "vr_engine_run_parallel_f64": (None, [P, SZ, P, SZ, P, U32, U8]),
    "vr_jit_compile": (P, [P, SZ, P]),
    "vr_jit_entry": (F64, [P, P, SZ, U64, F64]),
    "vr_jit_release": (None, [P]),
    "vr_bitmap_f64": (U64, [P, SZ, INT, F64, F64, P]),
    "vr_bitmap_f32": (U64, [P, SZ, INT, F32, F32, P]),
    "vr_bitmap_i64": (U64, [P, SZ, INT, I64, I64, P]),
    "vr_bitmap_i32": (U64, [P, SZ, INT, I32, I32, P]),
    "vr_column_filter": (U64, [P, P, SZ, INT]),
    "vr_select_from_bitmap": (U64, [P, SZ, P]),
    "vr_scalar_filter_gt_f64": (U64, [P, P, F64, SZ]),
    "vr_scalar_filter_gt_f32": (U64, [P, P, F32, SZ]),
    "vr_scalar_sum_f64": (F64, [P, SZ]),
    "vr_scalar_sum_f32": (F32, [P, SZ]),
    "vr_radix_sort_u64": (None, [P, P, SZ]),
    "vr_sort_f64": (None, [P, SZ, INT, P]),
    "vr_hash_join_i64": (U64, [P, SZ, P, P, SZ, P, P, P]),
    "vr_hash_agg_i32_f64": (SZ, [P, P, SZ, SZ, SZ, P, P, P, P, P, P, P]),
    "vr_hash_agg_u32_f64": (SZ, [P, P, SZ, SZ, SZ, P, P, P, P, P, P, P]),
    "vr_topk_f64": (None, [P, SZ, SZ, P, INT]),
    "vr_compound_f64": (None, [P, P, SZ, F64, F64, P]),
    "vr_encode": (U32, [U8, U8, U8, U8, C.c_uint16]),
    "vr_window_f64": (I64, [INT, P, SZ, P, SZ, INT, F64]),
    "vr_nullable_f64": (None, [P, SZ, P, P, P, P]),
    "vr_dict_build_f64": (SZ, [P, SZ, P, P]),
    "vr_dict_build_i64": (SZ, [P, SZ, P, P]),
    "vr_dict_decode_f64": (None, [P, SZ, P, SZ, SZ, SZ, P]),
    "vr_dict_decode_i64": (None, [P, SZ, P, SZ, SZ, SZ, P]),
    "vr_rle_build_f64": (SZ, [P, SZ, P, P]),
    "vr_rle_build_i64": (SZ, [P, SZ, P, P]),
    "vr_rle_decode_f64": (None, [P, P, SZ, SZ, SZ, P]),
    "vr_rle_decode_i64": (None, [P, P, SZ, SZ, SZ, P]),
    "vr_delta_build_i64": (None, [P, SZ, P, P]),
    "vr_delta_build_i32": (None, [P, SZ, P, P]),
    "vr_delta_decode_i64": (None, [I64, P, SZ, SZ, SZ, P]),
    "vr_delta_decode_i32": (None, [I32, P, SZ, SZ, SZ, P]),
    "vr_bitpack_build": (SZ, [P, SZ, P, P]),
    "vr_bitpack_decode": (None, [P, SZ, U32, SZ, SZ, SZ, P]),
    "vr_for_build": (None, [P, SZ, P, P, P]),
    "vr_for_decode": (None, [F64, F64, P, SZ, SZ, SZ, P]),
    "vr_nullable_i64": (None, [P, SZ, P, P, P, P]),
    "vr_time_window_sum_f64": (I64, [P, P, SZ, I64, I64, P, SZ]),
    "vr_engine_run_segs_f64": (INT, [P, P, INT, P, SZ, P]),
    "vr_engine_run_segs_i64": (INT, [P, P, INT, P, SZ, P]),
}
if R is not None:
    for _n, (_r, _a) in _RSIGS.items():
        _f = getattr(R, _n)
        _f.restype, _f.argtypes = _r, _a

NDIG = O.vo_acc_digits()
CMP = {"EQ": 0, "NE": 1, "LT": 2, "LE": 3, "GT": 4, "GE": 5, "TRUE": 6}


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


def digits_to_fraction(d) -> Fraction:
    v = 0
    for i in range(len(d) - 1, -1, -1):
        v = (v << 32) + int(d[i])
    return Fraction(v, 1 << 1074)


def exact_sum(x: np.ndarray) -> Fraction:
    x = np.ascontiguousarray(x, dtype=np.float64)
    d = np.zeros(NDIG, np.int64)
    O.vo_exact_sum_f64(ptr(x), x.size, ptr(d))
    return digits_to_fraction(d)


def mt_uniform_f64(seed: int, n: int, lo: float = 0.0, hi: float = 1000.0) -> np.ndarray:
    out = np.empty(n, np.float64)
    O.vo_mt64_uniform_f64(seed, n, lo, hi, ptr(out))
    return out


def mt_uniform_i64(seed: int, n: int, lo: int, hi: int) -> np.ndarray:
    out = np.empty(n, np.int64)
    O.vo_mt64_uniform_i64(seed, n, lo, hi, ptr(out))
    return out


def mt_uniform_i32(seed: int, n: int, lo: int, hi: int) -> np.ndarray:
    out = np.empty(n, np.int32)
    O.vo_mt64_uniform_i32(seed, n, lo, hi, ptr(out))
    return out


_TCODE = {np.dtype(np.float64): "f64", np.dtype(np.float32): "f32",
          np.dtype(np.int64): "i64", np.dtype(np.int32): "i32"}


def relational(a: np.ndarray, op_a: int, t_a, b: np.ndarray | None = None, op_b: int = 6, t_b=0):
    """COUNT / exact SUM / MIN / MAX of ``a`` over rows passing both predicates.

    Returns a dict: count (int), sum (Fraction for floats, wrapping int64 for
    ints), min, max (None when no non-NaN row passes)."""
    tc = _TCODE[a.dtype]
    fn = getattr(O, "vo_rel_" + tc)
    cnt = np.zeros(1, np.uint64)
    dig = np.zeros(NDIG, np.int64)
    mn = np.zeros(1, a.dtype)
    mx = np.zeros(1, a.dtype)
    isum = np.zeros(1, np.int64)
    anyv = np.zeros(1, np.int32)
    fn(ptr(a), ptr(b), a.size, op_a, t_a, op_b, t_b if b is not None else 0,
       ptr(cnt), ptr(dig), ptr(mn), ptr(mx), ptr(isum), ptr(anyv))
    isf = tc in ("f64", "f32")
    return {
        "count": int(cnt[0]),
        "sum": digits_to_fraction(dig) if isf else int(isum[0]),
        "min": mn[0] if anyv[0] else None,
        "max": mx[0] if anyv[0] else None,
    }


def f64_bits(x: float) -> int:
    return int(np.array([x], np.float64).view(np.uint64)[0])


def bits_f64(b: int) -> float:
    return float(np.array([b & (2**64 - 1)], np.uint64).view(np.float64)[0])


def ulp(x: float) -> float:
    return float(np.spacing(np.float64(abs(x)))) if x != 0 else 5e-324


def seg_ptrs(segs):
    """(array of segment pointers, array of row counts) for vo_relw_run /
    vr_engine_run_segs_*; the arrays must stay alive for the call."""
    p = np.array([a.ctypes.data for a in segs], np.uint64)
    n = np.array([a.size for a in segs], np.uint64)
    return p, n

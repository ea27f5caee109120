"""Relational filter + aggregate over device-resident columns
(voxel_gpu_scan in include/voxel_c.h).

``a``/``b`` are device buffers: torch CUDA tensors, or (pointer, length)
pairs from ``voxel_gpu_malloc``.  The compound predicate is
``op_a(a[i], t_a) AND op_b(b[i], t_b)`` as VectorFilter::Apply* +
ColumnFilter::AndWith select rows (ops/filter.hpp:42-288, 540-547); the
result is COUNT, SUM(a), MIN(a), MAX(a) over those rows.
"""
from __future__ import annotations

import ctypes as C
import struct

from ._lib import Agg, ScanDesc, check, lib

F64, F32, I64, I32 = 0, 1, 2, 3
EQ, NE, LT, LE, GT, GE, TRUE = 0, 1, 2, 3, 4, 5, 6
OPS = {"eq": EQ, "ne": NE, "lt": LT, "le": LE, "gt": GT, "ge": GE, "true": TRUE,
       "==": EQ, "!=": NE, "<": LT, "<=": LE, ">": GT, ">=": GE}

_TORCH_DT = {"torch.float64": F64, "torch.float32": F32, "torch.int64": I64, "torch.int32": I32}


def threshold_bits(dtype: int, value) -> int:
    """Bit pattern of ``value`` in the column type (f32/i32 in the low 32 bits)."""
    if dtype == F64:
        return struct.unpack("<Q", struct.pack("<d", float(value)))[0]
    if dtype == F32:
        return struct.unpack("<I", struct.pack("<f", float(value)))[0]
    if dtype == I64:
        return int(value) & 0xFFFFFFFFFFFFFFFF
    return int(value) & 0xFFFFFFFF


_NP_DT = {"float64": F64, "float32": F32, "int64": I64, "int32": I32}


def _dev(col):
    """(pointer, rows, dtype) of a DeviceArray or a torch CUDA tensor."""
    from .devarray import DeviceArray
    if isinstance(col, DeviceArray):
        dt = _NP_DT.get(col.dtype.name)
        if dt is None:
            raise TypeError("column dtype %s not supported" % col.dtype)
        return col.ptr, col.n, dt
    dt = _TORCH_DT.get(str(col.dtype))
    if dt is None:
        raise TypeError("column dtype %s not supported" % col.dtype)
    if not col.is_cuda or not col.is_contiguous():
        raise TypeError("columns must be contiguous device tensors")
    return col.data_ptr(), col.numel(), dt


def make_desc(a, op_a, t_a, b=None, op_b=TRUE, t_b=0, blocks: int = 0, unroll: int = 0,
              nontemporal: int = 0, layout: int = 0) -> ScanDesc:
    pa, n, dt = _dev(a)
    pb = None
    if b is not None:
        pb, nb, dtb = _dev(b)
        if nb != n or dtb != dt:
            raise ValueError("columns a and b must have the same length and dtype")
    op_a = OPS.get(op_a, op_a) if isinstance(op_a, str) else op_a
    op_b = OPS.get(op_b, op_b) if isinstance(op_b, str) else op_b
    d = ScanDesc()
    d.dtype, d.op_a, d.op_b, d.blocks = dt, op_a, op_b if b is not None else TRUE, blocks
    d.unroll, d.nontemporal, d.layout = unroll, nontemporal, layout
    d.a, d.b, d.n = pa, pb, n
    d.t_a, d.t_b = threshold_bits(dt, t_a), threshold_bits(dt, t_b)
    return d


def _result(dt: int, r: Agg) -> dict:
    out = {"count": int(r.count)}
    if dt in (F64, F32):
        out.update(sum=r.sum, sum_hi=r.sum_hi, sum_lo=r.sum_lo,
                   min=r.fmin if r.has_minmax else None, max=r.fmax if r.has_minmax else None)
    else:
        out.update(sum=int(r.isum), min=int(r.imin) if r.has_minmax else None,
                   max=int(r.imax) if r.has_minmax else None)
    return out


def _words(valid):
    """Device pointer of a validity-word tensor (int64/uint64, NullBitmap layout) or None."""
    if valid is None:
        return None
    from .devarray import DeviceArray
    if isinstance(valid, DeviceArray):
        if valid.dtype.itemsize != 8:
            raise TypeError("validity words must be 64-bit")
        return C.c_void_p(valid.ptr)
    if not valid.is_cuda or not valid.is_contiguous() or valid.element_size() != 8:
        raise TypeError("validity words must be a contiguous 64-bit device tensor")
    return C.c_void_p(valid.data_ptr())


def scan(a, op_a, t_a, b=None, op_b=TRUE, t_b=0, stream=None, blocks: int = 0, valid_a=None, valid_b=None) -> dict:
    """Runs one filter+aggregate and returns {count, sum, min, max}.  With
    validity words (bit i % 64 of word i // 64, 1 = valid) for a and/or b,
    rows that are null in either column are skipped (voxel_gpu_scan_valid)."""
    d = make_desc(a, op_a, t_a, b, op_b, t_b, blocks)
    r = Agg()
    st = None if stream is None else C.c_void_p(stream)
    if valid_a is not None or valid_b is not None:
        for v in (valid_a, valid_b):
            if v is not None and v.numel() < (d.n + 63) // 64:
                raise ValueError("validity words cover fewer rows than the column")
        if lib.voxel_gpu_scan_valid(C.byref(d), _words(valid_a), _words(valid_b), st, C.byref(r)) != 0:
            check("voxel_gpu_scan_valid")
    elif lib.voxel_gpu_scan(C.byref(d), st, C.byref(r)) != 0:
        check("voxel_gpu_scan")
    return _result(d.dtype, r)


def nullable_reduce(col, valid=None, stream=None) -> dict:
    """NullableSegment<T>::SumValid / MinValid / MaxValid / ValidCount /
    NullCountVal (data/nulls.hpp:241-420) of a device column."""
    from ._lib import NullableStats
    p, n, dt = _dev(col)
    if valid is not None and valid.numel() < (n + 63) // 64:
        raise ValueError("validity words cover fewer rows than the column")
    r = NullableStats()
    st = None if stream is None else C.c_void_p(stream)
    if lib.voxel_gpu_nullable_reduce(dt, C.c_void_p(p), _words(valid), n, C.byref(r), st) != 0:
        check("voxel_gpu_nullable_reduce")
    out = {"valid_count": int(r.valid_count), "null_count": int(r.null_count)}
    if dt in (F64, F32):
        out.update(sum=r.sum, min=r.min, max=r.max)
    else:
        out.update(sum=int(r.isum), min=int(r.imin), max=int(r.imax))
    return out


def nullable_compact(col, valid, out, stream=None) -> int:
    """NullableSegment<T>::CompactValid: the valid rows of ``col`` in order
    into ``out`` (a device tensor of the same dtype, len(col) entries);
    returns how many."""
    p, n, dt = _dev(col)
    po, no, dto = _dev(out)
    if no < n or dto != dt:
        raise ValueError("out must hold len(col) entries of the column's dtype")
    if valid.numel() < (n + 63) // 64:
        raise ValueError("validity words cover fewer rows than the column")
    cnt = C.c_uint64(0)
    st = None if stream is None else C.c_void_p(stream)
    if lib.voxel_gpu_nullable_compact(dt, C.c_void_p(p), _words(valid), n, C.c_void_p(po), C.byref(cnt), st) != 0:
        check("voxel_gpu_nullable_compact")
    return int(cnt.value)


def launch(desc: ScanDesc, stream=None) -> None:
    """Enqueues a scan on ``stream`` (a hipStream_t handle as int, or None)."""
    st = None if stream is None else C.c_void_p(stream)
    if lib.voxel_gpu_scan_launch(C.byref(desc), st) != 0:
        check("voxel_gpu_scan_launch")


def result(dtype: int, stream=None) -> dict:
    r = Agg()
    st = None if stream is None else C.c_void_p(stream)
    if lib.voxel_gpu_scan_result(st, C.byref(r)) != 0:
        check("voxel_gpu_scan_result")
    return _result(dtype, r)

"""Columnar encodings on the GPU (data/encoding.hpp:104-527): DecodeBatch and
filter + aggregate scans over encoded device columns (voxel_gpu_decode,
voxel_gpu_scan_encoded in include/voxel_c.h).

Arrays are device tensors in the reference's member layout: Dictionary
(Dictionary, Indices), RLE (Values, RunLengths), Delta (Deltas, BaseValue),
BitPacked (PackedData, BitWidth, ValueCount), FrameOfReference
(EncodedValues, BaseValue, Scale).
"""
from __future__ import annotations

import ctypes as C

from ._lib import Agg, EncodedCol, check, lib
from .scan import _result, threshold_bits

DICTIONARY, RLE, DELTA, BITPACKED, FOR = 1, 2, 3, 4, 6
_DT = {"torch.float64": 0, "torch.float32": 1, "torch.int64": 2, "torch.int32": 3}


class Encoded:
    """One encoded device column (keeps its tensors alive)."""

    def __init__(self, kind: int, values, index=None, rows: int = 0, dtype=None, bit_width: int = 0,
                 base: float = 0.0, scale: float = 1.0, delta_base: int = 0):
        self.kind, self._keep = kind, (values, index)
        c = EncodedCol()
        c.kind = kind
        c.dtype = _DT.get(str(dtype if dtype is not None else values.dtype), 0)
        c.values = values.data_ptr() if values is not None and values.numel() else None
        c.index = index.data_ptr() if index is not None and index.numel() else None
        c.nvalues = values.numel() if values is not None else 0
        c.rows, c.bit_width, c.base, c.scale, c.delta_base = rows, bit_width, base, scale, delta_base
        self.col = c

    @property
    def value_dtype(self) -> int:
        return {BITPACKED: 2, FOR: 0}.get(self.kind, self.col.dtype)

    def decode(self, offset: int, count: int, out) -> int:
        """IEncoding::DecodeBatch(out, offset, count) into device tensor ``out``."""
        n = int(lib.voxel_gpu_decode(C.byref(self.col), offset, count, C.c_void_p(out.data_ptr()), None))
        if n < 0:
            check("voxel_gpu_decode")
        return n

    def scan(self, op, t) -> dict:
        """Filter op(value, t) + COUNT/SUM/MIN/MAX over the decoded rows."""
        from .scan import OPS
        op = OPS.get(op, op) if isinstance(op, str) else op
        dt = self.value_dtype
        r = Agg()
        tb = threshold_bits(dt, t) if self.kind != BITPACKED else int(t) & 0xFFFFFFFF
        if lib.voxel_gpu_scan_encoded(C.byref(self.col), op, tb, None, C.byref(r)) != 0:
            check("voxel_gpu_scan_encoded")
        return _result(dt, r)


def dictionary(dictionary, indices) -> Encoded:
    return Encoded(DICTIONARY, dictionary, indices, rows=indices.numel())


def rle(values, run_lengths) -> Encoded:
    return Encoded(RLE, values, run_lengths, rows=int(run_lengths.sum().item()) if run_lengths.numel() else 0)


def delta(deltas, base_value: int = 0) -> Encoded:
    return Encoded(DELTA, deltas, None, rows=deltas.numel(), delta_base=int(base_value))


def bitpacked(packed_words, bit_width: int, value_count: int) -> Encoded:
    return Encoded(BITPACKED, packed_words, None, rows=value_count, bit_width=bit_width)


def frame_of_reference(encoded_values, base_value: float, scale: float) -> Encoded:
    return Encoded(FOR, encoded_values, None, rows=encoded_values.numel(), base=base_value, scale=scale)

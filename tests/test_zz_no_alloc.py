"""voxel_gpu_scan_noalloc (the allocation-free scan) against voxel_gpu_scan.

Both run the same scan kernels; the allocation-free one keeps its partials in
a static device buffer and writes its voxel_gpu_agg into caller device
memory.  Bar: every field bit-identical to voxel_gpu_scan's on the same
input (the merge order depends only on the grid size, which both pick the
same way), on two calls in a row, the second over fewer rows, so the second
call reads none of the first call's partials.  Calls go through the C ABI.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F64, F32, I64, I32 = 0, 1, 2, 3
AGG_BYTES = 80


def _column(torch, dtype, n, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    if dtype in (F64, F32):
        t = torch.rand(n, dtype=torch.float64, device="cuda", generator=g) * 2000.0 - 500.0
        return t if dtype == F64 else t.to(torch.float32)
    t = torch.randint(-2**31, 2**31 - 1, (n,), dtype=torch.int64, device="cuda", generator=g)
    return t if dtype == I64 else t.to(torch.int32)


def _agg_bytes(a) -> bytes:
    return bytes(C.string_at(C.addressof(a), AGG_BYTES))


def _noalloc(lib, desc, torch, stream=None):
    from voxel_amd._lib import Agg, last_error
    out = torch.zeros(AGG_BYTES, dtype=torch.uint8, device="cuda")
    s = stream.cuda_stream if stream is not None else None
    rc = lib.voxel_gpu_scan_noalloc(C.byref(desc), C.c_void_p(out.data_ptr()), C.c_void_p(s) if s else None)
    assert rc == 0, last_error()
    (stream or torch.cuda.current_stream()).synchronize()
    raw = out.cpu().numpy().tobytes()
    return Agg.from_buffer_copy(raw), raw


def _reference(lib, desc):
    from voxel_amd._lib import Agg, last_error
    a = Agg()
    assert lib.voxel_gpu_scan(C.byref(desc), None, C.byref(a)) == 0, last_error()
    return a


CASES = [  # dtype, op_a, t_a, op_b, t_b  (a > t_a AND b < t_b, and single-column forms)
    (F64, "gt", 292.8932188, "lt", 707.1067812),
    (F32, "ge", 5.0125628, "le", 994.9874371),
    (I64, "gt", 1 << 30, "lt", 1 << 29),
    (I32, "ne", 7, "true", 0),
]


@pytest.mark.parametrize("case", CASES, ids=["f64", "f32", "i64", "i32"])
def test_noalloc_matches_scan_two_calls(gpu, case):
    import torch
    from voxel_amd import scan
    dtype, op_a, t_a, op_b, t_b = case
    for n, seed in ((1 << 22) + 13, 11), (1000, 12):       # second call: fewer rows
        a = _column(torch, dtype, n, seed)
        b = _column(torch, dtype, n, seed + 100) if op_b != "true" else None
        desc = scan.make_desc(a, op_a, t_a, b, op_b, t_b)
        got, raw = _noalloc(gpu, desc, torch)
        ref = _reference(gpu, desc)
        assert raw == _agg_bytes(ref), (n, got.count, ref.count, got.sum, ref.sum)
        # the count against a host recount of the same rows
        ah = a.cpu().numpy()
        if dtype in (F64, F32):
            m = ah > np.array(t_a, dtype=ah.dtype) if op_a == "gt" else ah >= np.array(t_a, dtype=ah.dtype)
        else:
            m = (ah > t_a) if op_a == "gt" else (ah != t_a)
        if b is not None:
            bh = b.cpu().numpy()
            m &= (bh < np.array(t_b, dtype=bh.dtype)) if op_b == "lt" else (bh <= np.array(t_b, dtype=bh.dtype))
        assert got.count == int(m.sum())


def test_noalloc_empty_stream_and_max_blocks(gpu):
    import torch
    from voxel_amd import scan
    from voxel_amd._lib import last_error
    a = _column(torch, F64, (1 << 21) + 5, 21)
    # explicit grid at the static buffer's limit, on a non-default stream
    desc = scan.make_desc(a, "lt", 100.0, blocks=4096)
    st = torch.cuda.Stream()
    got, raw = _noalloc(gpu, desc, torch, stream=st)
    assert raw == _agg_bytes(_reference(gpu, desc))
    # no rows: every field zero, as voxel_gpu_scan returns
    e = scan.make_desc(a[:0], "gt", 0.0)
    got, raw = _noalloc(gpu, e, torch)
    assert raw == _agg_bytes(_reference(gpu, e)) == bytes(AGG_BYTES)
    # repeated calls take no device memory
    free0 = torch.cuda.mem_get_info()[0]
    out = torch.zeros(AGG_BYTES, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(5):
        assert gpu.voxel_gpu_scan_noalloc(C.byref(desc), C.c_void_p(out.data_ptr()), None) == 0, last_error()
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0
    # limits: more workgroups than the buffer holds, ticket layouts
    big = scan.make_desc(a, "lt", 100.0, blocks=4097)
    assert gpu.voxel_gpu_scan_noalloc(C.byref(big), C.c_void_p(out.data_ptr()), None) == -1
    assert "at most 4096" in last_error()
    lay = scan.make_desc(a, "lt", 100.0, layout=1)
    assert gpu.voxel_gpu_scan_noalloc(C.byref(lay), C.c_void_p(out.data_ptr()), None) == -1
    assert "layout 1" in last_error()

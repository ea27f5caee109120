"""GPU parity of the radix sort (vx_radix.hip) through the C ABI.

Bar: bit-exact.  A stable argsort is unique, so the GPU permutation must equal
the reference's (golden vectors from oracle/_ref) and the oracle's
(vo_radix_sort_u64 / vo_mergesort_f64, pinned against numpy's stable sort).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import oracle as o

pytestmark = pytest.mark.gpu


def gpu_sort_pairs(keys: np.ndarray, vals: np.ndarray | None = None):
    import torch
    from voxel_amd._lib import lib
    n = keys.size
    kin = torch.from_numpy(keys.view(np.int64)).cuda()
    vin = torch.from_numpy(vals.view(np.int32)).cuda() if vals is not None else None
    kout = torch.empty(max(n, 1), dtype=torch.int64, device="cuda")
    vout = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
    rc = lib.voxel_gpu_sort_pairs_u64(kin.data_ptr() if n else None, vin.data_ptr() if vin is not None else None,
                                      kout.data_ptr(), vout.data_ptr(), n, None)
    assert rc == 0, lib.voxel_last_error()
    torch.cuda.synchronize()
    return kout[:n].cpu().numpy().view(np.uint64), vout[:n].cpu().numpy().view(np.uint32)


def test_radix_golden(gpu, golden):
    g = golden["radix_u64"]
    k, i = gpu_sort_pairs(np.array(g["in"], np.uint64))
    assert k.tolist() == g["keys"] and i.tolist() == g["idx"]


@pytest.mark.parametrize("n", [0, 1, 2, 63, 4095, 4096, 4097, 8191, 8192, 8193, 65537, 100_003, 1_000_000])
def test_radix_random_vs_oracle(gpu, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 2**64, n, dtype=np.uint64)
    vals = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    k, v = gpu_sort_pairs(keys, vals)
    kk, vv = keys.copy(), vals.copy()
    o.O.vo_radix_sort_u64(o.ptr(kk), o.ptr(vv), n)
    assert np.array_equal(k, kk) and np.array_equal(v, vv)


@pytest.mark.parametrize("kind", ["dups", "equal", "low16", "high_only", "zipfish"])
def test_radix_skewed_keys(gpu, kind):
    rng = np.random.default_rng(5)
    n = 300_001
    if kind == "dups":
        keys = rng.integers(0, 40, n).astype(np.uint64)
    elif kind == "equal":
        keys = np.full(n, 0xDEADBEEF, np.uint64)            # every pass skipped: copy
    elif kind == "low16":
        keys = rng.integers(0, 2**16, n).astype(np.uint64)  # six passes skipped
    elif kind == "high_only":
        keys = (rng.integers(0, 256, n).astype(np.uint64) << np.uint64(56))
    else:
        keys = (np.floor(rng.pareto(0.1, n)).clip(0, 2**32 - 1)).astype(np.uint64)
    k, v = gpu_sort_pairs(keys)
    ref = np.argsort(keys, kind="stable")
    assert np.array_equal(v, ref.astype(np.uint32)) and np.array_equal(k, keys[ref])


def test_argsort_f64_golden(gpu, golden):
    from voxel_amd._lib import lib
    for name in ("sort_c_test", "sort_ties"):
        g = golden[name]
        x = np.array(g["data"], np.float64)
        out = np.zeros(x.size, np.uint32)
        lib.voxel_sort_ascending(x.ctypes.data, x.size, out.ctypes.data)
        assert out.tolist() == g["asc"], name
        lib.voxel_sort_descending(x.ctypes.data, x.size, out.ctypes.data)
        assert out.tolist() == g["desc"], name


@pytest.mark.parametrize("n", [1, 7, 5000, 262_147])
def test_argsort_f64_vs_oracle(gpu, n):
    from voxel_amd._lib import lib
    rng = np.random.default_rng(n + 1)
    x = np.round(rng.standard_normal(n) * 100, 1)
    x[::17] = -0.0
    x[::29] = 0.0
    x[::101] = np.inf
    x[::103] = -np.inf
    for desc in (0, 1):
        out = np.zeros(n, np.uint32)
        (lib.voxel_sort_descending if desc else lib.voxel_sort_ascending)(x.ctypes.data, n, out.ctypes.data)
        ref = np.zeros(n, np.uint32)
        o.O.vo_mergesort_f64(o.ptr(x), n, desc, o.ptr(ref))
        assert np.array_equal(out, ref), desc


def test_topk_golden_and_random(gpu, golden):
    from voxel_amd._lib import lib
    g = golden["topk3"]
    x = np.array(g["data"], np.float64)
    out = np.zeros(3, np.float64)
    lib.voxel_topk(x.ctypes.data, x.size, 3, out.ctypes.data, 1)
    assert out.tolist() == g["largest"] == [9.0, 6.0, 5.0]
    rng = np.random.default_rng(3)
    y = rng.standard_normal(100_000)
    out = np.zeros(50, np.float64)
    lib.voxel_topk(y.ctypes.data, y.size, 50, out.ctypes.data, 0)
    assert np.array_equal(out, np.sort(y)[:50])
    lib.voxel_topk(y.ctypes.data, y.size, 50, out.ctypes.data, 1)
    assert np.array_equal(out, np.sort(y)[::-1][:50])


def test_full_size_sort_properties(gpu):
    """2^29 uniform u64 keys with row-index payload: output keys are
    non-decreasing, the payload is a permutation whose gather reproduces the
    keys, and equal neighbours keep increasing payloads (stability)."""
    import torch
    from voxel_amd._lib import lib
    n = 1 << 29
    hi = torch.empty(n, dtype=torch.int64, device="cuda")
    lo = torch.empty(n, dtype=torch.int64, device="cuda")
    assert lib.voxel_gpu_fill_uniform(2, hi.data_ptr(), n, 7, 0.0, 2.0**32, None) == 0
    assert lib.voxel_gpu_fill_uniform(2, lo.data_ptr(), n, 8, 0.0, 2.0**32, None) == 0
    torch.cuda.synchronize()
    kin = (hi << 32) | lo            # full 64-bit keys (top bit set on half of them)
    del hi, lo
    kout = torch.empty_like(kin)
    vout = torch.empty(n, dtype=torch.int32, device="cuda")
    assert lib.voxel_gpu_sort_pairs_u64(kin.data_ptr(), None, kout.data_ptr(), vout.data_ptr(), n, None) == 0
    torch.cuda.synchronize()
    ku = kout ^ (-(2**63))            # unsigned order as signed order
    assert bool((ku[1:] >= ku[:-1]).all())
    idx = vout.to(torch.int64) & 0xFFFFFFFF
    assert bool(torch.equal(kin[idx], kout))
    seen = torch.zeros(n, dtype=torch.int32, device="cuda")
    seen.index_add_(0, idx, torch.ones(n, dtype=torch.int32, device="cuda"))
    assert bool((seen == 1).all())
    eq = kout[1:] == kout[:-1]
    assert bool((idx[1:][eq] > idx[:-1][eq]).all())

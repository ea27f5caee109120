"""GPU tests of the learned-CDF sort (vx_learned.hip) through the C ABI.

Bars: the sort is bit-exact with a stable argsort (numpy kind='stable', the
order RadixSort<u64>::Sort with indices gives, ops/sort.hpp:83-160) and with
voxel_gpu_sort_pairs_u64 on every distribution; the network's forward pass,
gradient and Adam step match the float64 CPU model (oracle/mlp.py) within
FP32 tolerances stated per test.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu


def _lib():
    from voxel_amd._lib import lib
    return lib


def gpu_sort(keys, vals=None, model=0, learned=True):
    import torch
    from voxel_amd._lib import LearnedStats
    lib = _lib()
    n = keys.size
    kd = torch.from_numpy(keys.view(np.int64)).cuda()
    vd = torch.from_numpy(vals.view(np.int32)).cuda() if vals is not None else None
    ko = torch.empty(max(n, 1), dtype=torch.int64, device="cuda")
    vo = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
    st = LearnedStats()
    if learned:
        rc = lib.voxel_gpu_sort_pairs_u64_learned(kd.data_ptr(), vd.data_ptr() if vd is not None else None,
                                                  ko.data_ptr(), vo.data_ptr(), n, model, C.byref(st), None)
    else:
        rc = lib.voxel_gpu_sort_pairs_u64(kd.data_ptr(), vd.data_ptr() if vd is not None else None,
                                          ko.data_ptr(), vo.data_ptr(), n, None)
    assert rc == 0, lib.voxel_last_error()
    torch.cuda.synchronize()
    return ko[:n].cpu().numpy().view(np.uint64), vo[:n].cpu().numpy().view(np.uint32), st.as_dict()


def distribution(name, n, seed):
    rng = np.random.default_rng(seed)
    if name == "uniform":
        return rng.integers(0, 2**64, n, dtype=np.uint64)
    if name == "lognormal":
        return np.minimum(np.exp(rng.normal(0.0, 2.5, n)) * 1e9, 2.0**63).astype(np.uint64)
    if name == "zipf":
        return rng.zipf(1.3, n).astype(np.uint64)
    if name == "sorted95":
        k = np.sort(rng.integers(0, 2**40, n, dtype=np.uint64))
        i = rng.integers(0, n, n // 20)
        j = rng.integers(0, n, n // 20)
        k[i], k[j] = k[j].copy(), k[i].copy()
        return k
    if name == "narrow":                                  # few distinct values, heavy duplicates
        return rng.integers(1000, 1100, n).astype(np.uint64)
    raise ValueError(name)


@pytest.mark.parametrize("name", ["uniform", "lognormal", "zipf", "sorted95", "narrow"])
def test_learned_sort_bit_exact(gpu, name):
    n = 1 << 22
    keys = distribution(name, n, 11)
    order = np.argsort(keys, kind="stable")
    k, v, st = gpu_sort(keys)
    assert np.array_equal(k, keys[order]), st
    assert np.array_equal(v, order.astype(np.uint32)), st
    k2, v2, _ = gpu_sort(keys, learned=False)
    assert np.array_equal(k, k2) and np.array_equal(v, v2)


@pytest.mark.parametrize("n", [0, 1, 1000, 65_535, 65_536, 65_537, 3_000_001])
def test_learned_sort_sizes_and_payload(gpu, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 2**64, n, dtype=np.uint64)
    keys[: n // 3] = keys[n // 2] if n else 0             # a duplicated key
    vals = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    order = np.argsort(keys, kind="stable")
    for model in (0, 1):
        k, v, st = gpu_sort(keys, vals, model=model)
        assert np.array_equal(k, keys[order]) and np.array_equal(v, vals[order]), (model, st)


def test_learned_sort_all_equal_and_two_values(gpu):
    n = 1 << 20
    keys = np.full(n, 7, np.uint64)
    k, v, st = gpu_sort(keys)
    assert np.array_equal(k, keys) and np.array_equal(v, np.arange(n, dtype=np.uint32))
    keys = np.where(np.arange(n) % 3 == 0, 2**63 + 5, 3).astype(np.uint64)
    order = np.argsort(keys, kind="stable")
    k, v, st = gpu_sort(keys)
    assert np.array_equal(k, keys[order]) and np.array_equal(v, order.astype(np.uint32))


def test_learned_sort_large_matches_lsd(gpu):
    """64M lognormal keys: identical to the plain LSD sort (keys and payload),
    keys non-decreasing, payload a permutation."""
    import torch
    lib = _lib()
    n = 1 << 26
    keys = distribution("lognormal", n, 5)
    kd = torch.from_numpy(keys.view(np.int64)).cuda()
    outs = []
    for learned in (True, False):
        ko = torch.empty(n, dtype=torch.int64, device="cuda")
        vo = torch.empty(n, dtype=torch.int32, device="cuda")
        if learned:
            rc = lib.voxel_gpu_sort_pairs_u64_learned(kd.data_ptr(), None, ko.data_ptr(), vo.data_ptr(), n, 0, None, None)
        else:
            rc = lib.voxel_gpu_sort_pairs_u64(kd.data_ptr(), None, ko.data_ptr(), vo.data_ptr(), n, None)
        assert rc == 0, lib.voxel_last_error()
        outs.append((ko, vo))
    torch.cuda.synchronize()
    assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])
    k = outs[0][0].cpu().numpy().view(np.uint64)
    assert np.all(k[1:] >= k[:-1])
    assert np.array_equal(np.sort(outs[0][1].cpu().numpy()), np.arange(n, dtype=np.int32))


def test_learned_model_beats_equal_width_buckets(gpu):
    """Lognormal keys: the learned CDF keeps buckets within the LDS capacity,
    the equal-width partition (no model) does not; both sort exactly."""
    n = 1 << 24
    keys = distribution("lognormal", n, 9)
    k0, v0, s0 = gpu_sort(keys, model=0)
    k1, v1, s1 = gpu_sort(keys, model=1)
    assert np.array_equal(k0, k1) and np.array_equal(v0, v1)
    assert s0["fell_back"] == 0 and s0["big_rows"] < n // 100, str(s0)
    assert s1["big_rows"] > n // 4 or s1["fell_back"] == 1, str(s1)
    assert s0["widest_group"] * 10 < s1["widest_group"], str((s0, s1))
    assert s0["final_loss"] < 0.02, str(s0)            # mean squared error of the logit CDF


def _params(seed):
    import torch
    lib = _lib()
    p = torch.empty(mlp.N_PARAMS, dtype=torch.float32, device="cuda")
    assert lib.voxel_gpu_mlp_param_count() == mlp.N_PARAMS
    assert lib.voxel_gpu_mlp_init(p.data_ptr(), seed) == 0
    return p


def test_mlp_forward_and_gradient_match_cpu(gpu):
    import torch
    lib = _lib()
    rng = np.random.default_rng(2)
    p = _params(7)
    n = 1000                                                # not a multiple of the 64-point slices
    x = np.sort(rng.random(n)).astype(np.float32)
    y = (x ** 2).astype(np.float32)
    xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    out = torch.empty(n, dtype=torch.float32, device="cuda")
    assert lib.voxel_gpu_mlp_forward(p.data_ptr(), xd.data_ptr(), n, out.data_ptr(), None) == 0
    g = torch.empty(mlp.N_PARAMS, dtype=torch.float32, device="cuda")
    l = torch.empty(1, dtype=torch.float32, device="cuda")
    assert lib.voxel_gpu_mlp_grad(p.data_ptr(), xd.data_ptr(), yd.data_ptr(), n, g.data_ptr(), l.data_ptr(), None) == 0
    torch.cuda.synchronize()
    pc = p.cpu().numpy().astype(np.float64)
    ref = mlp.forward(pc, x)
    assert np.max(np.abs(out.cpu().numpy() - ref)) <= 1e-5 * max(1.0, np.max(np.abs(ref)))
    gref = mlp.grad(pc, x, y)
    gg = g.cpu().numpy().astype(np.float64)
    # FP32 sums of 1000 terms: 1e-4 of the largest component, per component
    assert np.max(np.abs(gg - gref)) <= 1e-4 * np.max(np.abs(gref)), np.argmax(np.abs(gg - gref))
    assert abs(float(l.item()) - mlp.loss(pc, x, y)) <= 1e-5 * mlp.loss(pc, x, y)


def test_mlp_adam_step_and_training(gpu):
    import torch
    lib = _lib()
    rng = np.random.default_rng(3)
    n = 4096
    x = np.sort(rng.random(n)).astype(np.float32)
    y = np.sqrt(x).astype(np.float32)                        # a concave CDF
    xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    p = _params(1)
    p0 = p.cpu().numpy().astype(np.float64)
    lo = C.c_float()
    assert lib.voxel_gpu_mlp_train(p.data_ptr(), xd.data_ptr(), yd.data_ptr(), n, 1, C.c_float(0.01),
                                   C.byref(lo), None) == 0
    p1 = p.cpu().numpy().astype(np.float64)
    g = mlp.grad(p0, x, y)
    pref, _, _ = mlp.adam(p0, g, np.zeros_like(g), np.zeros_like(g), 1, 0.01)
    sel = np.abs(g) > 1e-5                                  # first Adam step = lr * sign(g) there
    assert np.max(np.abs(p1[sel] - pref[sel])) <= 1e-6
    assert np.all(np.abs(p1[~sel] - p0[~sel]) <= 0.01 + 1e-6)
    assert abs(lo.value - mlp.loss(p0, x, y)) <= 1e-5 * mlp.loss(p0, x, y)
    start = mlp.loss(p1, x, y)
    assert lib.voxel_gpu_mlp_train(p.data_ptr(), xd.data_ptr(), yd.data_ptr(), n, 300, C.c_float(0.01),
                                   C.byref(lo), None) == 0
    end = mlp.loss(p.cpu().numpy().astype(np.float64), x, y)
    assert end < start / 20 and end < 1e-3, (start, end)

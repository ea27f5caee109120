import os
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs through libvoxel_c")


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "reference_vectors.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def x42():
    """validate.cpp / jit_test.cpp input: mt19937_64(42) uniform [0,1000), 1e6 rows."""
    from oracle import oracle as o
    return o.mt_uniform_f64(42, 1_000_000)


@pytest.fixture(scope="session")
def gpu():
    """The C ABI library, on a machine with a usable GPU."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.init()
    torch.empty(1, device="cuda")
    torch.cuda.synchronize()
    from voxel_amd._lib import lib, last_error
    assert lib.voxel_gpu_available() == 1, "libvoxel_c found no usable GPU: " + last_error()
    return lib


@pytest.fixture(params=["interpreted", "compiled"])
def expr_mode(request, gpu):
    """Tier-0 loops with vector words on the library's interpreter kernels, or
    compiled whole at run time for every loop size (voxel_gpu_set_expr_jit)."""
    prev = gpu.voxel_gpu_set_expr_jit(0 if request.param == "interpreted" else 2)
    yield request.param
    gpu.voxel_gpu_set_expr_jit(prev)

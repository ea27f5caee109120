"""voxel_amd -- MI355X (gfx950) execution backend for VoxelVM.

The product is the C ABI library ``voxel_amd/libvoxel_c.so`` (include/voxel_c.h):
the reference's 64-symbol ABI plus ``voxel_gpu_*`` extensions.  This package
holds its sources (``csrc/``), its build (``build.py``), and the Python
surfaces that sit on top of the ABI:

* ``voxel_amd.voxel_py``    -- the ``voxel_py`` module surface (EngineF64, jit_run, ...)
* ``voxel_amd.voxel_query`` -- the query builder (``from_numpy(x).filter_gt(t).sum().run()``)
* ``voxel_amd.scan``        -- relational filter+aggregate over device columns
"""
from __future__ import annotations

__all__ = ["build", "lib", "scan"]


def build() -> str:
    from .build import build as _b
    return _b()


def __getattr__(name):
    if name == "lib":
        from ._lib import lib
        return lib
    raise AttributeError(name)

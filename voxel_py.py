"""``import voxel_py`` -- the module name the reference's Python callers
import (defined by bindings/voxel_bindings.cpp:25; imported by
bindings/voxel_query.py:8, voxel_market.py:13, voxel_sql.py:18,
voxel_data.py:3, voxel_stream.py:3).  Importing it yields the
``voxel_amd.voxel_py`` module itself, backed by libvoxel_c on the GPU.
"""
import sys

from voxel_amd import voxel_py as _impl

sys.modules[__name__] = _impl

"""Batched queries and the continuous-batching server (voxel_gpu_batch_* and
voxel_gpu_server_* in include/voxel_c.h).

Queries are scan descriptors (``voxel_amd.scan.make_desc``) over resident
device columns.  ``Batch.run`` executes a list of them in one persistent
kernel launch; ``Server.replay`` feeds an arrival schedule through the
continuous-batching loop (mode 0), one synchronous launch per query (mode 1)
or one launch per query queued without waiting (mode 2); ``Server.start`` /
``submit`` / ``poll`` / ``wait`` / ``stop`` serve queries as they arrive.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import Agg, ScanDesc, ServeStats, check, lib
from .scan import _result


def _descs(queries) -> C.Array:
    arr = (ScanDesc * len(queries))()
    for i, q in enumerate(queries):
        arr[i] = q
    return arr


class Batch:
    """One batch engine (own stream, graph, persistent kernel)."""

    def __init__(self, max_queries: int = 4096):
        self.h = lib.voxel_gpu_batch_create(max_queries)
        if not self.h:
            check("voxel_gpu_batch_create")
        self.max_queries = max_queries

    def close(self) -> None:
        if self.h:
            lib.voxel_gpu_batch_destroy(self.h)
            self.h = None

    def __del__(self, _lib=lib):          # the library outlives module teardown
        if getattr(self, "h", None):
            _lib.voxel_gpu_batch_destroy(self.h)
            self.h = None

    def config(self, unroll: int = 0, tile_rows: int = 0, blocks_per_cu: int = 0) -> None:
        if lib.voxel_gpu_batch_config(self.h, unroll, tile_rows, blocks_per_cu) != 0:
            check("voxel_gpu_batch_config")

    def get_config(self) -> dict:
        u, t, b = C.c_int(), C.c_uint32(), C.c_int()
        lib.voxel_gpu_batch_get_config(self.h, C.byref(u), C.byref(t), C.byref(b))
        return {"unroll": u.value, "tile_rows": t.value, "blocks_per_cu": b.value}

    def run_raw(self, queries):
        """(list of Agg structs, device ms)."""
        n = len(queries)
        out = (Agg * n)()
        ms = C.c_float()
        d = _descs(queries)
        if lib.voxel_gpu_batch_run(self.h, C.addressof(d), n, C.addressof(out), C.byref(ms)) != 0:
            check("voxel_gpu_batch_run")
        return list(out), ms.value

    def run(self, queries) -> list:
        raw, _ = self.run_raw(queries)
        return [_result(q.dtype, r) for q, r in zip(queries, raw)]

    def autotune(self, queries, reps: int = 3) -> dict:
        best, tried = C.c_double(), C.c_int()
        d = _descs(queries)
        if lib.voxel_gpu_batch_autotune(self.h, C.addressof(d), len(queries), reps, C.byref(best),
                                        C.byref(tried)) != 0:
            check("voxel_gpu_batch_autotune")
        return dict(self.get_config(), best_ms=best.value, configs_tried=tried.value)


class Server:
    """Continuous-batching request loop over resident columns.

    Live use: ``start()``, then ``submit(desc)`` from any thread (a ticket),
    ``poll(ticket)`` / ``wait(ticket)`` for its result, ``stop()`` for the
    statistics.  ``replay`` feeds a recorded arrival schedule through the
    same loop (mode 0), one synchronous launch per query (mode 1) or one
    launch per query queued without waiting (mode 2)."""

    def __init__(self, max_batch: int = 1024, memory_budget: int = 64 << 20, inflight: int = 2):
        self.h = lib.voxel_gpu_server_create(max_batch, memory_budget, inflight)
        if not self.h:
            check("voxel_gpu_server_create")

    def close(self) -> None:
        if self.h:
            lib.voxel_gpu_server_destroy(self.h)
            self.h = None

    def __del__(self, _lib=lib):          # the library outlives module teardown
        if getattr(self, "h", None):
            _lib.voxel_gpu_server_destroy(self.h)
            self.h = None

    def config(self, unroll: int = 0, tile_rows: int = 0, blocks_per_cu: int = 0) -> None:
        if lib.voxel_gpu_server_config(self.h, unroll, tile_rows, blocks_per_cu) != 0:
            check("voxel_gpu_server_config")

    def autotune(self, queries, reps: int = 3) -> dict:
        best, tried = C.c_double(), C.c_int()
        d = _descs(queries)
        if lib.voxel_gpu_server_autotune(self.h, C.addressof(d), len(queries), reps, C.byref(best),
                                         C.byref(tried)) != 0:
            check("voxel_gpu_server_autotune")
        return {"best_ms": best.value, "configs_tried": tried.value}

    def start(self) -> None:
        if lib.voxel_gpu_server_start(self.h) != 0:
            check("voxel_gpu_server_start")

    def submit(self, desc) -> int:
        t = int(lib.voxel_gpu_server_submit(self.h, C.byref(desc)))
        if t < 0:
            check("voxel_gpu_server_submit")
        return t

    def poll_raw(self, ticket: int):
        """The result (Agg) once the ticket has completed, else None."""
        r = Agg()
        rc = lib.voxel_gpu_server_poll(self.h, ticket, C.byref(r))
        if rc < 0:
            check("voxel_gpu_server_poll")
        return r if rc == 1 else None

    def wait_raw(self, ticket: int):
        """(Agg, seconds from submit to completion)."""
        r, lat = Agg(), C.c_double()
        if lib.voxel_gpu_server_wait(self.h, ticket, C.byref(r), C.byref(lat)) != 0:
            check("voxel_gpu_server_wait")
        return r, lat.value

    def stop(self) -> dict:
        st = ServeStats()
        if lib.voxel_gpu_server_stop(self.h, C.byref(st)) != 0:
            check("voxel_gpu_server_stop")
        return st.as_dict()

    def replay_raw(self, queries, arrivals, mode: int = 0):
        n = len(queries)
        arr = np.ascontiguousarray(arrivals, dtype=np.float64)
        assert arr.size == n
        out = (Agg * n)()
        lat = np.zeros(n, np.float64)
        st = ServeStats()
        d = _descs(queries)
        if lib.voxel_gpu_server_replay(self.h, C.addressof(d), arr.ctypes.data, n, mode, C.addressof(out),
                                       lat.ctypes.data, C.byref(st)) != 0:
            check("voxel_gpu_server_replay")
        return list(out), lat, st.as_dict()

    def replay(self, queries, arrivals, mode: int = 0):
        raw, lat, st = self.replay_raw(queries, arrivals, mode)
        return [_result(q.dtype, r) for q, r in zip(queries, raw)], lat, st

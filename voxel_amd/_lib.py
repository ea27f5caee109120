"""ctypes binding of libvoxel_c.so (the C ABI in include/voxel_c.h).

Loading fails loudly when the library has not been built: there is no
host-side fallback for any GPU entry point.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("VOXEL_LIB") or os.path.join(HERE, "libvoxel_c.so")
HEADER = os.path.join(HERE, "..", "include", "voxel_c.h")

# torch ships its own HIP runtime (same soname as /opt/rocm's).  Loading torch
# first makes this library bind to that already-loaded runtime, so a process
# that uses both holds one HIP runtime instead of two that contend for the GPU.
try:
    import torch  # noqa: F401
except ImportError:
    pass

if not os.path.exists(LIB_PATH):
    raise ImportError(
        "voxel_amd: %s is missing; build it with `python -m voxel_amd.build` "
        "(hipcc, gfx950). There is no CPU fallback." % LIB_PATH)

lib = C.CDLL(LIB_PATH)

P, U8, U16, U32, U64, I16, I32, I64, SZ, F64, INT = (
    C.c_void_p, C.c_uint8, C.c_uint16, C.c_uint32, C.c_uint64, C.c_int16, C.c_int32,
    C.c_int64, C.c_size_t, C.c_double, C.c_int)


class ScanDesc(C.Structure):
    _fields_ = [("dtype", C.c_int32), ("op_a", C.c_int32), ("op_b", C.c_int32), ("blocks", C.c_int32),
                ("unroll", C.c_int32), ("nontemporal", C.c_int32), ("layout", C.c_int32), ("pad_", C.c_int32),
                ("a", P), ("b", P), ("n", U64), ("t_a", U64), ("t_b", U64)]


class Agg(C.Structure):
    _fields_ = [("count", U64), ("sum", F64), ("sum_hi", F64), ("sum_lo", F64), ("isum", I64),
                ("fmin", F64), ("fmax", F64), ("imin", I64), ("imax", I64),
                ("has_minmax", C.c_int32), ("pad", C.c_int32)]


class LearnedStats(C.Structure):
    _fields_ = [("model", C.c_int32), ("log_grid", C.c_int32), ("fell_back", C.c_int32),
                ("partition_passes", C.c_int32), ("train_steps", C.c_int32), ("final_loss", C.c_float),
                ("buckets", U64), ("cells", U64), ("groups", U64), ("longest_bucket", U64), ("widest_group", U64),
                ("big_groups", U64), ("big_rows", U64), ("big_unsorted", U64),
                ("sample_ms", C.c_float), ("train_ms", C.c_float), ("table_ms", C.c_float),
                ("partition_ms", C.c_float), ("bucket_ms", C.c_float), ("fixup_ms", C.c_float)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class EncodedCol(C.Structure):
    """voxel_gpu_encoded_col."""
    _fields_ = [("kind", C.c_int32), ("dtype", C.c_int32), ("rows", U64), ("values", P), ("index", P),
                ("nvalues", U64), ("bit_width", U32), ("pad", U32), ("base", F64), ("scale", F64),
                ("delta_base", I64)]


class NullableStats(C.Structure):
    """voxel_gpu_nullable_stats."""
    _fields_ = [("valid_count", U64), ("null_count", U64), ("sum", F64), ("min", F64), ("max", F64),
                ("isum", I64), ("imin", I64), ("imax", I64)]


class JoinStats(C.Structure):
    _fields_ = [("mode", C.c_int32), ("fell_back", C.c_int32), ("partitions", C.c_uint32), ("pad", C.c_uint32),
                ("slots_per_partition", U64), ("max_partition_keys", U64), ("min_partition_keys", U64),
                ("build_ms", C.c_float), ("pad2", C.c_float), ("model", LearnedStats)]

    def as_dict(self) -> dict:
        d = {f: getattr(self, f) for f, _ in self._fields_ if not f.startswith("pad") and f != "model"}
        d["model"] = self.model.as_dict()
        return d


class ServeStats(C.Structure):
    _fields_ = [("queries", U64), ("batches", U64), ("bytes", U64), ("max_batch_seen", U64),
                ("admission_waits", U64), ("makespan_s", F64), ("qps", F64), ("gbps", F64), ("mean_batch", F64),
                ("p50_s", F64), ("p99_s", F64), ("mean_s", F64), ("max_s", F64)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


SIGS = {
    # engine (f64)
    "voxel_engine_create_f64": (P, []),
    "voxel_engine_destroy": (None, [P]),
    "voxel_engine_add_segment": (SZ, [P, P, SZ]),
    "voxel_engine_load_program": (None, [P, P, SZ]),
    "voxel_engine_run": (None, [P]),
    "voxel_engine_run_parallel": (None, [P, SZ, U32, U8]),
    "voxel_engine_set_scalar": (None, [P, SZ, U64]),
    "voxel_engine_get_scalar": (U64, [P, SZ]),
    "voxel_engine_set_scalar_f64": (None, [P, SZ, F64]),
    "voxel_engine_get_scalar_f64": (F64, [P, SZ]),
    "voxel_engine_k_lanes": (SZ, []),
    # engine (i64)
    "voxel_engine_create_i64": (P, []),
    "voxel_engine_destroy_i64": (None, [P]),
    "voxel_engine_add_segment_i64": (SZ, [P, P, SZ]),
    "voxel_engine_load_program_i64": (None, [P, P, SZ]),
    "voxel_engine_run_i64": (None, [P]),
    "voxel_engine_set_scalar_i64": (None, [P, SZ, U64]),
    "voxel_engine_get_scalar_i64": (U64, [P, SZ]),
    # JIT
    "voxel_jit_run": (F64, [P, SZ, P, SZ, F64]),
    # encoders
    "voxel_instr_vload": (U32, [U8, U8, U8, U8]),
    "voxel_instr_vfilter_gt": (U32, [U8, U8, U8]),
    "voxel_instr_vfilter_ge": (U32, [U8, U8, U8]),
    "voxel_instr_vfilter_lt": (U32, [U8, U8, U8]),
    "voxel_instr_vfilter_le": (U32, [U8, U8, U8]),
    "voxel_instr_vsum": (U32, [U8, U8]),
    "voxel_instr_vred_min": (U32, [U8, U8]),
    "voxel_instr_vred_max": (U32, [U8, U8]),
    "voxel_instr_vcount": (U32, [U8, U8]),
    "voxel_instr_addf": (U32, [U8, U8, U8]),
    "voxel_instr_add": (U32, [U8, U8, U8]),
    "voxel_instr_sub": (U32, [U8, U8, U8]),
    "voxel_instr_mul": (U32, [U8, U8, U8]),
    "voxel_instr_mov": (U32, [U8, I16]),
    "voxel_instr_movr": (U32, [U8, U8]),
    "voxel_instr_cmp": (U32, [U8, U8]),
    "voxel_instr_jnz": (U32, [I16]),
    "voxel_instr_jz": (U32, [I16]),
    "voxel_instr_jmp": (U32, [I16]),
    "voxel_instr_halt": (U32, []),
    # null bitmap
    "voxel_null_create": (P, [SZ]),
    "voxel_null_destroy": (None, [P]),
    "voxel_null_is_null": (INT, [P, SZ]),
    "voxel_null_set_null": (None, [P, SZ]),
    "voxel_null_set_valid": (None, [P, SZ]),
    "voxel_null_null_count": (SZ, [P]),
    "voxel_null_valid_count": (SZ, [P]),
    # aggregator
    "voxel_agg_create": (P, []),
    "voxel_agg_destroy": (None, [P]),
    "voxel_agg_init": (None, [P, SZ]),
    "voxel_agg_accumulate": (None, [P, P, P, SZ]),
    "voxel_agg_group_count": (SZ, [P]),
    # sort / topk
    "voxel_sort_ascending": (None, [P, SZ, P]),
    "voxel_sort_descending": (None, [P, SZ, P]),
    "voxel_topk": (None, [P, SZ, SZ, P, INT]),
    # dictionary
    "voxel_dict_create": (P, []),
    "voxel_dict_destroy": (None, [P]),
    "voxel_dict_build": (None, [P, P, SZ]),
    "voxel_dict_cardinality": (SZ, [P]),
    "voxel_dict_decode_batch": (None, [P, P, SZ, SZ]),
    # pool
    "voxel_pool_create": (P, [U32]),
    "voxel_pool_destroy": (None, [P]),
    "voxel_pool_wait_all": (None, [P]),
    "voxel_pool_thread_count": (SZ, [P]),
    # error
    "voxel_last_error": (C.c_char_p, []),
    # extensions
    "voxel_gpu_available": (INT, []),
    "voxel_gpu_malloc": (P, [SZ]),
    "voxel_gpu_free": (None, [P]),
    "voxel_gpu_memcpy_h2d": (INT, [P, P, SZ]),
    "voxel_gpu_memcpy_d2h": (INT, [P, P, SZ]),
    "voxel_gpu_synchronize": (INT, []),
    "voxel_gpu_fill_uniform": (INT, [INT, P, SZ, U64, F64, F64, P]),
    "voxel_gpu_engine_create": (P, [INT]),
    "voxel_gpu_engine_add_segment_raw": (SZ, [P, P, SZ]),
    "voxel_gpu_engine_set_resident": (INT, [P, SZ, INT]),
    "voxel_gpu_engine_add_device_segment": (SZ, [P, P, SZ]),
    "voxel_gpu_jit_lowerable": (INT, [P, SZ]),
    "voxel_gpu_jit_run_r2": (F64, [P, SZ, P, SZ, U64, F64]),
    "voxel_gpu_engine_lowered": (INT, [P]),
    "voxel_gpu_set_expr_jit": (INT, [INT]),
    "voxel_gpu_sort_pairs_u64": (INT, [P, P, P, P, SZ, P]),
    "voxel_gpu_argsort_f64": (INT, [P, SZ, INT, P, P]),
    "voxel_gpu_gather_f64": (INT, [P, P, SZ, P, P]),
    "voxel_gpu_f64_to_i64": (INT, [P, SZ, P, P]),
    "voxel_gpu_filter_bitmap": (INT, [INT, P, SZ, INT, U64, U64, P, P, P]),
    "voxel_gpu_filter_bitmap2": (INT, [INT, P, P, SZ, INT, U64, INT, U64, P, P, P]),
    "voxel_gpu_bitmap_combine": (INT, [P, P, SZ, INT, P, P]),
    "voxel_gpu_select_from_bitmap": (INT, [P, SZ, P, P, P]),
    "voxel_gpu_compact": (INT, [INT, P, SZ, INT, U64, P, P, P, P]),
    "voxel_gpu_join_build": (P, [P, P, SZ, P]),
    "voxel_gpu_join_destroy": (None, [P]),
    "voxel_gpu_join_probe_sum": (INT, [P, P, SZ, P, P, P]),
    "voxel_gpu_join_probe_sum_launch": (INT, [P, P, SZ, P]),
    "voxel_gpu_join_probe_sum_result": (INT, [P, P, P, P]),
    "voxel_gpu_join_probe_pairs": (INT, [P, P, SZ, P, P, P, P]),
    "voxel_gpu_join_build_ex": (P, [P, P, SZ, INT, P, P]),
    "voxel_gpu_join_set_probe_partitioning": (INT, [P, INT]),
    "voxel_gpu_groupby_f64": (INT, [INT, P, P, SZ, SZ, P, P, P, P, P, P, P, P]),
    "voxel_gpu_groupby_f64_general": (INT, [INT, P, P, SZ, SZ, P, P, P, P, P, P, P, P]),
    "voxel_gpu_decode": (I64, [P, U64, U64, P, P]),
    "voxel_gpu_scan_encoded": (INT, [P, INT, U64, P, P]),
    "voxel_gpu_scan_valid": (INT, [P, P, P, P, P]),
    "voxel_gpu_nullable_reduce": (INT, [INT, P, P, SZ, P, P]),
    "voxel_gpu_nullable_compact": (INT, [INT, P, P, SZ, P, P, P]),
    "voxel_gpu_window_delta_f64": (I64, [P, SZ, P, SZ, P]),
    "voxel_gpu_window_sum_f64": (I64, [P, SZ, P, SZ, INT, P]),
    "voxel_gpu_window_mean_f64": (I64, [P, SZ, P, SZ, INT, P]),
    "voxel_gpu_window_stddev_f64": (I64, [P, SZ, P, SZ, INT, P]),
    "voxel_gpu_window_variance_f64": (I64, [P, SZ, P, SZ, INT, P]),
    "voxel_gpu_window_quantile_f64": (I64, [P, SZ, P, SZ, INT, F64, P]),
    "voxel_gpu_time_window_sum_f64": (I64, [P, P, SZ, I64, I64, P, SZ, P]),
    "voxel_gpu_agg_results": (INT, [P, P, P, P, P, P, P, SZ]),
    "voxel_gpu_mlp_param_count": (INT, []),
    "voxel_gpu_mlp_init": (INT, [P, U64]),
    "voxel_gpu_mlp_forward": (INT, [P, P, SZ, P, P]),
    "voxel_gpu_mlp_grad": (INT, [P, P, P, SZ, P, P, P]),
    "voxel_gpu_mlp_train": (INT, [P, P, P, SZ, INT, C.c_float, P, P]),
    "voxel_gpu_sort_pairs_u64_learned": (INT, [P, P, P, P, SZ, INT, P, P]),
    "voxel_gpu_learned_tune": (INT, [INT, INT, INT, INT]),
    "voxel_gpu_batch_create": (P, [SZ]),
    "voxel_gpu_batch_destroy": (None, [P]),
    "voxel_gpu_batch_config": (INT, [P, INT, U32, INT]),
    "voxel_gpu_batch_get_config": (INT, [P, P, P, P]),
    "voxel_gpu_batch_launch": (INT, [P, P, SZ]),
    "voxel_gpu_batch_done": (INT, [P]),
    "voxel_gpu_batch_wait": (INT, [P, P, P]),
    "voxel_gpu_batch_run": (INT, [P, P, SZ, P, P]),
    "voxel_gpu_batch_autotune": (INT, [P, P, SZ, INT, P, P]),
    "voxel_gpu_server_create": (P, [SZ, U64, INT]),
    "voxel_gpu_server_destroy": (None, [P]),
    "voxel_gpu_server_config": (INT, [P, INT, U32, INT]),
    "voxel_gpu_server_autotune": (INT, [P, P, SZ, INT, P, P]),
    "voxel_gpu_server_replay": (INT, [P, P, P, SZ, INT, P, P, P]),
    "voxel_gpu_server_start": (INT, [P]),
    "voxel_gpu_server_submit": (I64, [P, P]),
    "voxel_gpu_server_poll": (INT, [P, I64, P]),
    "voxel_gpu_server_wait": (INT, [P, I64, P, P]),
    "voxel_gpu_server_stop": (INT, [P, P]),
    "voxel_gpu_scan": (INT, [C.POINTER(ScanDesc), P, C.POINTER(Agg)]),
    "voxel_gpu_scan_launch": (INT, [C.POINTER(ScanDesc), P]),
    "voxel_gpu_scan_result": (INT, [P, C.POINTER(Agg)]),
    "voxel_gpu_scan_noalloc": (INT, [C.POINTER(ScanDesc), P, P]),
}

for _name, (_res, _args) in SIGS.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


def header_symbols() -> list[str]:
    """Every function name include/voxel_c.h declares."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(voxel_[a-z0-9_]+)\s*\(", text)))


def last_error() -> str:
    v = lib.voxel_last_error()
    return v.decode() if v else ""


def check(what: str = "") -> None:
    """Raises RuntimeError when the last library call reported an error."""
    e = last_error()
    if e:
        raise RuntimeError((what + ": " if what else "") + e)

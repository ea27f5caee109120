"""voxel_py -- the Python module surface of VoxelVM (reference:
bindings/voxel_bindings.cpp:25-396), backed by libvoxel_c on the GPU.

Only the hot-path members are GPU work: EngineF64/EngineI64 ``run`` and
``run_parallel``, ``jit_run``, ``sort_ascending``, ``HashAggregator``,
``HashJoin``, ``VectorFilterF64``.  Everything here calls the C ABI in
include/voxel_c.h through ctypes; nothing computes a query result in Python.

Differences from the pybind module (documented per member):
* ``add_segment`` keeps a reference to the array (the reference does not,
  voxel_bindings.cpp:107-110, leaving a dangling pointer) and requires a
  C-contiguous array of the engine's dtype (TypeError otherwise) instead of
  silently converting to a temporary;
* calls the GPU backend cannot lower raise RuntimeError with the reason.
"""
from __future__ import annotations

import ctypes as C
import enum
import os

import numpy as np

from ._lib import check, lib

__version__ = "1.0.0"

# Filter comparison constants, bytecode/opcodes.hpp:415-420 (voxel_bindings.cpp:48-51).
CMP_EQ, CMP_NE, CMP_LT, CMP_LE, CMP_GT, CMP_GE = 0, 1, 2, 3, 4, 5


class DataType(enum.IntEnum):
    """Column types the pybind module exports (voxel_bindings.cpp:29-39;
    values of core/types.hpp:32-53)."""
    Null, Bool, Int8, Int16, Int32, Int64 = 0, 1, 2, 3, 4, 5
    Uint8, Uint16, Uint32, Uint64, Float32, Float64 = 6, 7, 8, 9, 10, 11
    Decimal128, String, Binary, Date, Timestamp = 14, 20, 21, 30, 32


class Opcode(enum.IntEnum):
    """The opcode values the pybind module exports (voxel_bindings.cpp:42-50)."""
    NOP, HALT, MOV, ADD, SUB, MUL = 0x00, 0x01, 0x10, 0x20, 0x21, 0x22
    CMP, JMP, JZ, JNZ = 0x40, 0x50, 0x51, 0x52
    VLOAD, VSTORE, VADD, VFILTER_GT, VSUM, ADDF = 0x70, 0x71, 0x80, 0xC5, 0xD0, 0x2A


# py::enum_::export_values(): the members are module attributes too
globals().update({m.name: m for m in DataType})
globals().update({m.name: m for m in Opcode})


def _word(op: int, rd: int = 0, ra: int = 0, rb: int = 0, imm: int = 0) -> int:
    """[op:8][rd:4][ra:4][rb:4][imm12:12] (bytecode/instruction.hpp:64-71)."""
    return ((op & 0xFF) | ((rd & 0xF) << 8) | ((ra & 0xF) << 12) | ((rb & 0xF) << 16)
            | ((imm & 0xFFF) << 20))


class Instruction:
    """A 32-bit instruction word with the constructors voxel_py exposes."""

    __slots__ = ("raw",)

    def __init__(self, raw: int = 0):
        self.raw = int(raw) & 0xFFFFFFFF

    def op(self) -> int: return self.raw & 0xFF
    def rd(self) -> int: return (self.raw >> 8) & 0xF
    def ra(self) -> int: return (self.raw >> 12) & 0xF
    def rb(self) -> int: return (self.raw >> 16) & 0xF
    def imm12(self) -> int: return (self.raw >> 20) & 0xFFF

    def __repr__(self) -> str:
        return "Instruction(0x%08x)" % self.raw

    # control / scalar
    @staticmethod
    def nop(): return Instruction(_word(0x00))
    @staticmethod
    def halt(): return Instruction(_word(0x01))
    @staticmethod
    def mov(rd, imm): return Instruction(_word(0x10, rd, 0, 0, imm))
    @staticmethod
    def movr(rd, ra): return Instruction(_word(0x11, rd, ra))
    @staticmethod
    def add(rd, ra, rb): return Instruction(_word(0x20, rd, ra, rb))
    @staticmethod
    def sub(rd, ra, rb): return Instruction(_word(0x21, rd, ra, rb))
    @staticmethod
    def mul(rd, ra, rb): return Instruction(_word(0x22, rd, ra, rb))
    @staticmethod
    def addf(rd, ra, rb): return Instruction(_word(0x2A, rd, ra, rb))
    @staticmethod
    def cmp(ra, rb): return Instruction(_word(0x40, 0, ra, rb))
    @staticmethod
    def jmp(off): return Instruction(_word(0x50, imm=off))
    @staticmethod
    def jz(off): return Instruction(_word(0x51, imm=off))
    @staticmethod
    def jnz(off): return Instruction(_word(0x52, imm=off))
    @staticmethod
    def jle(off): return Instruction(_word(0x5A, imm=off))
    @staticmethod
    def jge(off): return Instruction(_word(0x5C, imm=off))
    # vector
    @staticmethod
    def vload(vd, ra, seg_id, count): return Instruction(_word(0x70, vd, ra, 0, ((seg_id & 0xF) << 8) | (count & 0xFF)))
    @staticmethod
    def vstore(vs, ra, seg_id, count): return Instruction(_word(0x71, vs, ra, 0, ((seg_id & 0xF) << 8) | (count & 0xFF)))
    @staticmethod
    def vadd(vd, va, vb): return Instruction(_word(0x80, vd, va, vb))
    @staticmethod
    def vsub(vd, va, vb): return Instruction(_word(0x81, vd, va, vb))
    @staticmethod
    def vmul(vd, va, vb): return Instruction(_word(0x82, vd, va, vb))
    @staticmethod
    def vmin(vd, va, vb): return Instruction(_word(0x87, vd, va, vb))
    @staticmethod
    def vmax(vd, va, vb): return Instruction(_word(0x88, vd, va, vb))
    @staticmethod
    def vfilter(vd, va, rb, mode): return Instruction(_word(0xC0, vd, va, rb, mode & 0x7))
    @staticmethod
    def vfilter_gt(vd, va, rb): return Instruction(_word(0xC5, vd, va, rb))
    @staticmethod
    def vfilter_ge(vd, va, rb): return Instruction(_word(0xC6, vd, va, rb))
    @staticmethod
    def vsum(rd, va): return Instruction(_word(0xD0, rd, va))
    @staticmethod
    def vred_min(rd, va): return Instruction(_word(0xD5, rd, va))
    @staticmethod
    def vred_max(rd, va): return Instruction(_word(0xD6, rd, va))
    @staticmethod
    def vcount(rd, va): return Instruction(_word(0xD7, rd, va))


def _code_array(code) -> np.ndarray:
    return np.ascontiguousarray(np.array([int(w) & 0xFFFFFFFF for w in code], dtype=np.uint32))


def _as_column(arr, dtype) -> np.ndarray:
    if not isinstance(arr, np.ndarray) or arr.dtype != dtype or not arr.flags.c_contiguous:
        raise TypeError("add_segment needs a C-contiguous numpy array of %s" % np.dtype(dtype).name)
    return arr


class _Engine:
    _dtype = np.float64
    _create = "voxel_engine_create_f64"
    _destroy = "voxel_engine_destroy"
    _add = "voxel_engine_add_segment"
    _load = "voxel_engine_load_program"
    _run = "voxel_engine_run"
    _set = "voxel_engine_set_scalar"
    _get = "voxel_engine_get_scalar"

    def __init__(self):
        self._h = getattr(lib, self._create)()
        self._keep = []          # borrowed arrays stay alive as long as the engine
        self._code = None

    def __del__(self, _lib=lib):          # the library outlives module teardown
        h = getattr(self, "_h", None)
        if h:
            getattr(_lib, self._destroy)(h)
            self._h = None

    def add_segment(self, arr) -> int:
        a = _as_column(arr, self._dtype)
        self._keep.append(a)
        return int(getattr(lib, self._add)(self._h, a.ctypes.data, a.size))

    def load_program(self, code) -> None:
        self._code = _code_array(code)
        getattr(lib, self._load)(self._h, self._code.ctypes.data, self._code.size)

    def run(self) -> None:
        getattr(lib, self._run)(self._h)
        check("run")

    def get_scalar(self, i: int) -> int:
        return int(getattr(lib, self._get)(self._h, i))

    def set_scalar(self, i: int, v: int) -> None:
        getattr(lib, self._set)(self._h, i, int(v) & 0xFFFFFFFFFFFFFFFF)


class EngineF64(_Engine):
    """Engine<f64> (voxel_bindings.cpp:105-171)."""
    k_lanes = 4

    def run_parallel(self, seg_id: int, threads: int = 0, result_reg: int = 0) -> None:
        lib.voxel_engine_run_parallel(self._h, seg_id, threads, result_reg)
        check("run_parallel")

    def get_scalar_f64(self, i: int) -> float:
        return float(lib.voxel_engine_get_scalar_f64(self._h, i))

    def set_scalar_f64(self, i: int, v: float) -> None:
        lib.voxel_engine_set_scalar_f64(self._h, i, float(v))

    # Window family (voxel_bindings.cpp:127-180, Engine<f64>::Window*,
    # exec/engine.hpp:183-300): host arrays in, the first N outputs written
    # into `out`, N returned; the work runs on the GPU.
    def window_delta(self, data, out) -> int:
        return _window(lib.voxel_gpu_window_delta_f64, data, out)

    def window_sum(self, data, out, window) -> int:
        return _window(lib.voxel_gpu_window_sum_f64, data, out, _u8(window))

    def window_mean(self, data, out, window) -> int:
        return _window(lib.voxel_gpu_window_mean_f64, data, out, _u8(window))

    def window_stddev(self, data, out, window) -> int:
        return _window(lib.voxel_gpu_window_stddev_f64, data, out, _u8(window))

    def window_variance(self, data, out, window) -> int:
        return _window(lib.voxel_gpu_window_variance_f64, data, out, _u8(window))

    def window_quantile(self, data, out, window, quantile) -> int:
        return _window(lib.voxel_gpu_window_quantile_f64, data, out, _u8(window), float(quantile))

    def time_window_sum(self, timestamps, values, start_time, duration_sec, out) -> int:
        ts = _convert(timestamps, np.int64)
        v = _convert(values, np.float64)
        o = _out_array(out)
        cnt = min(ts.size, v.size)
        dts, dv, do = _Dev(ts), _Dev(v), _Dev(o)
        n = int(lib.voxel_gpu_time_window_sum_f64(dts.ptr, dv.ptr, cnt, int(start_time), int(duration_sec), do.ptr,
                                                  o.size, None))
        if n < 0:
            check("time_window_sum")
        do.to_host(o)
        return n


class EngineI64(_Engine):
    """Engine<i64> (voxel_bindings.cpp:174-189)."""
    k_lanes = 4
    _dtype = np.int64
    _create = "voxel_engine_create_i64"
    _destroy = "voxel_engine_destroy_i64"
    _add = "voxel_engine_add_segment_i64"
    _load = "voxel_engine_load_program_i64"
    _run = "voxel_engine_run_i64"
    _set = "voxel_engine_set_scalar_i64"
    _get = "voxel_engine_get_scalar_i64"

    def get_scalar_i64(self, i: int) -> int:
        v = self.get_scalar(i)
        return v - (1 << 64) if v >= (1 << 63) else v


def jit_compile(code) -> dict:
    """voxel_bindings.cpp:329-340: reports whether the fused loop lowers.
    No machine code exists: code_size and code_ptr are 0."""
    c = _code_array(code)
    ok = bool(lib.voxel_gpu_jit_lowerable(c.ctypes.data, c.size))
    if not ok:
        return {"compiled": False}
    return {"compiled": True, "code_size": 0, "code_ptr": 0}


def jit_run(code, data, threshold: float, total_count: int) -> float:
    """voxel_bindings.cpp:342-366: R0 = R1 = 0, R2 = total_count,
    R3 = threshold, one segment of len(data) rows; returns R0 as f64."""
    c = _code_array(code)
    d = _as_column(data, np.float64)
    if not lib.voxel_gpu_jit_lowerable(c.ctypes.data, c.size):
        raise RuntimeError("JIT compilation failed")
    v = _jit_run_r2(c, d, float(threshold), int(total_count))
    check("jit_run")
    return v


def _jit_run_r2(code: np.ndarray, data: np.ndarray, threshold: float, r2: int) -> float:
    # voxel_jit_run sets R2 = data_len; the pybind form sets R2 = total_count
    # and passes len(data) as the segment's row count.  Rows at or past
    # len(data) read as zero; the loop bound comes from R2.
    if r2 == data.size:
        return float(lib.voxel_jit_run(code.ctypes.data, code.size, data.ctypes.data, data.size, threshold))
    return float(lib.voxel_gpu_jit_run_r2(code.ctypes.data, code.size, data.ctypes.data, data.size,
                                          C.c_uint64(r2 & 0xFFFFFFFFFFFFFFFF), threshold))


# ---------------------------------------------------------------------------
# Operators (voxel_bindings.cpp:224-300).  Host arrays in, host arrays out;
# the work runs on the GPU through the C ABI.
# ---------------------------------------------------------------------------
def _convert(arr, dtype) -> np.ndarray:
    """py::array_t<T> argument conversion (forcecast): any array-like becomes a
    C-contiguous array of T."""
    return np.ascontiguousarray(arr, dtype=dtype)


class _Dev:
    """A device copy of a host array (freed with the object)."""

    def __init__(self, host: np.ndarray):
        self.n = host.size
        self.bytes = max(host.nbytes, 8)
        self.ptr = lib.voxel_gpu_malloc(self.bytes)
        if not self.ptr:
            check("voxel_gpu_malloc")
            raise RuntimeError("voxel_gpu_malloc failed")
        if host.nbytes and lib.voxel_gpu_memcpy_h2d(self.ptr, host.ctypes.data, host.nbytes) != 0:
            check("voxel_gpu_memcpy_h2d")

    @staticmethod
    def empty(nbytes: int) -> "_Dev":
        return _Dev(np.zeros(max(nbytes, 8), np.uint8))

    def to_host(self, out: np.ndarray) -> np.ndarray:
        if out.nbytes and lib.voxel_gpu_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes) != 0:
            check("voxel_gpu_memcpy_d2h")
        return out

    def __del__(self, _lib=lib):          # the library outlives module teardown
        if getattr(self, "ptr", None):
            _lib.voxel_gpu_free(self.ptr)
            self.ptr = None


def _u8(v) -> int:
    """pybind11's u8 argument conversion: an int in 0..255, else TypeError."""
    if isinstance(v, (bool, float)) or not isinstance(v, (int, np.integer)) or not 0 <= int(v) <= 255:
        raise TypeError("incompatible function arguments: window must be an integer in 0..255")
    return int(v)


def _out_array(out) -> np.ndarray:
    """py::array_t<f64> output: written in place when it already is a
    C-contiguous f64 array, else into a converted copy (as pybind11 does)."""
    if isinstance(out, np.ndarray) and out.dtype == np.float64 and out.flags.c_contiguous and out.flags.writeable:
        return out
    return _convert(out, np.float64)


def _window(fn, data, out, *args) -> int:
    x = _convert(data, np.float64)
    o = _out_array(out)
    d = _Dev(x)
    od = _Dev.empty(o.nbytes)
    n = int(fn(d.ptr, x.size, od.ptr, o.size, *args, None))
    if n < 0:
        check("window")
    if n > 0 and lib.voxel_gpu_memcpy_d2h(o.ctypes.data, od.ptr, n * 8) != 0:
        check("voxel_gpu_memcpy_d2h")
    return n


class VectorFilterF64:
    """VectorFilter<f64> (ops/filter.hpp:42-78): apply_gt -> (bitmap words, pass count)."""

    def apply_gt(self, arr, threshold: float):
        x = _convert(arr, np.float64)
        words = np.zeros((x.size + 63) // 64, np.uint64)
        d = _Dev(x)
        w = _Dev.empty(words.nbytes)
        cnt = np.zeros(1, np.uint64)
        tb = int(np.array([threshold], np.float64).view(np.uint64)[0])
        if lib.voxel_gpu_filter_bitmap(0, d.ptr, x.size, CMP_GT, tb, 0, w.ptr, cnt.ctypes.data, None) != 0:
            check("apply_gt")
        w.to_host(words)
        return words, int(cnt[0])


class Arena:
    """Arena (voxel_bindings.cpp:235-237).  The GPU operators allocate device
    memory themselves; the object is kept for signature compatibility."""

    def __init__(self, block_size: int = 65536):
        self.block_size = block_size


class HashAggregator:
    """HashAggregator<u32, f64> (ops/aggregate.hpp:71-301) over voxel_agg_*."""

    def __init__(self, arena: Arena):
        self._arena = arena
        self.h = lib.voxel_agg_create()

    def __del__(self, _lib=lib):          # the library outlives module teardown
        if getattr(self, "h", None):
            _lib.voxel_agg_destroy(self.h)
            self.h = None

    def init(self, expected_groups: int) -> None:
        lib.voxel_agg_init(self.h, expected_groups)

    def accumulate(self, keys, values) -> None:
        k = _convert(keys, np.uint32)
        v = _convert(values, np.float64)
        lib.voxel_agg_accumulate(self.h, k.ctypes.data, v.ctypes.data, min(k.size, v.size))
        check("accumulate")

    def group_count(self) -> int:
        return int(lib.voxel_agg_group_count(self.h))


def sort_ascending(arr) -> np.ndarray:
    """SortOperator::SortAscending (ops/sort.hpp:506-523): stable argsort, u32."""
    x = _convert(arr, np.float64)
    out = np.zeros(x.size, np.uint32)
    lib.voxel_sort_ascending(x.ctypes.data, x.size, out.ctypes.data)
    check("sort_ascending")
    return out


def topk_select(arr, k: int, largest: bool) -> np.ndarray:
    """TopKSelect (ops/sort.hpp): the k largest (or smallest) values, sorted."""
    x = _convert(arr, np.float64)
    out = np.zeros(k, np.float64)
    lib.voxel_topk(x.ctypes.data, x.size, k, out.ctypes.data, 1 if largest else 0)
    check("topk_select")
    return out


class HashJoin:
    """HashJoin<u32> (ops/join.hpp:16-214) on the GPU join (keys widened to
    i64, which keeps key equality and the first-match rule)."""

    def __init__(self, arena: Arena):
        self._arena = arena
        self.h = None

    def __del__(self, _lib=lib):          # the library outlives module teardown
        if getattr(self, "h", None):
            _lib.voxel_gpu_join_destroy(self.h)
            self.h = None

    def build(self, keys) -> None:
        k = _convert(keys, np.uint32).astype(np.int64)
        d = _Dev(k)
        if self.h:
            lib.voxel_gpu_join_destroy(self.h)
        self.h = lib.voxel_gpu_join_build(d.ptr, None, k.size, None)
        if not self.h:
            check("HashJoin.build")
            raise RuntimeError("HashJoin.build failed")

    def probe(self, probe_keys):
        p = _convert(probe_keys, np.uint32).astype(np.int64)
        if self.h is None:
            return np.zeros(0, np.uint32), np.zeros(0, np.uint32), 0
        d = _Dev(p)
        left = _Dev.empty(p.size * 4)
        right = _Dev.empty(p.size * 4)
        m = np.zeros(1, np.uint64)
        if lib.voxel_gpu_join_probe_pairs(self.h, d.ptr, p.size, left.ptr, right.ptr, m.ctypes.data, None) != 0:
            check("HashJoin.probe")
        n = int(m[0])
        lo = left.to_host(np.zeros(n, np.uint32))
        ro = right.to_host(np.zeros(n, np.uint32))
        return lo, ro, n

    def contains(self, key: int) -> bool:
        _, _, n = self.probe(np.array([key], np.uint32))
        return n > 0


class SegmentF64:
    """Segment<f64> (data/segment.hpp:25-60; voxel_bindings.cpp:97-102): a
    borrowed view of f64 rows.  This one keeps its array alive."""

    def __init__(self, data=None, size=None):
        if data is None:
            self._a = np.zeros(0, np.float64)
        else:
            a = _as_column(data, np.float64)
            self._a = a if size is None else a[: int(size)]

    def __getitem__(self, i: int) -> float:
        return float(self._a[i])

    def __len__(self) -> int:
        return int(self._a.size)

    @property
    def size(self) -> int:
        return int(self._a.size)


class NullBitmap:
    """voxel::NullBitmap through the C ABI's voxel_null_* (bit 1 = valid;
    voxel_bindings.cpp:192-200)."""

    def __init__(self, count: int):
        self._n = int(count)
        self._h = lib.voxel_null_create(self._n)

    def __del__(self, _lib=lib):          # the library outlives module teardown
        h = getattr(self, "_h", None)
        if h:
            _lib.voxel_null_destroy(h)
            self._h = None

    def is_null(self, row: int) -> bool:
        return bool(lib.voxel_null_is_null(self._h, row))

    def set_null(self, row: int) -> None:
        lib.voxel_null_set_null(self._h, row)

    def set_valid(self, row: int) -> None:
        lib.voxel_null_set_valid(self._h, row)

    def null_count(self) -> int:
        return int(lib.voxel_null_null_count(self._h))

    def valid_count(self) -> int:
        return int(lib.voxel_null_valid_count(self._h))

    def set_all_valid(self) -> None:
        for r in range(self._n):
            lib.voxel_null_set_valid(self._h, r)

    def set_all_null(self) -> None:
        for r in range(self._n):
            lib.voxel_null_set_null(self._h, r)


def cpu_info() -> dict:
    """platform::QueryCpuInfo's fields (voxel_bindings.cpp:313-324) for the
    host this process runs on (Linux /proc and /sys)."""
    flags, phys = set(), set()
    try:
        with open("/proc/cpuinfo") as f:
            pid = core = None
            for line in f:
                k, _, v = line.partition(":")
                k = k.strip()
                if k == "flags" and not flags:
                    flags = set(v.split())
                elif k == "physical id":
                    pid = v.strip()
                elif k == "core id":
                    core = v.strip()
                    phys.add((pid, core))
    except OSError:
        pass

    def cache(level: int) -> int:
        base = "/sys/devices/system/cpu/cpu0/cache"
        try:
            for d in sorted(os.listdir(base)):
                p = os.path.join(base, d)
                with open(os.path.join(p, "level")) as f:
                    if int(f.read()) != level:
                        continue
                with open(os.path.join(p, "type")) as f:
                    if f.read().strip() == "Instruction":
                        continue
                with open(os.path.join(p, "size")) as f:
                    t = f.read().strip()
                return int(t[:-1]) * 1024 if t.endswith("K") else int(t)
        except (OSError, ValueError):
            pass
        return 0

    logical = os.cpu_count() or 0
    return {"logical_cores": logical, "physical_cores": len(phys) or logical,
            "l1_cache": cache(1), "l2_cache": cache(2), "l3_cache": cache(3),
            "has_avx2": "avx2" in flags, "has_avx512": "avx512f" in flags, "has_fma": "fma" in flags}


# Members of the pybind module that are outside this backend's path: host
# data-format, I/O and profiling classes.  Asking for one names the reason.
_ABSENT = {
    "DictionaryEncodingF64": "host-side column encodings are in voxel_amd.encoding (GPU DecodeBatch and scans)",
    "RLEEncodingF64": "host-side column encodings are in voxel_amd.encoding (GPU DecodeBatch and scans)",
    "ThreadPool": "host thread pools run no query here (the C ABI's voxel_pool_* remain)",
    "InstructionProfiler": "the per-opcode cycle profiler belongs to the CPU interpreter",
    "MmapSegment": "file I/O (mmap segments) is outside the GPU execution path",
    "ChunkedReader": "file I/O (chunked readers) is outside the GPU execution path",
    "StreamingEngine": "the reference's StreamingEngine never resets its offset registers (streaming.hpp:49-55); "
                       "batches of rows go through EngineF64.run",
}


def __getattr__(name):
    if name in _ABSENT:
        raise AttributeError("voxel_py.%s is not provided by the GPU backend: %s" % (name, _ABSENT[name]))
    raise AttributeError("module 'voxel_py' has no attribute %r" % name)

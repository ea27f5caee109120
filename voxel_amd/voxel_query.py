"""voxel_query -- the high-level query builder of VoxelVM
(reference: bindings/voxel_query.py:10-228), running on the GPU backend.

    import voxel_amd.voxel_query as vq
    vq.from_numpy(data).filter_gt(500.0).sum().run()

The builder emits the same instruction words as the reference builder and
executes them through ``voxel_py.EngineF64`` (libvoxel_c), so ``run()``
returns what the reference returns for the same data, including its quirks:

* ``count().run()`` yields the integer count's register bits read as an f64
  (voxel_query.py:162,191-193: ``_result_is_int`` is reset before the check);
* ``min()``/``max()`` fold group values into R0 = 0.0 with an integer compare
  of bit patterns (voxel_query.py:111-135), so ``min()`` of non-negative data
  is 0.0;
* the last ``filter_*`` threshold wins (each writes R3).

Use ``voxel_amd.scan`` for relational (SQL) results instead.
"""
from __future__ import annotations

import numpy as np

from . import voxel_py as _vx

# register roles of the emitted programs (voxel_query.py:14-22)
_ACC, _OFF, _TOT, _THR, _LANES = 0, 1, 2, 3, 4
_FIRST_FREE = 5

# filter word per comparison: GT/GE have dedicated opcodes, the rest use the
# generic VFILTER with the comparison code in imm12 (voxel_query.py:74-86)
_FILTER_WORD = {
    _vx.CMP_GT: lambda: _vx.Instruction.vfilter_gt(1, 0, _THR),
    _vx.CMP_GE: lambda: _vx.Instruction.vfilter_ge(1, 0, _THR),
}


class Query:
    """One column, a filter, and a reduction loop (voxel_query.py:10-193)."""

    def __init__(self):
        self._engine = None
        self._seg_id = 0
        self._code: list[int] = []
        self._next_reg = _FIRST_FREE
        self._cmp = None
        self._total_count = 0
        self._data = None

    # -- filters: store the threshold in R3 now, remember the comparison --
    def _filter(self, cmp: int, threshold) -> "Query":
        self._engine.set_scalar_f64(_THR, float(threshold))
        self._cmp = cmp
        return self

    def filter_gt(self, t): return self._filter(_vx.CMP_GT, t)
    def filter_ge(self, t): return self._filter(_vx.CMP_GE, t)
    def filter_eq(self, t): return self._filter(_vx.CMP_EQ, t)
    def filter_ne(self, t): return self._filter(_vx.CMP_NE, t)
    def filter_lt(self, t): return self._filter(_vx.CMP_LT, t)
    def filter_le(self, t): return self._filter(_vx.CMP_LE, t)

    # -- loop emission --
    def _load_and_filter(self) -> int:
        """VLOAD V0 (+ VFILTER into V1); returns the vector the reduction reads."""
        self._code.append(_vx.Instruction.vload(0, _OFF, self._seg_id, 0).raw)
        if self._cmp is None:
            return 0
        make = _FILTER_WORD.get(self._cmp)
        word = make() if make else _vx.Instruction.vfilter(1, 0, _THR, self._cmp)
        self._code.append(word.raw)
        return 1

    def _loop(self, reduce_word, accumulate_words, back: int) -> "Query":
        lanes = self._engine.k_lanes
        part = self._next_reg
        self._next_reg += 1
        v = self._load_and_filter()
        self._code.append(reduce_word(part, v).raw)
        for w in accumulate_words(part):
            self._code.append(w.raw)
        # ADD's third operand is a register index: R4, which the MOV at the
        # head of the program sets to the lane count (voxel_query.py:93,165)
        self._code.append(_vx.Instruction.add(_OFF, _OFF, lanes).raw)
        self._code.append(_vx.Instruction.cmp(_OFF, _TOT).raw)
        self._code.append(_vx.Instruction.jnz(back).raw)
        return self

    def sum(self):
        return self._loop(_vx.Instruction.vsum, lambda r: [_vx.Instruction.addf(_ACC, _ACC, r)], -6)

    def count(self):
        return self._loop(_vx.Instruction.vcount, lambda r: [_vx.Instruction.add(_ACC, _ACC, r)], -6)

    def _select(self, reduce_word, branch):
        return self._loop(reduce_word, lambda r: [_vx.Instruction.cmp(_ACC, r), branch(2),
                                                  _vx.Instruction.movr(_ACC, r)], -8)

    def min(self): return self._select(_vx.Instruction.vred_min, _vx.Instruction.jle)
    def max(self): return self._select(_vx.Instruction.vred_max, _vx.Instruction.jge)

    # -- operator requests served at run() time --
    def sort(self, ascending: bool = True):
        self._sort_ascending = ascending
        return self

    def topk(self, k: int, largest: bool = True):
        self._topk = (k, largest)
        return self

    def group_by(self, group_keys):
        self._group_keys = group_keys
        return self

    def run(self):
        """Finish the program (MOV R4,lanes first, HALT last), run it, and
        return the result the reference returns (voxel_query.py:155-193)."""
        eng = self._engine
        eng.set_scalar_f64(_ACC, 0.0)
        eng.set_scalar(_OFF, 0)
        eng.set_scalar(_TOT, self._total_count)
        self._code.insert(0, _vx.Instruction.mov(_LANES, eng.k_lanes).raw)
        if len(self._code) == 1:
            self.sum()
        self._code.append(_vx.Instruction.halt().raw)
        eng.load_program(self._code)
        eng.run()
        if hasattr(self, "_group_keys"):
            agg = _vx.HashAggregator(_vx.Arena())
            agg.init(128)
            agg.accumulate(self._group_keys, self._data)
            return agg.group_count()
        if hasattr(self, "_sort_ascending"):
            return _vx.sort_ascending(self._data)
        if hasattr(self, "_topk"):
            return _vx.topk_select(self._data, self._topk[0], self._topk[1])
        return eng.get_scalar_f64(_ACC)

    @property
    def code(self) -> list[int]:
        return list(self._code)


def from_numpy(data: np.ndarray) -> Query:
    """A query over a numpy column (converted to f64 as the reference does)."""
    if data.dtype != np.float64:
        data = data.astype(np.float64)
    data = np.ascontiguousarray(data)
    q = Query()
    q._engine = _vx.EngineF64()
    q._data = data
    q._seg_id = q._engine.add_segment(data)
    q._total_count = len(data)
    return q


def from_list(values) -> Query:
    return from_numpy(np.array(values, dtype=np.float64))


def filter_and_sum(data, threshold):
    return from_numpy(data).filter_gt(threshold).sum().run()


def filter_and_count(data, threshold):
    return from_numpy(data).filter_gt(threshold).count().run()


def group_sum(data, groups):
    q = from_numpy(data)
    q._group_keys = np.array(groups, dtype=np.uint32)
    return q.group_by(q._group_keys).run()

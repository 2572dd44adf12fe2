"""Autotuner: the engines' tile settings, chosen by timing per size bucket.

Trajectories are bit-identical for every setting (kicks are integer sums, and every list holds
all pairs within the cutoff), so tuning only changes speed.

- ``EngineTuner`` picks the single-system step kernel's list variant and work-group block
  (``GpuEngine3.tune`` / ``GpuEngine.tune``) per bucket (dimension, size class floor(log2 n),
  several species or one).  Each candidate runs a few steps on the caller's own system.
- ``BatchTuner`` picks, per (dimension, size class), the batched engine's list variant (half list
  within work-groups, or every pair at both ends) and cell occupancy (cells 1, 1.25 or 1.5 times
  rc + skin wide: fewer, fuller cells or more, emptier ones), then for the half list the block
  (work-group tile of 256, 512 or 1024 slots: larger blocks keep more pairs inside one work-group's
  LDS), timed on a sample of the caller's systems of that bucket, and installs them with
  ``BatchEngine.set_tuning``.

Choices are kept per bucket in memory and, if a path is given, in a JSON file.
"""
from __future__ import annotations

import json
import os
import time
from dataclasses import dataclass, field
from typing import Callable, Dict, Iterable, List, Optional, Sequence, Tuple

ENGINE_CANDIDATES: Tuple[Tuple[int, int], ...] = ((1, 4096), (1, 2048), (1, 1024), (1, 512), (1, 256), (0, 256))
BATCH_CANDIDATES: Tuple[Tuple[int, float], ...] = ((1, 1.0), (1, 1.25), (1, 1.5), (0, 1.0), (0, 1.25))
BATCH_BLOCKS: Tuple[int, ...] = (256, 512, 1024)


def size_class(n: int) -> int:
    """floor(log2 n): the size bucket's index (the batched engine's own rule)."""
    return max(0, int(n).bit_length() - 1)


def _load(path: Optional[str]) -> dict:
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def _save(path: Optional[str], data: dict) -> None:
    if not path:
        return
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


@dataclass
class Trial:
    setting: tuple
    rate: float  # atom-steps/s


@dataclass
class EngineTuner:
    """Per-bucket (variant, block) for the single-system engines."""

    path: Optional[str] = None
    candidates: Sequence[Tuple[int, int]] = ENGINE_CANDIDATES
    steps: int = 30
    warmup: int = 10
    table: Dict[str, List] = field(default_factory=dict)
    trials: Dict[str, List[Trial]] = field(default_factory=dict)

    def __post_init__(self):
        self.table = _load(self.path).get("engine", {})

    @staticmethod
    def bucket(dim: int, n: int, nspecies: int) -> str:
        return f"{dim}d/{size_class(n)}/{'multi' if nspecies > 1 else 'one'}"

    def choose(self, dim: int, n: int, nspecies: int, make_engine: Callable[[], object]) -> Tuple[int, int]:
        """The fastest (variant, block) for this bucket; `make_engine()` returns a fresh engine on
        the caller's state (Verlet mode), which each candidate steps and closes."""
        key = self.bucket(dim, n, nspecies)
        if key in self.table:
            v, b = self.table[key]
            return int(v), int(b)
        trials = []
        for variant, block in self.candidates:
            e = make_engine()
            try:
                e.tune(variant, block)
                e.step(self.warmup)
                t0 = time.perf_counter()
                e.step(self.steps)  # returns when the device is done
                el = time.perf_counter() - t0
            finally:
                e.close()
            trials.append(Trial((variant, block), n * self.steps / el))
        best = max(trials, key=lambda t: t.rate)
        self.table[key] = list(best.setting)
        self.trials[key] = trials
        data = _load(self.path)
        data["engine"] = self.table
        _save(self.path, data)
        return best.setting

    def tune_engine(self, engine_cls, sys, state, mode) -> Tuple[int, int]:
        """Choose for (sys, state) and return the setting; the caller applies it with tune()."""
        dim = 3 if len(sys.box_len) == 3 else 2
        return self.choose(dim, state.n(), sys.species(), lambda: engine_cls(sys, state, mode))


@dataclass
class BatchTuner:
    """Per-(dimension, size class) (variant, cell factor, block) for the batched engine: the
    (variant, cell factor) candidates at block 256 first, then the other blocks for the best of
    them if it is the half list (a block is tried only where the sample's mean size is at least
    four blocks, and kept only if it is block_margin faster)."""

    path: Optional[str] = None
    candidates: Sequence[Tuple[int, float]] = BATCH_CANDIDATES
    blocks: Sequence[int] = BATCH_BLOCKS
    block_margin: float = 0.03
    per_bucket: int = 8
    steps: int = 200
    chunk: int = 100
    table: Dict[str, List] = field(default_factory=dict)
    trials: Dict[str, List[Trial]] = field(default_factory=dict)
    seconds: float = 0.0

    def __post_init__(self):
        self.table = _load(self.path).get("batch", {})

    @staticmethod
    def bucket(dim: int, n: int) -> str:
        return f"{dim}d/{size_class(n)}"

    def tune(self, systems: Iterable[Tuple[object, object]], skin: float = 0.3) -> Dict[str, List]:
        """Time every candidate on up to `per_bucket` of the given (system, state) pairs of each
        bucket not yet in the table; returns the table."""
        from .batch import BatchEngine

        groups: Dict[str, List[Tuple[object, object]]] = {}
        for sys, st in systems:
            dim = 3 if len(sys.box_len) == 3 else 2
            key = self.bucket(dim, st.n())
            if key not in self.table and len(groups.setdefault(key, [])) < self.per_bucket:
                groups[key].append((sys, st))
        if not groups:
            return self.table
        t_start = time.perf_counter()
        probe = BatchEngine(budget_bytes=64 << 20, max_systems=1, graph_chunk=0)
        try:
            need = max(sum(probe.bytes_needed(s, st.n(), skin) for s, st in g) for g in groups.values())
        finally:
            probe.close()
        # the arena splits its budget 16 / 80 / 4 between slots, lists and cells: room for any mix
        eng = BatchEngine(budget_bytes=int(need * 8) + (64 << 20), max_systems=self.per_bucket, graph_chunk=self.chunk)
        try:
            for key, sample in sorted(groups.items()):
                dim, cls = int(key[0]), int(key.split("/")[1])
                atoms = sum(st.n() for _, st in sample)
                trials = []

                def trial(variant: int, factor: float, block: int) -> None:
                    eng.set_tuning(dim, cls, variant, factor, block)
                    ids = [eng.add(s, st, self.steps + self.chunk, skin) for s, st in sample]
                    if any(i < 0 for i in ids):
                        for i in ids:
                            if i >= 0:
                                eng.remove(i)
                        return
                    eng.step_all(self.chunk)  # first lists and graph capture
                    t0 = time.perf_counter()
                    eng.step_all(self.steps)
                    el = time.perf_counter() - t0
                    for i in ids:
                        eng.remove(i)
                    trials.append(Trial((variant, factor, block), atoms * self.steps / el))

                for variant, factor in self.candidates:
                    trial(variant, factor, self.blocks[0])
                if trials:
                    first = max(trials, key=lambda t: t.rate)
                    variant, factor, _ = first.setting
                    if variant != 0:
                        for block in self.blocks[1:]:
                            if atoms >= 4 * block * len(sample):
                                trial(variant, factor, block)
                    # a larger block must beat the smallest by a margin above the timing noise
                    best = max(trials, key=lambda t: t.rate if t.setting[2] == first.setting[2] else t.rate / (1.0 + self.block_margin))
                    self.table[key] = list(best.setting)
                    self.trials[key] = trials
        finally:
            eng.close()
        self.seconds += time.perf_counter() - t_start
        data = _load(self.path)
        data["batch"] = self.table
        _save(self.path, data)
        return self.table

    def install(self, engine) -> None:
        """Apply the table to a BatchEngine (for systems it admits from now on); entries written
        before blocks were tuned keep block 256."""
        for key, setting in self.table.items():
            variant, factor = setting[0], setting[1]
            block = setting[2] if len(setting) > 2 else self.blocks[0]
            dim, cls = int(key[0]), int(key.split("/")[1])
            engine.set_tuning(dim, cls, int(variant), float(factor), int(block))


def default_path() -> str:
    return os.path.join(os.path.expanduser("~"), ".cache", "micrographia_amd", "tuning.json")

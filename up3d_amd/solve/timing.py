"""Wall-clock accounting on the reference's `timings` schema
(pyfp3d/solve/timing.py:33-107): phases seed / assembly / precond /
linsolve / residual / kutta, plus wall and other = wall - sum(phases).

Device work is asynchronous, so a phase that ends with device work must
synchronise before it is closed; `phase(..., sync=True)` does that.
"""

from __future__ import annotations

import time
from typing import Dict, Iterable, List

PHASES = ("seed", "assembly", "precond", "linsolve", "residual", "kutta")


def new_timings() -> Dict[str, float]:
    t = {p: 0.0 for p in PHASES}
    t["wall"] = 0.0
    t["other"] = 0.0
    return t


class phase:
    """Accumulate the wall time of a `with` block into timings[key]."""

    __slots__ = ("_t", "_k", "_t0", "_sync")

    def __init__(self, timings: Dict[str, float], key: str, sync: bool = True):
        self._t = timings
        self._k = key
        self._sync = sync

    def __enter__(self):
        self._t0 = time.perf_counter()
        return self

    def __exit__(self, *exc):
        if self._sync:
            import torch

            if torch.cuda.is_available():
                torch.cuda.synchronize()
        self._t[self._k] += time.perf_counter() - self._t0
        return False


def finalize(timings: Dict[str, float], wall: float) -> Dict[str, float]:
    timings["wall"] = float(wall)
    timings["other"] = float(wall - sum(timings[p] for p in PHASES))
    return timings


def sum_timings(parts: Iterable[Dict[str, float]]) -> Dict[str, float]:
    total = new_timings()
    for p in parts:
        for k in total:
            total[k] += float(p.get(k, 0.0))
    return total


def step_delta(timings: Dict[str, float], prev: Dict[str, float]) -> Dict[str, float]:
    return {f"t_{p}": timings[p] - prev.get(p, 0.0) for p in PHASES}


def snapshot(timings: Dict[str, float]) -> Dict[str, float]:
    return {p: timings[p] for p in PHASES}


def records_to_columns(records: List[Dict]) -> Dict[str, list]:
    keys: List[str] = []
    for r in records:
        for k in r:
            if k not in keys:
                keys.append(k)
    return {k: [r.get(k) for r in records] for k in keys}

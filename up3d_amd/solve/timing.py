"""Phase accounting for the `timings` dict every nonlinear driver returns.

The schema (key names and meaning) is the reference's
(pyfp3d/solve/timing.py:33-107); the code here is this package's own:

    seed      Picard warm start inside a Newton driver
    assembly  matrix / Jacobian assembly and right-hand sides
    precond   preconditioner set-up (AMG hierarchy, factorisations)
    linsolve  Krylov and direct solves
    residual  state sweep, upwinding and residual evaluation
    kutta     Kutta targets and circulation updates
    wall      the driver's whole wall time
    other     wall minus the six phases, i.e. the time they leave unexplained

Phases are exclusive: no region is charged to two of them, so `other` is
never negative beyond timer jitter.  Device work is asynchronous, so a phase
that ends on a kernel launch has not ended until the device has; `phase`
waits for the device before it stops the clock (pass sync=False where the
caller has just synchronised).
"""

from __future__ import annotations

import math
import time
from contextlib import contextmanager
from typing import Dict, Iterable, Iterator, List, Mapping

PHASES = ("seed", "assembly", "precond", "linsolve", "residual", "kutta")
TOTALS = ("wall", "other")
SCHEMA = PHASES + TOTALS


def new_timings() -> Dict[str, float]:
    """Zeroed seconds for every key of the schema."""
    return dict.fromkeys(SCHEMA, 0.0)


def _device_idle() -> None:
    import torch

    if torch.cuda.is_available() and torch.cuda.is_initialized():
        torch.cuda.synchronize()


@contextmanager
def phase(timings: Dict[str, float], key: str, sync: bool = True) -> Iterator[None]:
    """Charge the wall time of a `with` block to timings[key]."""
    if key not in PHASES:
        raise KeyError(f"{key!r} is not a timing phase (one of {PHASES})")
    start = time.perf_counter()
    try:
        yield
    finally:
        if sync:
            _device_idle()
        timings[key] += time.perf_counter() - start


def finalize(timings: Dict[str, float], wall: float) -> Dict[str, float]:
    """Record the driver's wall time and the unexplained remainder."""
    wall = float(wall)
    accounted = math.fsum(timings[p] for p in PHASES)
    timings["wall"] = wall
    timings["other"] = wall - accounted
    return timings


def sum_timings(parts: Iterable[Mapping[str, float]]) -> Dict[str, float]:
    """Key-wise total of several drivers' timings (a continuation's levels)."""
    parts = list(parts)
    return {k: math.fsum(float(p.get(k, 0.0)) for p in parts) for k in SCHEMA}


def snapshot(timings: Mapping[str, float]) -> Dict[str, float]:
    """The phase counters as they stand (the `prev` of step_delta)."""
    return {p: float(timings[p]) for p in PHASES}


def step_delta(timings: Mapping[str, float],
               prev: Mapping[str, float]) -> Dict[str, float]:
    """Per-phase seconds since `prev`, keyed t_<phase> for step records."""
    now = snapshot(timings)
    return {"t_" + p: now[p] - float(prev.get(p, 0.0)) for p in PHASES}


def records_to_columns(records: Iterable[Mapping]) -> Dict[str, List]:
    """Step records (one dict per Newton step) as one list per key, keys in
    first-seen order, None where a record lacks the key."""
    records = list(records)
    names = dict.fromkeys(k for r in records for k in r)
    return {k: [r.get(k) for r in records] for k in names}

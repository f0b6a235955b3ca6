"""Autotuner for the hot path (no reference counterpart).

Two searches, each timed on the device with HIP events on the launch
stream, on the caller's own mesh and state:

  * tune_element_block: threads per block of the element-wise kernels
    (state, walk + rho_tilde, residual element terms, sensitivities,
    Jacobian element terms) over {64, 128, 192, 256}; the node gathers keep
    their fixed LDS row blocks.  The winner is applied process-wide
    (fp3d_set_element_block).
  * tune_gather_chunk: entries per LDS-staged chunk of the CSR matrix gather
    over {512, 1024} (fp3d_set_gather_chunk; the sums are unchanged).
  * tune_walk_width: upstream walks per lane over {1, 2, 3}
    (fp3d_set_walk_width; the selections are unchanged).
  * tune_linear_solver: GMRES restart x AMG Chebyshev degree x AMG coarsest
    size on one Newton linear system J d = -R, by wall time to the asked
    relative tolerance (AMG set-up included: it is paid once per rebuild).

Results are cached in a JSON file keyed by (n_tets, n_nodes, nnz, device
name): $UP3D_TUNE_CACHE, else ~/.cache/up3d_amd/tuning.json.
"""

from __future__ import annotations

import json
import os
import time
from pathlib import Path
from typing import Callable, Dict, Iterable, Optional, Sequence

import torch

from .. import _native

ELEMENT_BLOCKS = (64, 128, 192, 256)
GATHER_CHUNKS = (512, 1024)
WALK_WIDTHS = (1, 2, 3)


def _event_time(fn: Callable[[], None], reps: int) -> float:
    """Average milliseconds per call of fn on the current stream."""
    fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    s = torch.cuda.current_stream()
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def apply_element_block(threads: int) -> None:
    _native.call("fp3d_set_element_block", int(threads))


def tune_element_block(pipeline: Callable[[], None],
                       candidates: Sequence[int] = ELEMENT_BLOCKS,
                       reps: int = 5) -> Dict[str, object]:
    """Time `pipeline` (one linearisation pass through the element-wise
    kernels) at each block size, apply the fastest and return
    {"best": threads, "ms": {threads: ms}}."""
    prev = int(_native.lib().fp3d_get_element_block())
    times = {}
    try:
        for t in candidates:
            apply_element_block(t)
            times[int(t)] = _event_time(pipeline, reps)
    except Exception:
        apply_element_block(prev)
        raise
    best = min(times, key=times.get)
    apply_element_block(best)
    return {"best": best, "ms": times}


def apply_gather_chunk(rows: int) -> None:
    _native.call("fp3d_set_gather_chunk", int(rows))


def tune_gather_chunk(pipeline: Callable[[], None],
                      candidates: Sequence[int] = GATHER_CHUNKS,
                      reps: int = 5) -> Dict[str, object]:
    """Time `pipeline` (a pass through the matrix gather) at each staged
    chunk size, apply the fastest and return {"best": rows, "ms": {...}}."""
    prev = int(_native.lib().fp3d_get_gather_chunk())
    times = {}
    try:
        for c in candidates:
            apply_gather_chunk(c)
            times[int(c)] = _event_time(pipeline, reps)
    except Exception:
        apply_gather_chunk(prev)
        raise
    best = min(times, key=times.get)
    apply_gather_chunk(best)
    return {"best": best, "ms": times}


def apply_walk_width(walks: int) -> None:
    _native.call("fp3d_set_walk_width", int(walks))


def tune_walk_width(pipeline: Callable[[], None],
                    candidates: Sequence[int] = WALK_WIDTHS,
                    reps: int = 5) -> Dict[str, object]:
    """Time `pipeline` (a pass through the upstream walk) at each number of
    walks per lane, apply the fastest and return {"best": w, "ms": {...}}."""
    prev = int(_native.lib().fp3d_get_walk_width())
    times = {}
    try:
        for w in candidates:
            apply_walk_width(w)
            times[int(w)] = _event_time(pipeline, reps)
    except Exception:
        apply_walk_width(prev)
        raise
    best = min(times, key=times.get)
    apply_walk_width(best)
    return {"best": best, "ms": times}


def tune_linear_solver(matvec: Callable, rhs: torch.Tensor, A_pre,
                       precision: str = "f64", rtol: float = 1e-4,
                       restarts: Iterable[int] = (60, 100, 200),
                       degrees: Iterable[int] = (1, 2, 3),
                       max_coarse: Iterable[int] = (10, 1000),
                       maxiter: int = 10) -> Dict[str, object]:
    """Fastest (restart, Chebyshev degree, AMG coarsest size) for
    J d = rhs with matvec = J (device) and A_pre the host CSR matrix the AMG
    is built on.  Returns {"best": {...}, "trials": [...]}; a trial that
    does not reach rtol within maxiter restarts is ranked last."""
    from ..solve.amg import AMGHierarchy
    from ..solve.linear import gmres_dev

    basis = "f32" if precision == "f32" else "f64"
    trials = []
    for mc in max_coarse:
        for deg in degrees:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ml = AMGHierarchy(A_pre, precision=precision, degree=int(deg),
                              max_coarse=int(mc), seed=0)
            setup = time.perf_counter() - t0
            for rs in restarts:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                _, info, it = gmres_dev(matvec, rhs, ml.apply, rtol=rtol,
                                        restart=int(rs), maxiter=maxiter, basis=basis)
                torch.cuda.synchronize()
                solve = time.perf_counter() - t0
                trials.append({"restart": int(rs), "degree": int(deg),
                               "max_coarse": int(mc), "iterations": int(it),
                               "converged": info == 0, "setup_s": setup,
                               "solve_s": solve, "levels": list(ml.shapes)})
            del ml
    ok = [t for t in trials if t["converged"]] or trials
    best = min(ok, key=lambda t: t["solve_s"] + t["setup_s"] / 3.0)
    return {"best": {k: best[k] for k in ("restart", "degree", "max_coarse")},
            "trials": trials}


# Version of the cache entries (their layout, and the kernels whose launch
# parameters they hold: 3 = the four-lanes-per-row CSR gather); entries of
# another version, or missing a tuned dimension, are misses and are tuned
# again.
CACHE_VERSION = 3
TUNED_KEYS = ("element_block", "gather_chunk", "walk_width", "solver")


class TuningCache:
    """JSON cache of tuning results keyed by mesh shape and device."""

    def __init__(self, path: Optional[str] = None):
        default = Path.home() / ".cache" / "up3d_amd" / "tuning.json"
        self.path = Path(path or os.environ.get("UP3D_TUNE_CACHE", default))

    @staticmethod
    def key(n_tets: int, n_nodes: int, nnz: int) -> str:
        name = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "none"
        return f"{name}|{n_tets}|{n_nodes}|{nnz}"

    def load(self) -> Dict[str, dict]:
        try:
            return json.loads(self.path.read_text())
        except (OSError, ValueError):
            return {}

    def get(self, key: str) -> Optional[dict]:
        """The entry for `key` if it is complete and of this layout version."""
        e = self.load().get(key)
        if not isinstance(e, dict) or e.get("version") != CACHE_VERSION:
            return None
        if any(k not in e for k in TUNED_KEYS):
            return None
        return e

    def put(self, key: str, value: dict) -> None:
        data = self.load()
        data[key] = value
        try:
            self.path.parent.mkdir(parents=True, exist_ok=True)
            self.path.write_text(json.dumps(data, indent=1))
        except OSError:
            pass                      # a read-only home: the result is still returned


def tuned_newton_kwargs(entry: Optional[dict]) -> Dict[str, object]:
    """solve_newton keyword arguments from a cached tuning entry."""
    if not entry or "solver" not in entry:
        return {}
    b = entry["solver"]["best"]
    return {"gmres_restart": b["restart"], "amg_degree": b["degree"],
            "amg_max_coarse": b["max_coarse"]}


def tune_newton_workspace(ws, phi, m_inf: float, upwind_c: float = 1.5,
                          m_crit: float = 0.95, rho_floor: float = 0.05,
                          precond_matrix: str = "jacobian_spd", rtol: float = 1e-4,
                          restarts: Iterable[int] = (60, 100, 200),
                          degrees: Iterable[int] = (1, 2, 3),
                          max_coarse: Iterable[int] = (10, 1000),
                          element_blocks: Sequence[int] = ELEMENT_BLOCKS,
                          cache: Optional[TuningCache] = None) -> Dict[str, object]:
    """Tune a NewtonWorkspace at the state phi (full potential): element
    block size on the residual + Jacobian linearisation, then the linear
    solver on J d = -R with the AMG built on `precond_matrix`.  Cached
    (and returned from the cache when present)."""
    from ..device import to_dev
    from ..solve.common import csr_to_host, mask_dirichlet
    from ..solve.linear import scale

    st = ws.st
    key = TuningCache.key(st.n_tets, st.n_nodes, st.nnz)
    cache = cache or TuningCache()
    hit = cache.get(key)
    if hit is not None:
        apply_element_block(hit["element_block"]["best"])
        apply_gather_chunk(hit["gather_chunk"]["best"])
        apply_walk_width(hit["walk_width"]["best"])
        return dict(hit, cached=True)
    ws.set_mach(m_inf)
    s = ws.states[2]
    s.phi.copy_(phi if isinstance(phi, torch.Tensor) else to_dev(phi))
    s.phi.index_copy_(0, ws.dn, ws.dv)

    def linearise():
        ws.eval_residual(s, upwind_c, m_crit, rho_floor)
        ws.assemble_coupled(s, upwind_c, m_crit, rho_floor)

    eb = tune_element_block(linearise, element_blocks)
    gc = tune_gather_chunk(linearise)
    ww = tune_walk_width(linearise)
    linearise()
    upw = ws.upw
    se, _ = upw.sensitivities_at_dev(s.q2l, s.rho, s.up, ws.m_inf, upwind_c, m_crit,
                                     ws.gamma_air, rho_floor)
    if precond_matrix == "jacobian_spd":
        ws.op.jacobian_spd_dev(s.rt, s.grad, se, s.lim, 1.0 / ws.u_inf ** 2, 0.9,
                               ws.A_pic)
    else:
        ws.op.matrix_dev(s.rt, ws.A_pic)
    mask_dirichlet(st, ws.fm, ws.A_pic)
    A_pre = csr_to_host(st, ws.A_pic)
    rhs = torch.empty_like(s.R)
    scale(-1.0, s.R, rhs)
    prec = "f32" if ws.precision == "fp32" else "f64"
    sol = tune_linear_solver(ws.jac.matvec, rhs, A_pre, precision=prec, rtol=rtol,
                             restarts=restarts, degrees=degrees, max_coarse=max_coarse)
    entry = {"version": CACHE_VERSION, "element_block": eb, "gather_chunk": gc,
             "walk_width": ww, "solver": sol, "m_inf": float(m_inf),
             "precision": ws.precision}
    cache.put(key, entry)
    return dict(entry, cached=False)

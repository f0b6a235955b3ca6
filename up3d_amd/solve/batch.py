"""Lockstep execution of many Newton solves (the batched sweep of flow
conditions; no reference counterpart: the reference solves one condition
per call, pyfp3d/solve/newton.py:1144-1307).

Every solve is the single-condition driver itself run as a coroutine
(newton_steps / transonic_steps with stepped=True): it yields an
EvalRequest for each residual evaluation and an ArnoldiStep for each Krylov
step instead of running them.  The executor advances all live solves one
request at a time, in rounds:

  * the Arnoldi steps of a round are queued back to back -- through the
    batched C-ABI (one launch per kernel for all instances of one mesh:
    the Jacobian matvec, the orthogonalisation, the norm and the
    normalisation), with only the V-cycles left per instance (each
    instance has its own hierarchy); each solve then waits only for its
    own previous step, queued a round earlier, and does that step's Givens
    update on the host while the device runs the new steps (GMRES's
    one-step-ahead pipeline, shared by all solves);
  * the live residual evaluations of a round that share a mesh go through
    fp3d_batch_residual_eval (one launch per kernel of the evaluation chain
    for all of them, one synchronisation); the others run per instance;
  * a solve's host-only work (HostTask: the SA set-up of a new AMG
    hierarchy) runs on a worker thread while the other solves keep their
    rounds going; the solve resumes with its result in the first round
    after it finishes (host_workers=0 runs it inline).

A solve that returns leaves the round structure (the convergence mask is
the set of live coroutines).  Each solve queues exactly the device work, in
the same order, that it queues when run alone, so its result is the same
bits (tests/test_gpu_batch.py).
"""

from __future__ import annotations

import time
from collections import defaultdict
from typing import Dict, Hashable, List, Optional, Sequence, Tuple

import numpy as np
import torch

from .linear import ArnoldiStep, batch_key, batched_arnoldi
from .newton import EvalRequest, HostTask


class BatchExecutor:
    """Runs coroutines that yield EvalRequest / ArnoldiStep to completion.

    batched=True queues the Arnoldi steps of instances that share a mesh
    through the batched entry points (linear.batched_arnoldi); False queues
    each step by itself (the same kernels, one instance per launch).
    Statistics: rounds, Arnoldi steps and evaluations executed, the
    largest number of steps queued in one round and the batched groups."""

    def __init__(self, batched: bool = True, host_workers: int = 4):
        self.batched = bool(batched)
        self.host_workers = int(host_workers)
        self.n_host_tasks = 0
        self.n_rounds = 0
        self.n_steps = 0
        self.n_evals = 0
        self.max_width = 0
        self.n_batched_launch_groups = 0
        self.n_grouped_vcycles = 0
        self.n_batched_eval_groups = 0

    def run(self, coros: Dict[Hashable, object],
            on_done=None) -> Dict[Hashable, object]:
        """Advance every coroutine to its return value; on_done(key, value)
        is called as each finishes (it may add new coroutines with
        self.add)."""
        from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait

        self._coros = dict(coros)
        self._pending: Dict[Hashable, object] = {}
        self._results: Dict[Hashable, object] = {}
        self._tasks: Dict[Hashable, object] = {}
        self._on_done = on_done
        self._pool = (ThreadPoolExecutor(max_workers=self.host_workers)
                      if self.host_workers > 0 else None)
        try:
            for k in list(self._coros):
                self._advance(k, None)
            while self._pending or self._tasks:
                if self._tasks:
                    if not self._pending:     # nothing else to run: wait for one
                        wait(list(self._tasks.values()), return_when=FIRST_COMPLETED)
                    for k, fut in list(self._tasks.items()):
                        if fut.done():
                            del self._tasks[k]
                            self._advance(k, fut.result())
                if self._pending:
                    self._round()
        finally:
            if self._pool is not None:
                self._pool.shutdown(wait=True)
                self._pool = None
        return self._results

    def add(self, key, coro) -> None:
        """Start another coroutine (from on_done)."""
        if key in self._coros:
            raise KeyError(f"coroutine {key!r} already running")
        self._coros[key] = coro
        self._advance(key, None)

    def _advance(self, key, value) -> None:
        try:
            req = self._coros[key].send(value)
        except StopIteration as stop:
            del self._coros[key]
            self._results[key] = stop.value
            if self._on_done is not None:
                self._on_done(key, stop.value)
            return
        if isinstance(req, HostTask):
            self.n_host_tasks += 1
            if self._pool is None:
                self._advance(key, req.fn())
            else:
                self._tasks[key] = self._pool.submit(req.fn)
            return
        if not isinstance(req, (EvalRequest, ArnoldiStep)):
            raise TypeError(f"solve {key!r} yielded {type(req).__name__}")
        self._pending[key] = req

    def _round(self) -> None:
        self.n_rounds += 1
        items = list(self._pending.items())
        self._pending.clear()
        steps = [(k, r) for k, r in items if isinstance(r, ArnoldiStep)]
        evals = [(k, r) for k, r in items if isinstance(r, EvalRequest)]
        if steps:
            self.n_steps += len(steps)
            self.max_width = max(self.max_width, len(steps))
            if self.batched:
                groups = defaultdict(list)
                for k, r in steps:
                    groups[batch_key(r)].append(r)
                for key, rs in groups.items():
                    if key is not None and len(rs) > 1:
                        self.n_grouped_vcycles += batched_arnoldi(rs)
                        self.n_batched_launch_groups += 1
                    else:
                        for r in rs:
                            r.run()
            else:
                for _, r in steps:
                    r.run()
            # no synchronisation: each solve waits for its own earlier step
            # (queued a round before) while these run on the device
            for k, _ in steps:
                self._advance(k, None)
        groups = defaultdict(list)
        for k, r in evals:
            self.n_evals += 1
            groups[eval_key(r) if self.batched else None].append((k, r))
        for key, items in groups.items():
            if key is not None and len(items) > 1:
                batched_eval([r for _, r in items])
                self.n_batched_eval_groups += 1
            else:
                for _, r in items:
                    r.ws.eval_residual(r.s, *r.args)
            for k, r in items:
                self._advance(k, r.s)

    def stats(self) -> dict:
        return {"rounds": self.n_rounds, "arnoldi_steps": self.n_steps,
                "evaluations": self.n_evals, "max_steps_per_round": self.max_width,
                "batched_groups": self.n_batched_launch_groups,
                "grouped_vcycles": self.n_grouped_vcycles,
                "batched_eval_groups": self.n_batched_eval_groups,
                "host_tasks": self.n_host_tasks}


_EVAL_FIELDS = 24      # int64 words per fp3d_batch_residual_eval job


def eval_key(req: EvalRequest):
    """Residual evaluations that one batched launch can serve share this
    key: live evaluations (no frozen selection) with M_inf > 0 and no
    external right-hand side, on the same mesh store with the same u_inf,
    gamma, upwind_c, m_crit and rho_floor.  None otherwise."""
    ws = req.ws
    args = req.args
    frozen = args[3] if len(args) > 3 else None
    if frozen is not None or ws.m_inf is None or ws.m_inf <= 0.0 or ws.b_ext is not None:
        return None
    return (id(ws.st), float(ws.u_inf), float(ws.gamma_air)) + tuple(float(a) for a in args[:3])


def _bits(x: float) -> int:
    return int(np.array([x], dtype=np.float64).view(np.int64)[0])


def batched_eval(reqs) -> None:
    """Run the live residual evaluations of several solves that share an
    eval_key through fp3d_batch_residual_eval and fill each state as
    NewtonWorkspace.eval_residual does (one synchronisation for all)."""
    from .. import _native
    from ..device import empty, ptr, stream_handle, to_host

    ws0 = reqs[0].ws
    st = ws0.st
    upwind_c, m_crit, rho_floor = (float(a) for a in reqs[0].args[:3])
    outs = empty((len(reqs), 8))
    jobs = np.zeros((len(reqs), _EVAL_FIELDS), dtype=np.int64)
    for i, r in enumerate(reqs):
        ws, s = r.ws, r.s
        work = getattr(ws, "_batch_eval_work", None)
        if work is None:
            work = ws._batch_eval_work = empty(3 * 1024)
        jobs[i, :20] = (ptr(s.phi), ptr(s.grad), ptr(s.q2n), ptr(s.q2l), ptr(s.rho),
                        ptr(s.lim), ptr(s.gb), ptr(s.up), ptr(ws.upw.d_nu), ptr(s.rt),
                        ptr(s.se), ptr(s.su), ptr(ws.op.d_terms), ptr(ws.op.d_R), ptr(s.R),
                        ptr(ws.fm), ptr(work), ptr(outs[i]), _bits(ws.m_inf),
                        _bits(ws.q2cap))
    host = torch.from_numpy(jobs).pin_memory()
    dev = torch.empty_like(host, device=st.device)
    dev.copy_(host, non_blocking=True)
    _native.call("fp3d_batch_residual_eval", len(reqs), ptr(dev), st.n_tets, st.n_nodes,
                 ptr(st.el), ptr(st.B), ptr(st.V), ptr(st.fn), ptr(st.dcen), ptr(st.nodes),
                 ptr(st.inc_ptr), ptr(st.inc_pos), float(ws0.u_inf ** 2),
                 float(ws0.gamma_air), m_crit, upwind_c, rho_floor, 4, stream_handle())
    h = to_host(outs)
    for i, r in enumerate(reqs):
        ws, s = r.ws, r.s
        upw = ws.upw
        s.n_limited = int(h[i, 0])
        upw.nu_max, upw.n_supersonic, upw.n_floored = float(h[i, 1]), int(h[i, 2]), int(h[i, 3])
        s.sens_key = (ws.m_inf, upwind_c, m_crit, ws.gamma_air, rho_floor)
        s.n_floored = upw.n_floored
        s.nu_max = upw.nu_max
        s.n_nu_active = upw.n_supersonic
        s.r_norm = float(h[i, 5])
        s.merit = float(h[i, 4]) * float(h[i, 4])


# ---------------------------------------------------------------- the sweep
def solve_sweep_batched(meshes: Sequence, conditions: Sequence[Tuple[int, float, float]],
                        chains_per_mesh=4, precision: str = "fp32",
                        upwind_c: float = 2.0, dm: float = 0.05,
                        newton_kw: Optional[dict] = None, batched: bool = True,
                        log=None, coarse=None, coarse_min_tets: int = 500_000,
                        coarse_min_mach: float = 0.0,
                        mesh_by_mesh: bool = False) -> Dict[str, object]:
    """Every (mesh index, M_inf, alpha_deg) condition, solved in lockstep.

    The conditions of each mesh are ordered by (M_inf, alpha) and dealt
    round-robin into chains_per_mesh chains (one count, or one per mesh); a chain solves its conditions
    one after another, each warm-started from the chain's last converged
    condition as in solve_sweep (alpha continuation, then a Mach ramp; a
    cold ramp when that fails), and all chains of all meshes run side by
    side through one BatchExecutor.  A condition whose chain could not
    converge it is retried once at the end from the nearest converged
    condition of its mesh with a finer Mach step (as solve_sweep).  coarse:
    a companion_mesh() pair enabling solve_sweep's coarse route.
    mesh_by_mesh=True advances the chains of one mesh at a time (one
    executor run per mesh, meshes in order), so a round never mixes a small
    mesh's steps with a large mesh's.
    Returns solve_sweep's keys plus "executor" statistics."""
    from ..mesh.channel import boundary_nodes
    from ..post.surface import wall_force_coefficients
    from .sweep import _record, _solve_one_steps, coarse_route

    kw = {"precision": precision, "gmres_restart": 200, "gmres_maxiter": 5,
          "n_newton_max": 30, "amg_rebuild_every": 0,
          "precond_matrix": "jacobian_spd", "amg_max_coarse": 1000}
    kw.update(newton_kw or {})
    t_all = time.perf_counter()
    inst: List[Optional[dict]] = [None] * len(conditions)
    phis: Dict[int, np.ndarray] = {}
    dns = [boundary_nodes(m, ["inflow", "outflow"]) for m in meshes]
    t_mesh0 = {}
    t_mesh1 = {}

    def chain(k, ids):
        last = None
        for i in ids:
            _, m, a = conditions[i]
            t0 = time.perf_counter()
            r, route, subs = yield from _solve_one_steps(
                meshes[k], dns[k], m, a, last, upwind_c, dm, kw, False, stepped=True,
                coarse=coarse_route(meshes[k], coarse, m, coarse_min_tets,
                                    coarse_min_mach))
            inst[i] = _record(meshes[k], k, m, a, r, subs, route,
                              time.perf_counter() - t0, False, wall_force_coefficients)
            if r["converged"]:
                last = (r["phi"], m, a)
                phis[i] = r["phi"]
            if log is not None:
                log(f"[batch] mesh {k} M={m:.4f} alpha={a:.3f} "
                    f"converged={r['converged']} {inst[i]['seconds']:.2f}s")
        t_mesh1[k] = time.perf_counter()
        return None

    nch = ([int(chains_per_mesh)] * len(meshes) if np.isscalar(chains_per_mesh)
           else [int(c) for c in chains_per_mesh])
    if len(nch) != len(meshes) or min(nch) < 1:
        raise ValueError("chains_per_mesh: a positive count, or one per mesh")
    ex = BatchExecutor(batched=batched)
    groups = [[k] for k in range(len(meshes))] if mesh_by_mesh else [list(range(len(meshes)))]
    for group in groups:
        coros = {}
        for k in group:
            ids = [i for i, c in enumerate(conditions) if c[0] == k]
            ids.sort(key=lambda i: (conditions[i][1], conditions[i][2]))
            t_mesh0[k] = time.perf_counter()
            for c in range(min(nch[k], len(ids))):
                coros[(k, c)] = chain(k, ids[c::nch[k]])
        ex.run(coros)

    # retries: from the nearest converged condition of the mesh, finer step
    retry = {}
    for i, rec in enumerate(inst):
        if rec is None or rec["converged"]:
            continue
        k, m, a = conditions[i]
        done = [j for j in phis if conditions[j][0] == k and conditions[j][1] <= m]
        src = (min(done, key=lambda j: (m - conditions[j][1]) / 0.05
                   + abs(a - conditions[j][2]) / 0.5) if done else None)
        seed = None if src is None else (phis[src], conditions[src][1], conditions[src][2])

        def one(i=i, k=k, m=m, a=a, seed=seed):
            t0 = time.perf_counter()
            r, route, subs = yield from _solve_one_steps(
                meshes[k], dns[k], m, a, seed, upwind_c, dm / 4.0, kw, False,
                dalpha_max=0.25, stepped=True)
            inst[i] = _record(meshes[k], k, m, a, r, subs, route,
                              time.perf_counter() - t0, True, wall_force_coefficients)
            if r["converged"]:
                phis[i] = r["phi"]
            return None

        retry[("retry", i)] = one()
    if retry:
        ex.run(retry)
    mask = [bool(r["converged"]) for r in inst]
    per_mesh = [{"mesh": k, "n_tets": int(len(meshes[k].elements)),
                 "n_conditions": sum(1 for c in conditions if c[0] == k),
                 "seconds": t_mesh1.get(k, t_mesh0[k]) - t_mesh0[k]}
                for k in range(len(meshes))]
    return {"instances": inst, "converged_mask": mask,
            "n_converged": int(sum(mask)), "wall_s": time.perf_counter() - t_all,
            "per_mesh": per_mesh, "executor": ex.stats(),
            "chains_per_mesh": nch}

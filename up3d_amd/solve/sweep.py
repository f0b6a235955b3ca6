"""Batched sweep of flow conditions over meshes of uneven size (no
reference counterpart: the reference solves one condition per call).

The conditions of one mesh share everything that depends only on the
mesh: the device-resident store (geometry, colouring, CSR pattern, node
incidence, upwind graphs), the Term-3 and Krylov workspaces and the
tuned solver parameters.  Each condition is an instance with its own
state and convergence flag:

  * the instances of a mesh are ordered by (M_inf, alpha) and each is
    warm-started from the last converged instance of its mesh: alpha
    continuation at the neighbour's Mach (steps of at most 0.5 deg), then a
    short Mach ramp from the neighbour's M_inf instead of a cold ramp from
    0.6; when that fails the instance takes the cold ramp;
  * the convergence mask records which instances reached
    ||R_free||_inf < tol with 0 limited / floored elements; instances that
    failed are retried once at the end, warm-started from the nearest
    converged instance of their mesh with a finer Mach step.

Boundary conditions are the transonic channel's: Dirichlet
phi = x cos(alpha) + y sin(alpha) on the inflow and outflow planes.
"""

from __future__ import annotations

import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from ..constraints.dirichlet import freestream_phi
from ..mesh.channel import boundary_nodes, bump_channel
from .linear import drive
from .newton import transonic_steps

SWEEP_SHAPES = ((60, 20, 20), (96, 32, 32), (120, 40, 40), (153, 51, 51))


def sweep_meshes(shapes=SWEEP_SHAPES, jitter: float = 0.15, seed: int = 20260104,
                 rcm: bool = True):
    """The four jittered bump channels of the sweep (144,000 / 589,824 /
    1,152,000 / 2,387,718 tets)."""
    return [bump_channel(*s, jitter=jitter, seed=seed + k, rcm=rcm)
            for k, s in enumerate(shapes)]


def sweep_conditions(n_meshes: int = 4, per_mesh: int = 16, seed: int = 20260105,
                     m_range=(0.60, 0.86), alpha_range=(0.0, 3.0)
                     ) -> List[Tuple[int, float, float]]:
    """(mesh index, M_inf, alpha_deg), M_inf and alpha uniform from a seed."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_meshes):
        ms = rng.uniform(*m_range, size=per_mesh)
        al = rng.uniform(*alpha_range, size=per_mesh)
        out.extend((k, float(m), float(a)) for m, a in zip(ms, al))
    return out


def _level(mesh, dn, m_inf, alpha, phi, m_start, upwind_c, dm, newton_kw, verbose):
    """One Mach ramp m_start -> m_inf at alpha from phi (None: freestream)."""
    return drive(_level_steps(mesh, dn, m_inf, alpha, phi, m_start, upwind_c, dm,
                              newton_kw, verbose))


def _level_steps(mesh, dn, m_inf, alpha, phi, m_start, upwind_c, dm, newton_kw, verbose,
                 stepped=False):
    """_level as a coroutine (transonic_steps)."""
    dv = freestream_phi(mesh.nodes[dn], alpha)
    kw = dict(newton_kw)
    phi0 = freestream_phi(mesh.nodes, alpha) if phi is None else np.array(phi, dtype=np.float64)
    phi0[dn] = dv
    kw["phi_init"] = phi0
    kw["n_picard_seed"] = 0
    return (yield from transonic_steps(mesh.nodes, mesh.elements, dn, dv, m_inf,
                                       m_start=min(m_inf, m_start), dm=dm, dm_min=0.005,
                                       upwind_c=upwind_c, newton_kw=kw, verbose=verbose,
                                       stepped=stepped))


def _solve_one_steps(mesh, dn, m_inf, alpha, seed, upwind_c, dm, newton_kw, verbose,
                     dalpha_max: float = 0.5, stepped: bool = False, coarse=None):
    """(result, route, sub-results), as a coroutine.  With coarse =
    (companion mesh, its Dirichlet nodes): first the "coarse" route -- a
    cold Mach ramp to (m_inf, alpha) on the companion mesh, its P1
    interpolant on this mesh, one Newton level at m_inf.  Then (or without
    it) from a converged neighbour seed = (phi, M, alpha): alpha
    continuation at the neighbour's Mach in steps of at most dalpha_max
    degrees, then a Mach ramp to m_inf; when that fails (or without a
    seed), a cold Mach ramp from 0.6 at this alpha."""
    subs = []
    if coarse is not None:
        from ..mesh.transfer import interpolate_p1

        cm, cdn = coarse
        rc = yield from _level_steps(cm, cdn, m_inf, alpha, None, 0.6, upwind_c, dm,
                                     newton_kw, verbose, stepped)
        subs.append(rc)
        if rc["converged"]:
            phi0 = interpolate_p1(cm.nodes, cm.elements, rc["phi"], mesh.nodes)
            r = yield from _level_steps(mesh, dn, m_inf, alpha, phi0, m_inf, upwind_c, dm,
                                        newton_kw, verbose, stepped)
            subs.append(r)
            if r["converged"]:
                return r, "coarse", subs
    if seed is not None:
        phi, m0, a0 = seed
        n_a = int(np.ceil(abs(alpha - a0) / dalpha_max - 1e-12))
        ok = True
        for j in range(1, n_a + 1):
            r = yield from _level_steps(mesh, dn, m0, a0 + (alpha - a0) * j / n_a, phi, m0,
                                  upwind_c, dm, newton_kw, verbose, stepped)
            subs.append(r)
            if not r["converged"]:
                ok = False
                break
            phi = r["phi"]
        if ok:
            r = yield from _level_steps(mesh, dn, m_inf, alpha, phi, max(m0, 0.6), upwind_c, dm,
                                  newton_kw, verbose, stepped)
            subs.append(r)
            if r["converged"]:
                return r, "warm", subs
    r = yield from _level_steps(mesh, dn, m_inf, alpha, None, 0.6, upwind_c, dm, newton_kw,
                          verbose, stepped)
    subs.append(r)
    return r, "cold", subs


def _solve_one(mesh, dn, m_inf, alpha, seed, upwind_c, dm, newton_kw, verbose,
               dalpha_max: float = 0.5, coarse=None):
    return drive(_solve_one_steps(mesh, dn, m_inf, alpha, seed, upwind_c, dm, newton_kw,
                                  verbose, dalpha_max, coarse=coarse))


def coarse_route(mesh, coarse, m_inf, min_tets: int = 500_000, m_min: float = 0.0):
    """The companion (mesh, Dirichlet nodes) pair for the coarse route of a
    condition, or None: used for conditions with m_inf >= m_min on meshes of
    at least min_tets tets, where a Mach or alpha continuation on the mesh
    itself costs many of its Newton steps (a shock forming and moving across
    it, alpha jumps of several continuation levels)."""
    if coarse is None or len(mesh.elements) < min_tets or m_inf < m_min:
        return None
    return coarse


def companion_mesh(shape=(48, 16, 16), jitter: float = 0.15, seed: int = 20260110,
                   rcm: bool = True):
    """(mesh, Dirichlet nodes) of the coarse channel the coarse route
    solves on (73,728 tets)."""
    m = bump_channel(*shape, jitter=jitter, seed=seed, rcm=rcm)
    return m, boundary_nodes(m, ["inflow", "outflow"])


def solve_sweep(meshes: Sequence, conditions: Sequence[Tuple[int, float, float]],
                precision: str = "fp32", upwind_c: float = 2.0, dm: float = 0.05,
                newton_kw: Optional[dict] = None, retry: bool = True,
                log=None, verbose: bool = False, coarse=None,
                coarse_min_tets: int = 500_000,
                coarse_min_mach: float = 0.0) -> Dict[str, object]:
    """Solve every (mesh index, M_inf, alpha_deg) condition (coarse: a
    companion_mesh() pair that enables the coarse route on meshes of at
    least coarse_min_tets tets for M_inf >= coarse_min_mach, see
    coarse_route).  Returns
    {"instances": [...], "converged_mask", "n_converged", "wall_s",
    "per_mesh": [...]} (log: optional callable taking one progress line per
    instance) with, per instance, its mesh, M_inf, alpha, converged,
    residual, Newton steps, linear iterations, seconds, wall force
    (cf_x, cf_y on the floor) and whether it was a retry."""
    from ..post.surface import wall_force_coefficients

    kw = {"precision": precision, "gmres_restart": 200, "gmres_maxiter": 5,
          "n_newton_max": 30, "amg_rebuild_every": 0,
          "precond_matrix": "jacobian_spd", "amg_max_coarse": 1000}
    kw.update(newton_kw or {})
    t_all = time.perf_counter()
    inst = [None] * len(conditions)
    phis: Dict[int, np.ndarray] = {}
    per_mesh = []
    for k, mesh in enumerate(meshes):
        dn = boundary_nodes(mesh, ["inflow", "outflow"])
        ids = [i for i, c in enumerate(conditions) if c[0] == k]
        ids.sort(key=lambda i: (conditions[i][1], conditions[i][2]))
        t_mesh = time.perf_counter()
        last = None                      # (phi, M) of the last converged instance
        for i in ids:
            _, m, a = conditions[i]
            t0 = time.perf_counter()
            r, route, subs = _solve_one(mesh, dn, m, a, last, upwind_c, dm, kw, verbose,
                                        coarse=coarse_route(mesh, coarse, m,
                                                            coarse_min_tets,
                                                            coarse_min_mach))
            if torch.cuda.is_available():
                torch.cuda.synchronize()
            inst[i] = _record(mesh, k, m, a, r, subs, route,
                              time.perf_counter() - t0, False, wall_force_coefficients)
            if r["converged"]:
                last = (r["phi"], m, a)
                phis[i] = r["phi"]
            if log is not None:
                log(f"[sweep] mesh {k} ({len(mesh.elements)} tets) M={m:.4f} "
                    f"alpha={a:.3f} converged={r['converged']} "
                    f"{inst[i]['seconds']:.2f}s")
        per_mesh.append({"mesh": k, "n_tets": int(len(mesh.elements)),
                         "n_conditions": len(ids),
                         "seconds": time.perf_counter() - t_mesh})
    if retry:
        for i, rec in enumerate(inst):
            if rec is None or rec["converged"]:
                continue
            k, m, a = conditions[i]
            mesh = meshes[k]
            dn = boundary_nodes(mesh, ["inflow", "outflow"])
            done = [j for j in phis if conditions[j][0] == k and conditions[j][1] <= m]
            src = (min(done, key=lambda j: (m - conditions[j][1]) / 0.05
                       + abs(a - conditions[j][2]) / 0.5) if done else None)
            seed = None if src is None else (phis[src], conditions[src][1], conditions[src][2])
            t0 = time.perf_counter()
            r, route, subs = _solve_one(mesh, dn, m, a, seed, upwind_c, dm / 4.0, kw,
                                        verbose, dalpha_max=0.25)
            if torch.cuda.is_available():
                torch.cuda.synchronize()
            inst[i] = _record(mesh, k, m, a, r, subs, route,
                              time.perf_counter() - t0, True, wall_force_coefficients)
            if r["converged"]:
                phis[i] = r["phi"]
            if log is not None:
                log(f"[sweep] retry mesh {k} M={m:.4f} alpha={a:.3f} "
                    f"converged={r['converged']}")
    mask = [bool(r["converged"]) for r in inst]
    return {"instances": inst, "converged_mask": mask,
            "n_converged": int(sum(mask)), "wall_s": time.perf_counter() - t_all,
            "per_mesh": per_mesh}


def _record(mesh, k, m, a, r, subs, route, seconds, retried, wall_force_coefficients):
    wf = wall_force_coefficients(mesh.nodes, mesh.elements,
                                 mesh.boundary_faces["floor"], r["phi"], m_inf=m,
                                 alpha_deg=a)
    return {"mesh": k, "n_tets": int(len(mesh.elements)), "m_inf": m, "alpha_deg": a,
            "converged": bool(r["converged"]), "accept_reason": r.get("accept_reason"),
            "residual_inf": float(r["residual_history"][-1]) if r.get("residual_history") else None,
            "n_newton": int(sum(lr["n_newton"] for q in subs for lr in q["level_results"])),
            "n_lin_iters": int(sum(lr["n_lin_iters"] for q in subs
                                   for lr in q["level_results"])),
            "n_levels": int(sum(len(q["level_results"]) for q in subs)),
            "route": route, "seconds": seconds,
            "cf": [float(x) for x in wf["cf"]], "retried": retried}

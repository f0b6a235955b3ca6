"""Grid-sequenced transonic Newton solve.

(New functionality without a reference counterpart; the per-mesh solves
are the reference's Mach continuation, pyfp3d/solve/newton.py:1144-1307,
and its `upwind_c_post` dissipation stages.)  The coarsest mesh runs the
upward Mach ramp to the target M_inf at a dissipative upwind coefficient;
each finer mesh starts AT the target Mach from the P1 interpolant of the
previous mesh's potential, and the finest mesh then lowers upwind_c through
`upwind_c_post` to the production value.  The converged state on the finest
mesh is that mesh's discrete solution under the same acceptance rules as
solve_newton (||R_free||_inf < tol with the freeze/honesty logic); only the
path to it changes.  On a fine mesh (millions of tets) the shock then needs
to move by about one coarse cell instead of travelling with every Mach
level, which is what lets the Newton steps run at full length.
"""

from __future__ import annotations

import time
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from ..mesh.transfer import interpolate_p1
from .newton import solve_newton_transonic
from .timing import sum_timings


def solve_newton_sequenced(meshes: Sequence[dict], m_inf: float,
                           m_start: float = 0.6, dm: float = 0.05,
                           dm_min: float = 0.005, upwind_c_ramp: float = 3.0,
                           upwind_c_post: Sequence[float] = (2.0, 1.5),
                           m_crit: float = 0.95,
                           post_stages_on: str = "finest",
                           newton_kw: Optional[dict] = None,
                           tune: bool = False, tune_kw: Optional[dict] = None,
                           verbose: bool = False) -> Dict[str, object]:
    """meshes: coarse-to-fine list of dicts with keys nodes, elements,
    dirichlet_nodes and either dirichlet_values or a callable
    dirichlet_fn(nodes) -> values.  Returns the finest mesh's result dict
    (solve_newton_transonic keys) plus `mesh_results` (per-mesh summary),
    `timings_total` (all meshes; the interpolations count as `seed`) and
    `wall_s`.

    post_stages_on="finest" runs the upwind_c_post stages on the finest mesh
    only; "coarser" runs them on every mesh after the first and lets the
    finest mesh start directly at the last (production) upwind_c from the
    interpolated production-C solution, so the expensive mesh runs one
    Newton level instead of len(upwind_c_post) + 1.

    tune=True runs `up3d_amd.tune.tune_newton_workspace` on the finest mesh
    at its interpolated initial state (element block size, GMRES restart,
    AMG degree and coarsest size; cached per mesh shape and device) and
    solves the finest mesh with the result; `tuning` in the returned dict
    holds the choice, whether it came from the cache and its seconds."""
    if post_stages_on not in ("finest", "coarser"):
        raise ValueError(f"unknown post_stages_on {post_stages_on!r}")
    if len(meshes) < 1:
        raise ValueError("need at least one mesh")
    kw = dict(newton_kw or {})
    t_all = time.perf_counter()
    phi_prev = None
    prev = None
    mesh_results: List[dict] = []
    timing_parts = []
    result = None
    for level, mdef in enumerate(meshes):
        nodes = np.asarray(mdef["nodes"], dtype=np.float64)
        elements = np.asarray(mdef["elements"])
        dn = np.asarray(mdef["dirichlet_nodes"])
        if "dirichlet_values" in mdef:
            dv = np.asarray(mdef["dirichlet_values"], dtype=np.float64)
        else:
            dv = np.asarray(mdef["dirichlet_fn"](nodes), dtype=np.float64)
        t_seed = 0.0
        if prev is None:
            phi0 = kw.pop("phi_init", None)
            if phi0 is None:
                phi0 = nodes[:, 0].copy()
        else:
            t0 = time.perf_counter()
            phi0 = interpolate_p1(prev[0], prev[1], phi_prev, nodes)
            phi0[dn] = dv
            t_seed = time.perf_counter() - t0
        finest = level == len(meshes) - 1
        lvl_kw = dict(kw, phi_init=phi0, n_picard_seed=0)
        c_level, post = upwind_c_ramp, None
        if post_stages_on == "finest":
            post = list(upwind_c_post) if finest else None
        elif prev is not None:
            if finest and len(upwind_c_post):
                c_level = upwind_c_post[-1]
            else:
                post = list(upwind_c_post)
        tuning = None
        if tune and finest:
            from ..tune import tune_newton_workspace, tuned_newton_kwargs
            from .newton import NewtonWorkspace

            t0 = time.perf_counter()
            ws_t = NewtonWorkspace(nodes, elements, dn, dv,
                                   precision=kw.get("precision", "fp64"))
            tuning = tune_newton_workspace(
                ws_t, phi0, m_inf, upwind_c=c_level, m_crit=m_crit,
                precond_matrix=kw.get("precond_matrix", "picard"),
                **(tune_kw or {}))
            del ws_t
            lvl_kw.update(tuned_newton_kwargs(tuning))
            if torch.cuda.is_available():
                torch.cuda.synchronize()
            tuning = {"element_block": tuning["element_block"]["best"],
                      "solver": tuning["solver"]["best"],
                      "cached": bool(tuning.get("cached")),
                      "seconds": time.perf_counter() - t0}
            if verbose:
                print(f"[sequence] tuning {tuning}", flush=True)
        t0 = time.perf_counter()
        result = solve_newton_transonic(
            nodes, elements, dn, dv, m_inf,
            m_start=m_start if prev is None else m_inf, dm=dm, dm_min=dm_min,
            upwind_c=c_level, m_crit=m_crit, upwind_c_post=post,
            newton_kw=lvl_kw, verbose=verbose)
        if torch.cuda.is_available():
            torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        tt = dict(result["timings_total"])
        tt["seed"] = tt.get("seed", 0.0) + t_seed
        tt["wall"] = tt.get("wall", 0.0) + t_seed
        timing_parts.append(tt)
        mesh_results.append({
            "n_nodes": len(nodes), "n_tets": len(elements),
            "converged": bool(result["converged"]), "wall_s": wall,
            "transfer_s": t_seed,
            "levels": [(lr["m"], lr["upwind_c"], lr["n_newton"],
                        lr["residual_history"][-1], lr["converged"],
                        lr["n_lin_iters"]) for lr in result["level_results"]],
        })
        if verbose:
            print(f"[sequence] mesh {level}: {len(elements)} tets, converged="
                  f"{result['converged']}, {wall:.1f}s", flush=True)
        # a coarse level that ends short of convergence still seeds the next
        phi_prev = result["phi"]
        prev = (nodes, elements)
    out = dict(result)
    out["tuning"] = tuning
    out["mesh_results"] = mesh_results
    out["timings_total"] = sum_timings(timing_parts)
    out["wall_s"] = time.perf_counter() - t_all
    return out

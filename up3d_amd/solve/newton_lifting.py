"""Coupled Newton on (phi, Gamma) over a conforming wake cut
(pyfp3d/solve/newton.py:97-1141 solve_newton_lifting with n_st > 0, and
:1144-1307 solve_newton_transonic's Mach ramp).

Unknowns: the reduced potential phi_red (original nodes; a slave is
phi_master + Gamma_station) and one circulation per station.  Equations:

    R_red(phi_red, Gamma) = T^T R_cut(T phi_red + G Gamma; far field(Gamma))
    F(phi_red, Gamma)     = kutta_targets(phi_cut) - Gamma

Newton step with Gamma eliminated (the reference's (J_ff + B K) form):
    dGamma = K dphi + F,   (J_red + B K) dphi = -R_red - B F
with J_red x = T^T J_cut T x (device: expand masters to slaves, the cut
mesh's Jacobian operator -- Terms 1+2 CSR and the Term-3 gather table --,
add slave rows into master rows), K the probe-difference rows and
B = dR_red/dGamma (the slave jump and the far-field vortex column) by a
central difference of the frozen-selection residual (one station: two
residual evaluations).  GMRES on the device, preconditioned by SA-AMG of
the reduced SPD matrix T^T A T (Picard or clamped Jacobian).
"""

from __future__ import annotations

import time
from typing import Dict, Optional

import numpy as np
import scipy.sparse as sp
import torch

from ..constraints.dirichlet import farfield_dirichlet
from ..device import empty, to_dev, to_host
from .amg import AMGHierarchy
from .common import csr_to_host
from .linear import gmres_dev
from .newton import NewtonWorkspace
from .timing import finalize, new_timings, phase


class LiftingNewtonSystem:
    def __init__(self, mesh_cut, wc, alpha_deg: float, u_inf: float = 1.0,
                 gamma_air: float = 1.4, vortex_center=(0.25, 0.0),
                 precision: str = "fp64"):
        self.mesh, self.wc = mesh_cut, wc
        self.alpha, self.u_inf, self.center = float(alpha_deg), float(u_inf), vortex_center
        ff, _ = farfield_dirichlet(mesh_cut, wc, alpha_deg, np.zeros(wc.n_stations), u_inf,
                                   vortex_center)
        self.ff = ff
        self.ws = NewtonWorkspace(mesh_cut.nodes, mesh_cut.elements, ff, np.zeros(len(ff)),
                                  u_inf, gamma_air, None, precision)
        self.n_red = n = wc.n_nodes_orig
        self.n_st = wc.n_stations
        self.dev = self.ws.st.device
        self.d_master = to_dev(wc.master_nodes, torch.int64)
        self.d_slave = to_dev(wc.slave_nodes, torch.int64)
        self.dir_red = ff[ff < n]
        free = np.ones(n, dtype=bool)
        free[self.dir_red] = False
        self.free = free
        self.d_free = to_dev(free.astype(np.float64))
        self.d_dir = to_dev(self.dir_red, torch.int64)
        self.d_up = to_dev(wc.kutta_upper, torch.int64)
        self.d_lo = to_dev(wc.kutta_lower, torch.int64)
        cnt = np.bincount(wc.te_station, minlength=self.n_st).astype(np.float64)
        self.k_w = to_dev(1.0 / cnt[wc.te_station])
        self.d_te_station = to_dev(wc.te_station, torch.int64)
        rows = np.concatenate([np.arange(n), wc.slave_nodes])
        cols = np.concatenate([np.arange(n), wc.master_nodes])
        self.T = sp.csr_matrix((np.ones(len(rows)), (rows, cols)),
                               shape=(n + len(wc.slave_nodes), n))
        self.beta = 1.0
        self.B = None
        self.amg = None
        self.cut_buf = empty(len(mesh_cut.nodes))
        self.red_buf = empty(n)

    # ---------------------------------------------------------- maps
    def set_mach(self, m_inf: float, m_cap: float = 3.0):
        if not 0.0 <= m_inf < 1.0:
            raise ValueError(f"needs 0 <= M_inf < 1, got {m_inf}")
        self.ws.set_mach(m_inf, m_cap)
        self.m_inf = float(m_inf)
        self.beta = float(np.sqrt(1.0 - m_inf ** 2))

    def expand(self, x_red: torch.Tensor, out: torch.Tensor, gamma=None) -> torch.Tensor:
        n = self.n_red
        out[:n].copy_(x_red)
        out[n:].copy_(x_red[self.d_master])
        if gamma is not None:
            out[n:].add_(to_dev(self.wc.gamma_at_nodes(gamma)))
        return out

    def reduce(self, y_cut: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        out.copy_(y_cut[: self.n_red])
        out.index_add_(0, self.d_master, y_cut[self.n_red:])
        return out

    def full_phi(self, phi_red: torch.Tensor, gamma, out: torch.Tensor) -> torch.Tensor:
        self.expand(phi_red, out, gamma)
        ff, vals = farfield_dirichlet(self.mesh, self.wc, self.alpha, gamma, self.u_inf,
                                      self.center, beta=self.beta)
        out[to_dev(ff, torch.int64)] = to_dev(vals)
        return out

    def dirichlet_red(self, gamma) -> np.ndarray:
        ff, vals = farfield_dirichlet(self.mesh, self.wc, self.alpha, gamma, self.u_inf,
                                      self.center, beta=self.beta)
        keep = ff < self.n_red
        return vals[keep]

    def kutta_k(self, x_red: torch.Tensor) -> torch.Tensor:
        """K x: per-station mean of x[upper probe] - x[lower probe]."""
        d = (x_red[self.d_up] - x_red[self.d_lo]) * self.k_w
        out = torch.zeros(self.n_st, dtype=torch.float64, device=self.dev)
        return out.index_add_(0, self.d_te_station, d)

    # ------------------------------------------------------ residual
    def residual(self, phi_red: torch.Tensor, gamma, upwind_c, m_crit, rho_floor,
                 frozen=None, slot: int = 0):
        """(R_red device (0 on Dirichlet rows), F host, state)."""
        ws = self.ws
        s = ws.states[slot]
        self.full_phi(phi_red, gamma, s.phi)
        ws.eval_residual(s, upwind_c, m_crit, rho_floor, frozen)
        R = self.reduce(s.R, empty(self.n_red))
        R.mul_(self.d_free)
        F = to_host(self.kutta_k(s.phi[: self.n_red])) - np.asarray(gamma, dtype=np.float64)
        return R, F, s

    # ------------------------------------------------------- linearise
    def linearize(self, s, phi_red, gamma, upwind_c, m_crit, rho_floor, frozen,
                  precond_matrix: str = "jacobian_spd", rebuild_amg: bool = True,
                  eps: float = 1e-6):
        ws = self.ws
        ws.assemble_coupled(s, upwind_c, m_crit, rho_floor, frozen)
        if rebuild_amg or self.amg is None:
            se = None
            if precond_matrix == "jacobian_spd" and self.m_inf > 0.0:
                upw = ws.upw
                if frozen is not None:
                    se, _ = upw.frozen_sensitivities_dev(s.q2l, s.rho, frozen[0], frozen[1],
                                                         ws.m_inf, upwind_c, m_crit,
                                                         ws.gamma_air, rho_floor)
                else:
                    se, _ = upw.sensitivities_at_dev(s.q2l, s.rho, s.up, ws.m_inf, upwind_c,
                                                     m_crit, ws.gamma_air, rho_floor)
                ws.op.jacobian_spd_dev(s.rt, s.grad, se, s.lim, 1.0 / ws.u_inf ** 2, 0.9,
                                       ws.A_pic)
            else:
                ws.op.matrix_dev(s.rt, ws.A_pic)
            A_cut = csr_to_host(ws.st, ws.A_pic)
            A_red = (self.T.T @ (A_cut @ self.T)).tocsr()
            # Dirichlet rows and columns -> identity
            Dm = sp.diags(self.free.astype(np.float64))
            Did = sp.diags((~self.free).astype(np.float64))
            A_red = (Dm @ A_red @ Dm + Did).tocsr()
            self.amg = AMGHierarchy(A_red, precision="f32" if ws.precision == "fp32" else "f64",
                                    seed=0, max_coarse=1000)
        # B = dR_red / dGamma at the frozen selection of this step
        fz = frozen if frozen is not None else ws.freeze(s, upwind_c, m_crit, rho_floor)
        cols = []
        for j in range(self.n_st):
            e = np.zeros(self.n_st)
            e[j] = eps
            Rp, _, _ = self.residual(phi_red, np.asarray(gamma) + e, upwind_c, m_crit,
                                     rho_floor, fz, slot=1)
            Rp = Rp.clone()
            Rm, _, _ = self.residual(phi_red, np.asarray(gamma) - e, upwind_c, m_crit,
                                     rho_floor, fz, slot=1)
            cols.append(((Rp - Rm) / (2.0 * eps)).unsqueeze(1))
        self.B = torch.cat(cols, dim=1)

    def matvec(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
        xc = self.expand(x, self.cut_buf)
        yc = self.ws.jac.matvec(xc, empty(len(self.cut_buf)))
        self.reduce(yc, y)
        y.mul_(self.d_free)
        y[self.d_dir] = x[self.d_dir]
        y.add_(self.B @ self.kutta_k(x))
        return y

    def psolve(self, r: torch.Tensor, z: torch.Tensor) -> torch.Tensor:
        return self.amg.apply(r.contiguous(), z)


def solve_newton_lifting(mesh_cut, wc, m_inf: float, alpha_deg: float = 1.25,
                         u_inf: float = 1.0, gamma_air: float = 1.4,
                         upwind_c: float = 1.5, m_crit: float = 0.95,
                         m_cap: float = 3.0, rho_floor: float = 0.05,
                         vortex_center=(0.25, 0.0), phi_init=None, gamma_init=None,
                         n_newton_max: int = 40, tol_residual: float = 1e-10,
                         freeze_tol: Optional[float] = 1e-6, freeze_refresh_max: int = 4,
                         gmres_restart: int = 100, gmres_maxiter: int = 10,
                         eta: float = 1e-4, precond_matrix: str = "jacobian_spd",
                         precision: str = "fp64", system: Optional[LiftingNewtonSystem] = None,
                         verbose: bool = False) -> Dict[str, object]:
    """Newton at one Mach level on (phi_red, Gamma).  phi_init: cut-mesh or
    reduced potential (default: the incompressible lifting solution);
    converged when ||R_free||_inf and ||F||_inf are below tol_residual with
    0 limited / floored elements, on the live selection (a frozen
    selection that keeps changing is accepted as "assignment_cycle" after
    freeze_refresh_max refreshes).  Returns phi (cut mesh), phi_red, gamma,
    converged, accept_reason, residual_history, kutta_history, n_newton,
    n_limited, n_floored, n_gmres_total, mach2_max, timings, system."""
    # the system (device set-up of the cut mesh) is built before the clock
    # starts, as the reference builds its workspace before its wall clock
    sysm = system or LiftingNewtonSystem(mesh_cut, wc, alpha_deg, u_inf, gamma_air,
                                         vortex_center, precision)
    t_wall0 = time.perf_counter()
    timings = new_timings()
    n = wc.n_nodes_orig
    with phase(timings, "seed"):
        sysm.set_mach(m_inf, m_cap)
        if phi_init is None:
            from .picard import solve_laplace_lifting

            lap = solve_laplace_lifting(mesh_cut, wc, alpha_deg, u_inf, vortex_center)
            phi_init, gamma_init = lap["phi"], lap["gamma"]
        phi = to_dev(np.asarray(phi_init, dtype=np.float64)[:n])
        gamma = np.atleast_1d(np.asarray(gamma_init if gamma_init is not None
                                         else np.zeros(wc.n_stations),
                                         dtype=np.float64)).copy()
        phi[sysm.d_dir] = to_dev(sysm.dirichlet_red(gamma))
    frozen = None
    n_refresh = 0
    res_hist, kut_hist = [], []
    converged, reason = False, None
    n_gm = 0
    r_best = np.inf
    prev_live = None
    for it in range(n_newton_max):
        with phase(timings, "residual"):
            R, F, s = sysm.residual(phi, gamma, upwind_c, m_crit, rho_floor, frozen)
            r = float(R.abs().max().item())
            f = float(np.abs(F).max())
        res_hist.append(r)
        kut_hist.append(f)
        clamped = s.n_limited + s.n_floored
        if verbose:
            print(f"  lifting-newton {it}: |R|={r:.3e} |F|={f:.3e} gamma={gamma} "
                  f"lim/flr={s.n_limited}/{s.n_floored}" + (" [frozen]" if frozen else ""),
                  flush=True)
        if max(r, f) < tol_residual and clamped == 0:
            if frozen is None:
                converged, reason = True, "tol"
                break
            with phase(timings, "residual"):
                Rl, Fl, sl = sysm.residual(phi, gamma, upwind_c, m_crit, rho_floor, None,
                                           slot=1)
                live = max(float(Rl.abs().max().item()), float(np.abs(Fl).max()))
            if live < tol_residual and sl.n_limited + sl.n_floored == 0:
                converged, reason = True, "tol"
                break
            if n_refresh >= freeze_refresh_max or (prev_live is not None and live > 0.5 * prev_live):
                converged, reason = True, "assignment_cycle"
                break
            prev_live = live
            n_refresh += 1
            with phase(timings, "residual"):
                frozen = sysm.ws.freeze(sl, upwind_c, m_crit, rho_floor)
            continue
        if frozen is None and freeze_tol is not None and max(r, f) < freeze_tol and clamped == 0:
            with phase(timings, "residual"):
                frozen = sysm.ws.freeze(s, upwind_c, m_crit, rho_floor)
                R, F, s = sysm.residual(phi, gamma, upwind_c, m_crit, rho_floor, frozen)
        with phase(timings, "assembly"):
            sysm.linearize(s, phi, gamma, upwind_c, m_crit, rho_floor, frozen,
                           precond_matrix, rebuild_amg=(it % 3 == 0))
        with phase(timings, "linsolve"):
            rhs = -(R + sysm.B @ to_dev(F))
            rhs.mul_(sysm.d_free)
            d, info, n_it = gmres_dev(sysm.matvec, rhs, sysm.psolve, rtol=eta,
                                      restart=gmres_restart, maxiter=gmres_maxiter,
                                      basis="f32" if precision == "fp32" else "f64")
            n_gm += n_it
            dg = to_host(sysm.kutta_k(d)) + F
        merit0 = max(r, f)
        lam = 1.0
        best = None
        with phase(timings, "residual"):
            while True:
                p_t = phi + lam * d
                g_t = gamma + lam * dg
                p_t[sysm.d_dir] = to_dev(sysm.dirichlet_red(g_t))
                Rt, Ft, st = sysm.residual(p_t, g_t, upwind_c, m_crit, rho_floor, frozen, slot=2)
                m_t = max(float(Rt.abs().max().item()), float(np.abs(Ft).max()))
                if np.isfinite(m_t) and (best is None or m_t < best[2]):
                    best = (p_t, g_t, m_t)
                if np.isfinite(m_t) and m_t < merit0:
                    break
                lam *= 0.5
                if lam < 1.0 / 32:
                    break
        if best is None:
            raise RuntimeError("lifting Newton: every trial state was non-finite")
        phi, gamma = best[0], best[1]
        r_best = min(r_best, best[2])
    with phase(timings, "residual"):
        R, F, s = sysm.residual(phi, gamma, upwind_c, m_crit, rho_floor, None)
        phi_cut = to_host(s.phi)
        q2 = to_host(s.q2n)
    from ..physics.isentropic import mach_squared_field

    finalize(timings, time.perf_counter() - t_wall0)
    return {"phi": phi_cut, "phi_red": to_host(phi), "gamma": gamma,
            "cl_kj": 2.0 * float(np.mean(gamma)) / u_inf, "converged": converged,
            "accept_reason": reason, "residual_history": res_hist,
            "kutta_history": kut_hist, "n_newton": len(res_hist),
            "n_limited": s.n_limited, "n_floored": s.n_floored, "n_gmres_total": n_gm,
            "mach2_max": float(np.max(mach_squared_field(q2, m_inf, gamma_air))) if m_inf > 0 else 0.0,
            "timings": timings, "system": sysm}


def solve_newton_lifting_transonic(mesh_cut, wc, m_target: float, alpha_deg: float = 1.25,
                                   m_start: float = 0.60, dm: float = 0.05,
                                   dm_min: float = 0.01, verbose: bool = False,
                                   **kw) -> Dict[str, object]:
    """Upward Mach ramp of solve_newton_lifting (pyfp3d/solve/newton.py:
    1144-1307): each level warm-started from the last converged one, the
    Mach step halved below a failed level down to dm_min.  Returns the last
    level's dict plus `levels` and `target_reached`."""
    from .continuation import mach_schedule

    sched = [float(m) for m in mach_schedule(m_target, m_start=m_start, dm=dm)]
    levels, good, res, i = [], None, None, 0
    sysm = None
    while i < len(sched):
        m = sched[i]
        lkw = dict(kw)
        if good is not None:
            lkw.update(phi_init=good[0], gamma_init=good[1])
        res = solve_newton_lifting(mesh_cut, wc, m, alpha_deg, system=sysm,
                                   verbose=verbose, **lkw)
        sysm = res["system"]
        levels.append({"m_inf": m, "gamma": float(np.mean(res["gamma"])),
                       "converged": res["converged"], "n_newton": res["n_newton"],
                       "accept_reason": res["accept_reason"],
                       "mach_max": float(np.sqrt(res["mach2_max"])),
                       "n_limited": res["n_limited"], "n_floored": res["n_floored"]})
        if verbose:
            print(f"  [lifting ramp] M={m:.4f} conv={res['converged']} "
                  f"gamma={levels[-1]['gamma']:.6f}", flush=True)
        if res["converged"]:
            good = (res["phi"], res["gamma"])
            i += 1
            continue
        prev = m_start if i == 0 else sched[i - 1]
        half = 0.5 * (m - prev)
        if i == 0 or half < dm_min:
            break
        sched.insert(i, prev + half)
    out = dict(res)
    out["levels"] = levels
    out["mach_schedule"] = sched
    out["target_reached"] = bool(good is not None and i == len(sched))
    out.pop("system", None)
    return out

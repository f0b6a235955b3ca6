"""Subsonic lifting solve on a level-set wake with implicit (pressure-
equality) Kutta (pyfp3d/solve/picard_ls.py:297-892,
solve_multivalued_lifting, restated for the vortex far field).

Per outer iteration:
  * assemble the extended system: the plain part on the device
    (MultivaluedOperator.plain_weight_dev + PicardOperator.matrix_dev), the
    special-element / aux-row / Kutta correction C on the host (O(sheet));
  * Dirichlet far field = freestream + Prandtl-Glauert vortex of the
    current (relaxed) circulation on the far-field MAIN DOFs, aux DOFs free;
  * solve with device GMRES on y = [A_plain x_main; 0] + C x, preconditioned
    by a block lower-triangular sweep: one AMG V-cycle on the single-valued
    Picard matrix for the main block, then an exact solve of the small aux
    block C_aa (host LU) against r_aux - C_am z_main;
  * Gamma = mean TE jump (relaxed by omega_gamma into the far-field vortex),
    per-element densities refreshed (relaxed by omega_rho);
  * converged when both the Gamma and the density lag fall below their
    tolerances (after the first outer), as the reference does.
"""

from __future__ import annotations

import time
from typing import Dict, Optional

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

from .. import _native
from ..constraints.dirichlet import freestream_phi, vortex_phi_2d
from ..device import empty, ptr, stream_handle, to_dev, to_host, zeros
from ..physics.isentropic import q2_at_mach
from .amg import DeviceCSR
from .common import amg_from_device, mask_dirichlet
from .linear import gmres_dev
from .timing import finalize, new_timings, phase


class ExtendedSystem:
    """Device operator, right-hand side and preconditioner of one outer
    iteration of the extended (main + aux) system.  Closure "wake_ls" fills
    the sheet terms on the device into their fixed pattern
    (MultivaluedOperator.sheet_fill); the "continuity" weld closure keeps
    the host correction."""

    def __init__(self, mvop, rho_main_dev, special_rho, phi_ext, dir_dofs,
                 dir_vals, closure="wake_ls", te_kutta="pressure",
                 amg=None, damping_theta=None, damp_rows=None, x_prev=None,
                 external_rhs=None):
        op = mvop.op
        st = op.store
        self.mvop = mvop
        self.nm, self.nT = mvop.n_main, mvop.n_total
        nm, nT = self.nm, self.nT
        dev = st.device
        fm = np.ones(nT, dtype=np.uint8)
        fm[dir_dofs] = 0
        if np.any(np.asarray(dir_dofs) >= nm):
            raise ValueError("Dirichlet data on aux DOFs is not supported")
        self.free = fm.astype(bool)
        self.fm_main = to_dev(fm[:nm], torch.uint8)
        xD = np.zeros(nT)
        xD[dir_dofs] = dir_vals
        # plain part (special elements weighted 0) and the sheet terms
        self.vals = empty(op.nnz)
        op.matrix_dev(mvop.plain_weight_dev(rho_main_dev), self.vals)
        device_sheet = closure == "wake_ls"
        if device_sheet:
            S = mvop.sheet_fill(special_rho, te_kutta, phi_ext)
            C = None
        else:
            C = mvop.correction(special_rho, closure, te_kutta, phi_ext)
        d_damp = None
        if damping_theta is not None and damp_rows is not None:
            # theta * diag(E) on the damped free main rows (supersonic-zone
            # damping, pyfp3d/solve/picard_ls.py:626-640), with
            # theta diag x_prev on the right-hand side
            ip, ix = st.indptr_host, st.indices_host
            if not hasattr(mvop, "_diag_pos"):
                rows_ = np.repeat(np.arange(nm), np.diff(ip))
                mvop._diag_pos = np.flatnonzero(ix == rows_)
                mvop._d_diag_pos = to_dev(mvop._diag_pos.astype(np.int32), torch.int32)
            if device_sheet:
                if not hasattr(S, "_main_diag"):
                    ds = S.diagonal_slots(nm)
                    S._main_diag = (to_dev(np.flatnonzero(ds >= 0), torch.int64),
                                    to_dev(ds[ds >= 0], torch.int64))
                cd = torch.zeros(nm, dtype=torch.float64, device=dev)
                cd[S._main_diag[0]] = S.data[S._main_diag[1]]
                dA = to_host(self.vals[mvop._d_diag_pos.long()] + cd)
            else:
                dA = to_host(self.vals)[mvop._diag_pos] + C.diagonal()[:nm]
            d = np.zeros(nT)
            d[:nm] = damping_theta * dA
            d[~damp_rows] = 0.0
            d[dir_dofs] = 0.0
            np.clip(d, 0.0, None, out=d)
            d_damp = d
        # b = -E x_D on free rows, x_D on Dirichlet rows (D x_D = 0: the
        # damping vanishes on the Dirichlet rows, x_D off them)
        y = empty(nT)
        xDd = to_dev(xD)
        _native.call("fp3d_csr_spmv", nm, op.nnz, ptr(st.indptr), ptr(st.indices),
                     ptr(self.vals), ptr(xDd[:nm]), ptr(y[:nm]), stream_handle())
        y[nm:].zero_()
        if device_sheet:
            S.add_matvec(xDd, y)
            b = -to_host(y)
        else:
            b = -(to_host(y) + C @ xD)
        if external_rhs is not None:
            b += external_rhs
        if d_damp is not None:
            # (E + D) x = b + D x_prev
            b += d_damp * np.asarray(x_prev)
        b[dir_dofs] = dir_vals
        self.b = to_dev(b)
        # Dirichlet embedding: identity rows, zero columns; the damping on
        # the plain part's diagonal (free main rows only)
        mask_dirichlet(st, self.fm_main, self.vals)
        if d_damp is not None:
            _native.call("fp3d_csr_diag_shift", nm, ptr(mvop._d_diag_pos),
                         ptr(to_dev(d_damp[:nm])), 1.0, ptr(self.fm_main),
                         ptr(self.vals), stream_handle())
        if device_sheet:
            S.fill(*mvop.special_weights_dev(special_rho),
                   kv=mvop.kutta_values(phi_ext) if te_kutta == "pressure" else None,
                   fmask=to_dev(fm, torch.uint8))
            self.S = S
            self.C = None
            Caa = S.block_csc(nm, nm) if nT > nm else None
        else:
            Dm = sp.diags(fm.astype(np.float64))
            Cm = (Dm @ C @ Dm).tocsr()
            Cm.eliminate_zeros()
            self.S = None
            self.C = DeviceCSR(Cm)
            self.C_host = Cm
            self.C_am = DeviceCSR(Cm[nm:, :nm].tocsr()) if nT > nm else None
            Caa = Cm[nm:, nm:].tocsc() if nT > nm else None
        # preconditioner: AMG of the single-valued Picard matrix (all
        # elements, main DOFs) + exact aux-block solve
        if amg is None:
            A_sv = empty(op.nnz)
            w = rho_main_dev if rho_main_dev is not None else \
                torch.ones(st.n_tets, dtype=torch.float64, device=dev)
            op.matrix_dev(w, A_sv)
            mask_dirichlet(st, self.fm_main, A_sv)
            amg = amg_from_device(st, A_sv, "f64")
        self.amg = amg
        self.aux_lu = spla.splu(Caa) if nT > nm else None
        self._t_aux = empty(max(nT - nm, 1))

    def matvec(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
        st = self.mvop.op.store
        nm = self.nm
        _native.call("fp3d_csr_spmv", nm, self.mvop.op.nnz, ptr(st.indptr),
                     ptr(st.indices), ptr(self.vals), ptr(x[:nm]), ptr(y[:nm]),
                     stream_handle())
        if self.nT > nm:
            y[nm:].zero_()
        if self.S is not None:
            self.S.add_matvec(x, y)
            return y
        C = self.C
        _native.call("fp3d_amg_prolong_add", C.n, C.nnz, ptr(C.indptr),
                     ptr(C.indices), ptr(C.data), 0, ptr(x), ptr(y),
                     stream_handle())
        return y

    def psolve(self, r: torch.Tensor, z: torch.Tensor) -> torch.Tensor:
        nm = self.nm
        self.amg.apply(r[:nm].contiguous(), z[:nm])
        if self.nT > nm:
            t = self._t_aux
            if self.S is not None:
                # aux rows of C applied to [z_main; 0]: the (aux, main) block
                z[nm:].zero_()
                self.S.rows_matvec(nm, z, t)
            else:
                C = self.C_am
                _native.call("fp3d_csr_spmv", C.n, C.nnz, ptr(C.indptr),
                             ptr(C.indices), ptr(C.data), ptr(z[:nm]), ptr(t),
                             stream_handle())
            # one read-back of r_aux - t, the exact aux solve on the host,
            # and the solution back through a pinned staging buffer (reused
            # only after the next read-back has synchronised the stream)
            t.mul_(-1.0).add_(r[nm:])
            rhs = to_host(t)
            stage = getattr(self, "_aux_stage", None)
            if stage is None:
                stage = self._aux_stage = torch.empty(len(rhs), dtype=torch.float64,
                                                      pin_memory=True)
            stage.numpy()[:] = self.aux_lu.solve(rhs)
            z[nm:].copy_(stage, non_blocking=True)
        return z


def _freestream(points, alpha_deg, u_inf, lift_axis=1):
    """U (x cos(alpha) + x_l sin(alpha)), x_l the coordinate along
    `lift_axis` (1: the reference's x-y convention, 2: x-z)."""
    if lift_axis == 1:
        return freestream_phi(points, alpha_deg, u_inf)
    a = np.deg2rad(alpha_deg)
    return u_inf * (points[:, 0] * np.cos(a) + points[:, lift_axis] * np.sin(a))


def _farfield_main(nodes, farfield_faces, alpha_deg, gamma, u_inf, center, beta,
                   farfield="vortex", lift_axis=1):
    """(far-field main DOFs, Dirichlet values): freestream plus, for
    farfield='vortex', the Prandtl-Glauert 2-D vortex of circulation gamma
    (x-y plane only)."""
    ff = np.unique(np.asarray(farfield_faces))
    xy = nodes[ff]
    vals = _freestream(xy, alpha_deg, u_inf, lift_axis)
    if farfield == "vortex":
        if lift_axis != 1:
            raise ValueError("the vortex far field is the x-y plane's (lift_axis=1)")
        vals = vals + vortex_phi_2d(xy, gamma, center, beta=beta)
    elif farfield != "freestream":
        raise NotImplementedError(f"farfield={farfield!r}: 'vortex' or 'freestream'")
    return ff, vals


def solve_multivalued_lifting(mvop, mesh, m_inf: float, alpha_deg: float = 2.0,
                              u_inf: float = 1.0, gamma_air: float = 1.4,
                              vortex_center=(0.25, 0.0),
                              omega_gamma: float = 0.5, omega_rho: float = 0.5,
                              tol_gamma: float = 1e-6, tol_rho: float = 1e-6,
                              n_outer_max: int = 80, gamma_init: float = 0.0,
                              gamma_farfield_fixed: Optional[float] = None,
                              te_kutta: str = "pressure",
                              farfield: str = "vortex", lift_axis: int = 1,
                              upwind_c: float = 0.0, m_crit: float = 0.95,
                              m_cap: float = 3.0, rho_floor: float = 0.05,
                              damping_theta: Optional[float] = None,
                              tol_residual: Optional[float] = None,
                              phi_init: Optional[np.ndarray] = None,
                              linear_rtol: float = 1e-10,
                              gmres_restart: int = 60, gmres_maxiter: int = 50,
                              amg_rebuild_every: int = 5,
                              external_rhs: Optional[np.ndarray] = None,
                              verbose: bool = False) -> Dict[str, object]:
    """`mesh` needs .nodes and .boundary_faces["farfield"]; farfield
    'freestream' holds the far field at the freestream (no circulation
    feedback), in the x-`lift_axis` plane; external_rhs (n_total) is a
    boundary load on the right-hand side (a Neumann outflow).  upwind_c > 0
    (with m_inf > 0) runs the per-side artificial density with side-
    restricted walks, and damping_theta the theta*diag damping on the main
    rows of nu > 0 elements (the reference's supersonic-zone scope).
    Returns the reference's result keys: phi_ext, phi, gamma, cl_kj,
    te_jump, n_outer, converged, gamma_history, drho_history,
    residual_history, residual_norm, timings, mach2_max, nu_max,
    n_nu_active, n_limited, n_floored, n_gmres_total, n_gmres_stalled."""
    if farfield not in ("vortex", "freestream"):
        raise NotImplementedError(f"farfield={farfield!r}: 'vortex' or 'freestream'")
    if not 0.0 <= m_inf < 1.0:
        raise ValueError(f"needs 0 <= M_inf < 1, got {m_inf}")
    if len(mvop.cm.te_nodes) == 0:
        raise ValueError("the wake level set matched 0 TE nodes on this mesh")
    t_wall0 = time.perf_counter()
    timings = new_timings()
    beta = float(np.sqrt(1.0 - m_inf ** 2))
    op = mvop.op
    st = op.store
    use_upwind = upwind_c > 0.0 and m_inf > 0.0
    nodes = np.asarray(mesh.nodes, dtype=np.float64)
    ffaces = mesh.boundary_faces["farfield"]
    phi_ext = np.zeros(mvop.n_total)
    if phi_init is not None:
        phi_ext[:] = phi_init
    else:
        phi_ext[: mvop.n_main] = _freestream(nodes, alpha_deg, u_inf, lift_axis)
        phi_ext[mvop.cut_aux] = phi_ext[mvop.cut_nodes]
    gamma = float(gamma_init)
    gamma_history, drho_history, residual_history = [gamma], [], []
    w_dev = None           # device element weights (plain part / AMG)
    special = None         # host densities of the special copies
    rt_pair = None         # (upper, lower) device densities (upwind path)
    # the first outer damps every main row; later outers only the rows of
    # nu > 0 elements (pyfp3d/solve/picard_ls.py:562-566)
    damp_rows = None
    if damping_theta is not None:
        damp_rows = np.zeros(mvop.n_total, dtype=bool)
        damp_rows[: mvop.n_main] = True
    amg = None
    n_gmres_total = n_gmres_stalled = 0
    converged = False
    residual = np.inf
    dgamma = drho = np.inf
    nt = st.n_tets
    grad = empty((nt, 3))
    q2n = empty(nt)
    q2l = empty(nt)
    rho_new = empty(nt)
    lim = empty(nt, torch.uint8)
    own_plus = torch.tensor(mvop.in_upper, device=st.device)
    for outer in range(n_outer_max):
        with phase(timings, "assembly"):
            g_ff = gamma if gamma_farfield_fixed is None else gamma_farfield_fixed
            ff, vals = _farfield_main(nodes, ffaces, alpha_deg, g_ff, u_inf,
                                      vortex_center, beta, farfield, lift_axis)
            reuse = amg if (amg is not None and outer % amg_rebuild_every) else None
            sysm = ExtendedSystem(mvop, w_dev, special, phi_ext, ff, vals,
                                  "wake_ls", te_kutta, amg=reuse,
                                  damping_theta=damping_theta,
                                  damp_rows=damp_rows, x_prev=phi_ext,
                                  external_rhs=external_rhs)
            amg = sysm.amg
            x_prev = to_dev(np.where(sysm.free, phi_ext, 0.0))
            x_prev[to_dev(ff, torch.int64)] = to_dev(vals)
            r = empty(mvop.n_total)
            sysm.matvec(x_prev, r)
            rr = to_host(sysm.b - r)
            residual = float(np.abs(rr[sysm.free]).max())
            residual_history.append(residual)
        with phase(timings, "linsolve"):
            x, info, n_it = gmres_dev(sysm.matvec, sysm.b, sysm.psolve,
                                      x0=x_prev, rtol=linear_rtol,
                                      restart=gmres_restart, maxiter=gmres_maxiter)
            n_gmres_total += n_it
            n_gmres_stalled += int(info != 0)
            phi_ext = to_host(x)
        with phase(timings, "kutta"):
            g_new = float(np.mean(mvop.te_jump(phi_ext)))
            dgamma = abs(g_new - gamma)
            gamma = (1.0 - omega_gamma) * gamma + omega_gamma * g_new
            gamma_history.append(gamma)
        with phase(timings, "residual"):
            if use_upwind:
                ru, rl = mvop.side_rho_tilde_dev(phi_ext, m_inf, upwind_c, m_crit,
                                                 gamma_air, u_inf, m_cap, rho_floor)
                if damping_theta is not None:
                    act = mvop.nu_active_elements()
                    damp_rows = np.zeros(mvop.n_total, dtype=bool)
                    damp_rows[np.unique(mvop.el[act])] = True
                if rt_pair is None:
                    drho = 0.0
                    rt_pair = (ru, rl)
                else:
                    new_own = torch.where(own_plus, ru, rl)
                    old_own = torch.where(own_plus, *rt_pair)
                    drho = float((new_own - old_own).abs().max().item())
                    rt_pair = tuple(o + omega_rho * (n_ - o)
                                    for o, n_ in zip(rt_pair, (ru, rl)))
                w_dev, special = mvop.side_weights(*rt_pair)
            else:
                phid = to_dev(phi_ext[: mvop.n_main])
                op.state_dev(phid, u_inf ** 2, m_inf, gamma_air, np.inf, grad, q2n,
                             q2l, rho_new, lim)
                sp_new = mvop.special_densities(phi_ext, m_inf, gamma_air, u_inf)
                if w_dev is None:
                    drho = 0.0
                    w_dev = rho_new.clone()
                    special = sp_new
                else:
                    d_plain = float(((rho_new - w_dev) * mvop.d_plain_mask).abs().max().item())
                    d_sp = max([float(np.abs(a_ - b_).max()) if len(a_) else 0.0
                                for a_, b_ in zip(sp_new, special)] + [0.0])
                    drho = max(d_plain, d_sp)
                    w_dev = w_dev + omega_rho * (rho_new - w_dev)
                    special = tuple(o + omega_rho * (n_ - o) for o, n_ in zip(special, sp_new))
            drho_history.append(drho)
        if verbose:
            print(f"  ls-picard {outer:3d}: |R|={residual:.3e} gamma={gamma:.6f} "
                  f"dgamma={dgamma:.2e} drho={drho:.2e} gmres={n_it} info={info}"
                  + (f" nu_max={mvop.nu_max:.3f}" if use_upwind else ""), flush=True)
        res_ok = tol_residual is not None and residual < tol_residual
        if outer > 0 and ((dgamma < tol_gamma and drho < tol_rho) or res_ok):
            converged = True
            break
    with phase(timings, "residual"):
        mach2_max = mvop.own_side_mach2_max(phi_ext, m_inf, gamma_air, u_inf)
    finalize(timings, time.perf_counter() - t_wall0)
    return {
        "phi_ext": phi_ext, "phi": mvop.main_potential(phi_ext),
        "gamma": gamma, "cl_kj": 2.0 * gamma / u_inf,
        "te_jump": mvop.te_jump(phi_ext), "n_outer": len(drho_history),
        "converged": converged, "gamma_history": gamma_history,
        "drho_history": drho_history, "residual_history": residual_history,
        "residual_norm": residual, "timings": timings,
        "mach2_max": mach2_max, "farfield": farfield,
        "nu_max": mvop.nu_max if use_upwind else 0.0,
        "n_nu_active": mvop.n_nu_active if use_upwind else 0,
        "n_limited": mvop.n_limited if use_upwind else 0,
        "n_floored": mvop.n_floored if use_upwind else 0,
        "n_gmres_total": n_gmres_total, "n_gmres_stalled": n_gmres_stalled,
    }


def solve_multivalued_transonic(mvop, mesh, m_target: float,
                                alpha_deg: float = 1.25, u_inf: float = 1.0,
                                gamma_air: float = 1.4, m_start: float = 0.60,
                                dm: float = 0.05, upwind_c: float = 1.5,
                                m_crit: float = 0.95, damping_theta: float = 0.2,
                                omega_rho: float = 0.5, n_outer_seed: int = 120,
                                n_outer_level: int = 400, tol_rho: float = 1e-6,
                                tol_gamma: float = 1e-6,
                                tol_residual: Optional[float] = 1e-7,
                                farfield_lag: str = "live",
                                n_ff_polish: int = 0, n_outer_polish: int = 200,
                                **kwargs) -> Dict[str, object]:
    """Mach continuation of the level-set Picard (pyfp3d/solve/picard_ls.py:
    895-1044): each level is a warm-started solve_multivalued_lifting from
    the previous level's full extended state and circulation;
    farfield_lag="level" freezes the far-field vortex at the previous
    level's circulation (level 0 keeps the live vortex), and n_ff_polish
    rounds at the target Mach then re-converge with the vortex set to the
    newest circulation, n_outer_polish outers each (one step of the
    far-field outer map per round).  Returns the final level's dict plus
    `levels` (polish rounds appended at the target Mach) and
    `mach_schedule`."""
    from .continuation import mach_schedule
    from .timing import sum_timings

    if farfield_lag not in ("live", "level"):
        raise ValueError(f"farfield_lag={farfield_lag!r} unknown")
    schedule = mach_schedule(m_target, m_start=m_start, dm=dm)
    phi_ext, gamma, levels, res = None, 0.0, [], None
    t0 = time.perf_counter()
    lagged = farfield_lag == "level" and kwargs.get("farfield", "vortex") == "vortex"
    runs = [(m, n_outer_seed if k == 0 else n_outer_level, lagged and k > 0)
            for k, m in enumerate(schedule)]
    if lagged:
        runs += [(schedule[-1], n_outer_polish, True)] * int(n_ff_polish)
    for m, budget, fixed in runs:
        tl = time.perf_counter()
        res = solve_multivalued_lifting(
            mvop, mesh, m, alpha_deg=alpha_deg, u_inf=u_inf, gamma_air=gamma_air,
            upwind_c=upwind_c, m_crit=m_crit, damping_theta=damping_theta,
            omega_rho=omega_rho, n_outer_max=budget,
            tol_rho=tol_rho, tol_gamma=tol_gamma, tol_residual=tol_residual,
            gamma_init=gamma, phi_init=phi_ext,
            gamma_farfield_fixed=gamma if fixed else None,
            **kwargs)
        phi_ext, gamma = res["phi_ext"], res["gamma"]
        levels.append({"m_inf": float(m), "gamma": float(gamma),
                       "cl_kj": float(res["cl_kj"]),
                       "mach_max": float(np.sqrt(res["mach2_max"])),
                       "n_outer": int(res["n_outer"]),
                       "converged": bool(res["converged"]),
                       "n_limited": int(res["n_limited"]),
                       "n_floored": int(res["n_floored"]),
                       "residual_norm": float(res["residual_norm"]),
                       "wall_s": time.perf_counter() - tl,
                       "timings": res["timings"],
                       "n_lin_iters": int(res["n_gmres_total"])})
    res = dict(res)
    res["levels"] = levels
    res["mach_schedule"] = [float(m) for m in schedule]
    res["timings_total"] = sum_timings([lv["timings"] for lv in levels])
    res["wall_s"] = time.perf_counter() - t0
    return res

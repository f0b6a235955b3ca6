"""Non-lifting subsonic Picard solve on the device
(pyfp3d/solve/picard.py:271-417 solve_subsonic).

Per iteration: coloured assembly of A(rho) -> SA-AMG (host set-up) ->
device PCG on the free dofs (Dirichlet rows/cols embedded as identity, so
full-length vectors with zero Dirichlet entries see exactly A[free][:,free])
-> element velocity sweep -> isentropic density -> coloured residual.
Converges on ||rho(phi_new) - rho||_inf < tol_rho, as the reference.
"""

from __future__ import annotations

from typing import Dict, Optional

import numpy as np
import torch

from .. import _native
from ..device import empty, ptr, stream_handle, to_dev, to_host, zeros
from ..kernels.jacobian import PicardOperator
from .common import (Reducer, amg_from_device, dirichlet_setup, mask_dirichlet,
                     masked)
from .linear import axpby, pcg_dev


def solve_subsonic(nodes: np.ndarray, elements: np.ndarray,
                   dirichlet_nodes: np.ndarray, dirichlet_values: np.ndarray,
                   m_inf: float, gamma_air: float = 1.4, u_inf: float = 1.0,
                   body_source_rhs: Optional[np.ndarray] = None,
                   phi_init: Optional[np.ndarray] = None, omega: float = 1.0,
                   n_picard_max: int = 40, tol_rho: float = 1e-10,
                   rtol: float = 1e-12, maxiter: int = 3000,
                   precision: str = "f64") -> Dict[str, object]:
    """Same arguments and result keys as the reference (phi, n_picard,
    converged, residual_history, drho_history, rho, mach2_max, n_cg_total);
    `precision` selects float64 or float32 AMG matrices."""
    op = PicardOperator(nodes, elements)
    st = op.store
    n = op.n_nodes
    nt = op.n_tets
    fm, dn, dv, phi_d = dirichlet_setup(n, dirichlet_nodes, dirichlet_values,
                                        st.device)
    b_wall = None
    if body_source_rhs is not None:
        bw = np.asarray(body_source_rhs, dtype=np.float64)
        if bw.shape != (n,):
            raise ValueError(f"body_source_rhs must be ({n},), got {bw.shape}")
        b_wall = to_dev(bw)
    red = Reducer()
    u2 = u_inf ** 2
    grad = empty((nt, 3))
    q2n = empty(nt)
    q2l = empty(nt)
    rho_new = empty(nt)
    lim = empty(nt, torch.uint8)
    A = empty(st.nnz)
    Ad = empty(n)
    b = empty(n)
    Rm = empty(n)
    diff = empty(nt)
    if phi_init is not None:
        phi = to_dev(phi_init, torch.float64).clone()
        phi.index_copy_(0, dn, dv)
        op.state_dev(phi, u2, m_inf, gamma_air, np.inf, grad, q2n, q2l, rho_new, lim)
        rho = rho_new.clone()
    else:
        phi = phi_d.clone()
        rho = torch.ones(nt, dtype=torch.float64, device=st.device)
    residual_history, drho_history = [], []
    n_cg_total = 0
    converged = False
    for _ in range(n_picard_max):
        op.matrix_dev(rho, A)
        # b_free = b_wall - A[:, dir] vals (rows of dir dofs zeroed)
        _native.call("fp3d_csr_spmv", n, st.nnz, ptr(st.indptr), ptr(st.indices),
                     ptr(A), ptr(phi_d), ptr(Ad), stream_handle())
        if b_wall is not None:
            b.copy_(b_wall)
        else:
            b.zero_()
        axpby(-1.0, Ad, 1.0, b)
        masked(fm, b, b)
        mask_dirichlet(st, fm, A)
        ml = amg_from_device(st, A, precision)
        Amv = lambda x, y: _native.call(  # noqa: E731
            "fp3d_csr_spmv", n, st.nnz, ptr(st.indptr), ptr(st.indices), ptr(A),
            ptr(x), ptr(y), stream_handle())
        x, info, n_it = pcg_dev(Amv, b, ml.apply, rtol=rtol, maxiter=maxiter)
        if info != 0:
            raise RuntimeError(f"CG did not converge (info={info})")
        n_cg_total += n_it
        phi_new = x
        axpby(1.0, phi_d, 1.0, phi_new)        # x has zero Dirichlet entries
        if omega == 1.0:
            phi = phi_new
        else:
            axpby(omega, phi_new, 1.0 - omega, phi)
        op.state_dev(phi, u2, m_inf, gamma_air, np.inf, grad, q2n, q2l, rho_new,
                     lim)
        R = op.residual_dev(rho_new, grad)
        if b_wall is not None:
            axpby(-1.0, b_wall, 1.0, R)
        masked(fm, R, Rm)
        residual_history.append(red.norms(Rm)[1])
        diff.copy_(rho_new)
        axpby(-1.0, rho, 1.0, diff)
        drho = red.norms(diff)[1]
        drho_history.append(drho)
        rho = rho_new.clone()
        if drho < tol_rho:
            converged = True
            break
    op.state_dev(phi, u2, m_inf, gamma_air, np.inf, grad, q2n, q2l, rho_new, lim)
    m2 = empty(nt)
    _native.call("fp3d_mach_squared_field", nt, ptr(q2n), float(m_inf),
                 float(gamma_air), ptr(m2), stream_handle())
    mach2_max = red.norms(m2)[1]
    return {
        "phi": to_host(phi),
        "n_picard": len(residual_history),
        "converged": converged,
        "residual_history": residual_history,
        "drho_history": drho_history,
        "rho": to_host(st.to_original(rho)),
        "mach2_max": mach2_max,
        "n_cg_total": n_cg_total,
    }


def solve_laplace_lifting(mesh_cut, wc, alpha_deg: float = 0.0, u_inf: float = 1.0,
                          vortex_center=(0.25, 0.0), gamma_fixed=None,
                          omega_gamma: float = 0.9, tol_gamma: float = 1e-8,
                          max_kutta_updates: int = 30, rtol: float = 1e-10,
                          maxiter: int = 2000, kutta_estimator: str = "probe"
                          ) -> Dict[str, object]:
    """Incompressible lifting solve on a wake-cut mesh
    (pyfp3d/solve/picard.py:88-268): the cut mesh's stiffness matrix is
    assembled once on the device and reduced by the wake constraint; each
    Kutta update is one device PCG solve (SA-AMG V-cycle) of the reduced
    free system with the far-field vortex of the current Gamma; the first
    update is relaxed by omega_gamma, later ones take the per-station
    secant step to the affine map's fixed point.  Converged when the Kutta
    mismatch ||target - Gamma||_inf < tol_gamma.  Returns phi (cut mesh),
    gamma, n_kutta_updates, gamma_history, residual_norm, n_cg_total,
    kutta_converged."""
    from ..constraints.dirichlet import farfield_dirichlet
    from ..constraints.wake import WakeConstraint, kutta_targets
    from .amg import AMGHierarchy, DeviceCSR

    if kutta_estimator != "probe":
        raise NotImplementedError("only the probe Kutta estimator is built")
    op = PicardOperator(mesh_cut.nodes, mesh_cut.elements)
    A = op.assemble_matrix()
    con = WakeConstraint(A, wc)
    n_red = con.n_reduced
    b_zero = np.zeros(A.shape[0])
    dir_nodes, _ = farfield_dirichlet(mesh_cut, wc, alpha_deg, np.zeros(wc.n_stations),
                                      u_inf, vortex_center)
    dir_red, _ = con.to_reduced_dirichlet(dir_nodes, np.zeros(len(dir_nodes)))
    is_dir = np.zeros(n_red, dtype=bool)
    is_dir[dir_red] = True
    free = np.flatnonzero(~is_dir)
    Ar = con.A_reduced.tocsr()
    A_free = Ar[free][:, free].tocsr()
    A_cpl = Ar[free][:, dir_red].tocsr()
    ml = AMGHierarchy(A_free, precision="f64", seed=0)
    D = DeviceCSR(A_free)

    def solve_for(gamma):
        nd, vd = farfield_dirichlet(mesh_cut, wc, alpha_deg, gamma, u_inf, vortex_center)
        _, vr = con.to_reduced_dirichlet(nd, vd)
        b = con.reduced_rhs(b_zero, gamma)[free] - A_cpl @ vr
        x, info, it = pcg_dev(D.matvec, to_dev(b), ml.apply, rtol=rtol, maxiter=maxiter)
        if info != 0:
            raise RuntimeError(f"CG did not converge (info={info})")
        xh = to_host(x)
        phi = np.empty(n_red)
        phi[free] = xh
        phi[dir_red] = vr
        return phi, float(np.abs(b - A_free @ xh).max()) if len(xh) else 0.0, it

    if gamma_fixed is not None:
        gamma = np.broadcast_to(np.atleast_1d(np.asarray(gamma_fixed, dtype=np.float64)),
                                (wc.n_stations,)).copy()
        phi, res, it = solve_for(gamma)
        return {"phi": con.expand(phi, gamma), "gamma": gamma, "n_kutta_updates": 0,
                "gamma_history": [gamma.copy()], "residual_norm": res,
                "n_cg_total": it, "kutta_converged": True}
    gamma = np.zeros(wc.n_stations)
    history = [gamma.copy()]
    n_cg = 0
    converged = False
    prev = None
    phi = None
    res = np.inf
    for _ in range(max_kutta_updates):
        phi, res, it = solve_for(gamma)
        n_cg += it
        target = kutta_targets(con.expand(phi, gamma), wc)
        if float(np.abs(target - gamma).max()) < tol_gamma:
            converged = True
            break
        relaxed = gamma + omega_gamma * (target - gamma)
        if prev is None:
            new = relaxed
        else:
            dg = gamma - prev[0]
            slope = np.where(np.abs(dg) > 1e-14,
                             (target - prev[1]) / np.where(dg == 0, 1.0, dg), 0.0)
            den = 1.0 - slope
            new = np.where(np.abs(den) > 1e-3,
                           (target - slope * gamma) / np.where(den == 0, 1.0, den), relaxed)
        prev = (gamma, target)
        gamma = new
        history.append(gamma.copy())
    if not converged and phi is not None:
        phi, res, it = solve_for(gamma)
        n_cg += it
    return {"phi": con.expand(phi, gamma), "gamma": gamma,
            "n_kutta_updates": len(history) - 1, "gamma_history": history,
            "residual_norm": res, "n_cg_total": n_cg, "kutta_converged": converged}

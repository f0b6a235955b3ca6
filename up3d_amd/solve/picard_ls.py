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
    iteration of the extended (main + aux) system."""

    def __init__(self, mvop, rho_main_dev, special_rho, phi_ext, dir_dofs,
                 dir_vals, closure="wake_ls", te_kutta="pressure",
                 amg=None):
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
        # plain part (special elements weighted 0) and host correction
        self.vals = empty(op.nnz)
        op.matrix_dev(mvop.plain_weight_dev(rho_main_dev), self.vals)
        C = mvop.correction(special_rho, closure, te_kutta, phi_ext)
        # b = -E x_D on free rows, x_D on Dirichlet rows
        y = empty(nT)
        xDd = to_dev(xD)
        _native.call("fp3d_csr_spmv", nm, op.nnz, ptr(st.indptr), ptr(st.indices),
                     ptr(self.vals), ptr(xDd[:nm]), ptr(y[:nm]), stream_handle())
        y[nm:].zero_()
        yh = to_host(y) + C @ xD
        b = -yh
        b[dir_dofs] = dir_vals
        self.b = to_dev(b)
        # Dirichlet embedding: identity rows, zero columns
        mask_dirichlet(st, self.fm_main, self.vals)
        Dm = sp.diags(fm.astype(np.float64))
        Cm = (Dm @ C @ Dm).tocsr()
        Cm.eliminate_zeros()
        self.C = DeviceCSR(Cm)
        self.C_host = Cm
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
        Caa = Cm[nm:, nm:].tocsc()
        self.aux_lu = spla.splu(Caa) if self.nT > nm else None
        self.C_am = DeviceCSR(Cm[nm:, :nm].tocsr()) if self.nT > nm else None
        self._t_aux = empty(max(nT - nm, 1))

    def matvec(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
        st = self.mvop.op.store
        nm = self.nm
        _native.call("fp3d_csr_spmv", nm, self.mvop.op.nnz, ptr(st.indptr),
                     ptr(st.indices), ptr(self.vals), ptr(x[:nm]), ptr(y[:nm]),
                     stream_handle())
        if self.nT > nm:
            y[nm:].zero_()
        C = self.C
        _native.call("fp3d_amg_prolong_add", C.n, C.nnz, ptr(C.indptr),
                     ptr(C.indices), ptr(C.data), 0, ptr(x), ptr(y),
                     stream_handle())
        return y

    def psolve(self, r: torch.Tensor, z: torch.Tensor) -> torch.Tensor:
        nm = self.nm
        self.amg.apply(r[:nm].contiguous(), z[:nm])
        if self.nT > nm:
            C = self.C_am
            t = self._t_aux
            _native.call("fp3d_csr_spmv", C.n, C.nnz, ptr(C.indptr),
                         ptr(C.indices), ptr(C.data), ptr(z[:nm]), ptr(t),
                         stream_handle())
            rhs = to_host(r[nm:]) - to_host(t)
            z[nm:].copy_(to_dev(self.aux_lu.solve(rhs)))
        return z


def _farfield_main(nodes, farfield_faces, alpha_deg, gamma, u_inf, center, beta):
    ff = np.unique(np.asarray(farfield_faces))
    xy = nodes[ff]
    return ff, freestream_phi(xy, alpha_deg, u_inf) + vortex_phi_2d(
        xy, gamma, center, beta=beta)


def solve_multivalued_lifting(mvop, mesh, m_inf: float, alpha_deg: float = 2.0,
                              u_inf: float = 1.0, gamma_air: float = 1.4,
                              vortex_center=(0.25, 0.0),
                              omega_gamma: float = 0.5, omega_rho: float = 0.5,
                              tol_gamma: float = 1e-6, tol_rho: float = 1e-6,
                              n_outer_max: int = 80, gamma_init: float = 0.0,
                              te_kutta: str = "pressure",
                              farfield: str = "vortex",
                              phi_init: Optional[np.ndarray] = None,
                              linear_rtol: float = 1e-10,
                              gmres_restart: int = 60, gmres_maxiter: int = 50,
                              amg_rebuild_every: int = 5,
                              verbose: bool = False) -> Dict[str, object]:
    """`mesh` needs .nodes and .boundary_faces["farfield"].  Returns the
    reference's result keys: phi_ext, phi, gamma, cl_kj, te_jump, n_outer,
    converged, gamma_history, drho_history, residual_history,
    residual_norm, timings, mach2_max, n_gmres_total, n_gmres_stalled."""
    if farfield != "vortex":
        raise NotImplementedError("only the vortex far field is built")
    if not 0.0 <= m_inf < 1.0:
        raise ValueError(f"needs 0 <= M_inf < 1, got {m_inf}")
    if len(mvop.cm.te_nodes) == 0:
        raise ValueError("the wake level set matched 0 TE nodes on this mesh")
    t_wall0 = time.perf_counter()
    timings = new_timings()
    beta = float(np.sqrt(1.0 - m_inf ** 2))
    op = mvop.op
    st = op.store
    nodes = np.asarray(mesh.nodes, dtype=np.float64)
    ffaces = mesh.boundary_faces["farfield"]
    phi_ext = np.zeros(mvop.n_total)
    if phi_init is not None:
        phi_ext[:] = phi_init
    else:
        phi_ext[: mvop.n_main] = freestream_phi(nodes, alpha_deg, u_inf)
        phi_ext[mvop.cut_aux] = phi_ext[mvop.cut_nodes]
    gamma = float(gamma_init)
    gamma_history, drho_history, residual_history = [gamma], [], []
    rho_main = None        # device (n_tets,) densities of the main field
    special = None         # host densities of the special copies
    amg = None
    n_gmres_total = n_gmres_stalled = 0
    converged = False
    residual = np.inf
    q2cap = np.inf
    grad = empty((st.n_tets, 3))
    q2n = empty(st.n_tets)
    q2l = empty(st.n_tets)
    rho_new = empty(st.n_tets)
    lim = empty(st.n_tets, torch.uint8)
    for outer in range(n_outer_max):
        with phase(timings, "assembly"):
            ff, vals = _farfield_main(nodes, ffaces, alpha_deg, gamma, u_inf,
                                      vortex_center, beta)
            reuse = amg if (amg is not None and outer % amg_rebuild_every) else None
            sysm = ExtendedSystem(mvop, rho_main, special, phi_ext, ff, vals,
                                  "wake_ls", te_kutta, amg=reuse)
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
            phid = to_dev(phi_ext[: mvop.n_main])
            op.state_dev(phid, u_inf ** 2, m_inf, gamma_air, q2cap, grad, q2n,
                         q2l, rho_new, lim)
            sp_new = mvop.special_densities(phi_ext, m_inf, gamma_air, u_inf)
            if rho_main is None:
                drho = 0.0
                rho_main = rho_new.clone()
                special = sp_new
            else:
                d_plain = float(((rho_new - rho_main) * mvop.d_plain_mask).abs().max().item())
                d_sp = max([float(np.abs(a - b).max()) if len(a) else 0.0
                            for a, b in zip(sp_new, special)] + [0.0])
                drho = max(d_plain, d_sp)
                rho_main = rho_main + omega_rho * (rho_new - rho_main)
                special = tuple(o + omega_rho * (n - o) for o, n in zip(special, sp_new))
            drho_history.append(drho)
        if verbose:
            print(f"  ls-picard {outer:2d}: |R|={residual:.3e} gamma={gamma:.6f} "
                  f"dgamma={dgamma:.2e} drho={drho:.2e} gmres={n_it} info={info}",
                  flush=True)
        if outer > 0 and dgamma < tol_gamma and drho < tol_rho:
            converged = True
            break
    m2 = empty(st.n_tets)
    _native.call("fp3d_mach_squared_field", st.n_tets, ptr(q2n), float(m_inf),
                 float(gamma_air), ptr(m2), stream_handle())
    finalize(timings, time.perf_counter() - t_wall0)
    return {
        "phi_ext": phi_ext, "phi": mvop.main_potential(phi_ext),
        "gamma": gamma, "cl_kj": 2.0 * gamma / u_inf,
        "te_jump": mvop.te_jump(phi_ext), "n_outer": len(drho_history),
        "converged": converged, "gamma_history": gamma_history,
        "drho_history": drho_history, "residual_history": residual_history,
        "residual_norm": residual, "timings": timings,
        "mach2_max": float(m2.max().item()), "farfield": farfield,
        "n_gmres_total": n_gmres_total, "n_gmres_stalled": n_gmres_stalled,
    }

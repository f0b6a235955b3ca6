"""Discrete adjoint of the level-set lifting solve for dCL/dalpha (no
reference counterpart: the reference has no adjoint).

At a converged state phi (frozen per-side upwind selection, freestream
far field phi_D(alpha) on the far-field main DOFs):

    R(phi_free; phi_D(alpha)) = E phi - f(alpha) = 0,   CL = g . phi
    J^T lam = g_free
    dCL/dalpha = - lam . dR/dalpha,
    dR/dalpha = J[:, D] dphi_D/dalpha - df/dalpha

(f: the Neumann outflow load, CL linear in the TE jumps).

J^T on the device: the Terms 1+2 CSR of the plain elements is symmetric
(element matrices rho V B B^T + w (B g)(B g)^T), the plain Term 3 is
transposed as a COO -> CSR with rows and columns swapped, the host sheet /
Kutta correction is transposed on the host and uploaded.  J[:, D] v is
taken as a central difference of the frozen-selection residual along
v = dphi_D/dalpha (exact up to O(h^2); R is smooth at a frozen selection).
The forward (tangent) solve J dphi = -dR/dalpha gives the same number and
is returned beside it as the check.
"""

from __future__ import annotations

import time
from typing import Dict

import numpy as np
import scipy.sparse.linalg as spla
import torch

from .. import _native
from ..device import empty, ptr, stream_handle, to_dev, to_host
from .adjoint import _csr_from_coo
from .amg import DeviceCSR
from .linear import gmres_dev
from .newton_ls import LSNewtonSystem
from .picard_ls import _farfield_main


class LSTransposeOperator:
    """x -> J^T x for a linearised LSNewtonSystem, and the transposed
    block preconditioner (AMG main block + exact aux block)."""

    def __init__(self, sysm: LSNewtonSystem):
        self.s = sysm
        st = sysm.mv.op.store
        self.st = st
        nm, nT = sysm.nm, sysm.nT
        grad, up, su, inv_u2 = sysm.t3_inputs
        active = torch.nonzero(su != 0.0).flatten().to(torch.int32)
        na = int(active.numel())
        rows = empty(16 * na, torch.int32)
        cols = empty(16 * na, torch.int32)
        vals = empty(16 * na)
        if na:
            _native.call("fp3d_fill_term3_coo", na, ptr(active), ptr(st.el),
                         ptr(st.B), ptr(st.V), ptr(grad), ptr(up.to(torch.int32)),
                         ptr(su), float(inv_u2), ptr(rows), ptr(cols), ptr(vals),
                         stream_handle())
        self.T3t = _csr_from_coo(cols, rows, vals, nm) if na else None
        self.fmd = sysm.nop.free_mask.to(torch.float64)
        Jc = sysm.Jc_host
        self.JcT = DeviceCSR(Jc.T.tocsr())
        self.ma_T = DeviceCSR(Jc[:nm, nm:].T.tocsr()) if nT > nm else None
        self.luT = spla.splu(Jc[nm:, nm:].T.tocsc()) if nT > nm else None
        self._t = empty(nm)
        self._ta = empty(max(nT - nm, 1))

    def matvec(self, x, y):
        s, st = self.s, self.st
        nm = s.nm
        _native.call("fp3d_csr_spmv", nm, st.nnz, ptr(st.indptr), ptr(st.indices),
                     ptr(s.nop.data), ptr(x[:nm]), ptr(y[:nm]), stream_handle())
        if self.T3t is not None:
            xm = x[:nm] * self.fmd
            self.T3t.matvec(xm, self._t)
            y[:nm].add_(self._t * self.fmd)
        if s.nT > nm:
            y[nm:].zero_()
        J = self.JcT
        _native.call("fp3d_amg_prolong_add", J.n, J.nnz, ptr(J.indptr),
                     ptr(J.indices), ptr(J.data), 0, ptr(x), ptr(y),
                     stream_handle())
        return y

    def psolve(self, r, z):
        s = self.s
        nm = s.nm
        s._amg.apply(r[:nm].contiguous(), z[:nm])
        if s.nT > nm:
            C = self.ma_T
            _native.call("fp3d_csr_spmv", C.n, C.nnz, ptr(C.indptr), ptr(C.indices),
                         ptr(C.data), ptr(z[:nm]), ptr(self._ta), stream_handle())
            rhs = to_host(r[nm:]) - to_host(self._ta)
            z[nm:].copy_(to_dev(self.luT.solve(rhs)))
        return z


def lift_alpha_sensitivity(case, phi_ext, m_inf: float, alpha_deg: float,
                           upwind_c: float = 1.5, m_crit: float = 0.95,
                           gamma_air: float = 1.4, rtol: float = 1e-12,
                           restart: int = 60, maxiter: int = 40,
                           tangent: bool = True) -> Dict[str, float]:
    """dCL/dalpha (per degree) of a PlateWingCase state by the adjoint (and
    the tangent check).  Returns cl, dcl_dalpha, dcl_dalpha_tangent,
    n_gmres_adjoint / tangent, seconds_adjoint / tangent, n_limited."""
    mv = case.mvop
    mesh = case.mesh
    u_inf = case.u_inf
    f = case.outflow_load(alpha_deg)
    sysm = LSNewtonSystem(mv, m_inf, upwind_c, m_crit, gamma_air, u_inf,
                          external_rhs=f)
    phi = np.array(phi_ext, dtype=np.float64)
    nodes = mesh.nodes
    ff, vals = _farfield_main(nodes, mesh.boundary_faces["farfield"], alpha_deg, 0.0,
                              u_inf, (0.0, 0.0), 1.0, "freestream", 2)
    phi[ff] = vals
    free = np.ones(mv.n_total, dtype=bool)
    free[ff] = False
    frozen = sysm.freeze(phi)
    R, sides, C, w = sysm.residual(phi, frozen, sens=True)
    n_lim = mv.n_limited + mv.n_floored
    sysm.linearize(phi, sides, C, w, ff)
    g = case.lift_gradient()
    g[~free] = 0.0
    # dR/dalpha along dphi_D/dalpha (per degree)
    a = np.deg2rad(alpha_deg)
    v = np.zeros(mv.n_total)
    v[ff] = u_inf * np.deg2rad(1.0) * (-nodes[ff, 0] * np.sin(a) + nodes[ff, 2] * np.cos(a))
    h = 1e-6 * max(1.0, float(np.abs(phi).max())) / max(float(np.abs(v).max()), 1e-300)
    dR = (sysm.residual(phi + h * v, frozen)[0] - sysm.residual(phi - h * v, frozen)[0]) / (2 * h)
    dR -= case.outflow_load(alpha_deg, derivative=True)
    dR[~free] = 0.0
    out = {"cl": case.lift_coefficient(phi), "n_limited_or_floored": int(n_lim)}
    t0 = time.perf_counter()
    T = LSTransposeOperator(sysm)
    lam, info, n_it = gmres_dev(T.matvec, to_dev(g), T.psolve, rtol=rtol,
                                restart=restart, maxiter=maxiter)
    torch.cuda.synchronize()
    out["dcl_dalpha"] = -float(np.dot(to_host(lam), dR))
    out["n_gmres_adjoint"] = int(n_it)
    out["adjoint_info"] = int(info)
    out["seconds_adjoint"] = time.perf_counter() - t0
    if tangent:
        t0 = time.perf_counter()
        d, info_t, n_t = sysm.solve_step(-dR, rtol, restart, maxiter, "block")
        out["dcl_dalpha_tangent"] = float(np.dot(g, d))
        out["n_gmres_tangent"] = int(n_t)
        out["tangent_info"] = int(info_t)
        out["seconds_tangent"] = time.perf_counter() - t0
    return out


class PlateLift(torch.autograd.Function):
    """CL(alpha) of the embedded-plate case as an autograd Function: the
    forward pass runs the level-set Newton, the backward pass the adjoint."""

    @staticmethod
    def forward(ctx, alpha_deg, case, m_inf, solve_kw):
        from .newton_ls import solve_multivalued_newton

        a = float(alpha_deg.item())
        kw = dict(solve_kw or {})
        kw.setdefault("phi_init", case.initial_state(a))
        kw.setdefault("farfield", "freestream")
        kw.setdefault("lift_axis", 2)
        kw.setdefault("external_rhs", case.outflow_load(a))
        res = solve_multivalued_newton(case.mvop, case.mesh, m_inf, alpha_deg=a, **kw)
        if not res["converged"]:
            raise RuntimeError("level-set Newton did not converge")
        ctx.state = (case, res["phi_ext"], m_inf, a, kw.get("upwind_c", 1.5),
                     kw.get("m_crit", 0.95))
        return alpha_deg.new_tensor(case.lift_coefficient(res["phi_ext"]))

    @staticmethod
    def backward(ctx, gout):
        case, phi, m_inf, a, c, mc = ctx.state
        s = lift_alpha_sensitivity(case, phi, m_inf, a, upwind_c=c, m_crit=mc,
                                   tangent=False)
        return gout * s["dcl_dalpha"], None, None, None

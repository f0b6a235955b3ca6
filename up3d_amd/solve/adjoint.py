"""Discrete adjoint of the channel Newton solve, behind torch.autograd.

(New functionality: the reference has no adjoint.)  For the converged
state phi(p) of R_free(phi, p) = 0 with parameters p = (floor-node y
coordinates, M_inf, alpha) and an output functional F(phi, p) -- the wall
pressure force cf (pyfp3d/post/surface.py:289-363), weighted x-force and
lift -- the gradient is

    J_ff^T lambda = (dF/dphi)_free                       (one GMRES solve)
    dF/dp = dF/dp|_phi + dF/dphi_D . dphi_D/dp - lambda^T (dR/dp + J_fD dphi_D/dp)

taken at the FROZEN upstream selection and branch of the converged state
(the same frozen linearisation the Newton finish uses).  Device pieces:
J_ff^T as one CSR, the free rows of J12 + T3^T (Terms 1+2 are symmetric;
Term 3 through its COO sorted by column), so the adjoint GMRES takes the
one-call Arnoldi step; dR/dM_inf from the branch-wise d rho_tilde/dM
kernel and the coloured residual scatter; lambda^T dR/dx from
fp3d_adjoint_dx; the Dirichlet chains from the unmasked J^T lambda.  The
functional's own derivatives (a 2-D wall sum) use torch autograd on the
wall nodes, on the device.  alpha is in degrees (freestream_phi
convention); d/d alpha is per degree.
"""

from __future__ import annotations

import ctypes as C
from typing import Dict, Optional

import numpy as np
import torch

from .. import _native
from ..constraints.dirichlet import freestream_phi
from ..device import empty, ptr, stream_handle, to_dev, to_host, zeros
from ..kernels.jacobian import PicardOperator
from ..kernels.upwind import UpwindOperator
from .amg import DeviceCSR
from .common import Reducer, dirichlet_setup
from .linear import axpby, dot, gmres_dev
from .newton import solve_newton


def _csr_from_coo(rows: torch.Tensor, cols: torch.Tensor, vals: torch.Tensor,
                  n: int) -> DeviceCSR:
    """Device CSR from COO triplets: stable sort by (row, col); duplicate
    entries stay separate (the SpMV sums them in a fixed order)."""
    key = rows.long() * n + cols.long()
    order = torch.argsort(key, stable=True)
    r = rows.long()[order]
    csr = DeviceCSR.__new__(DeviceCSR)
    counts = torch.bincount(r, minlength=n)
    indptr = torch.zeros(n + 1, dtype=torch.int64, device=rows.device)
    indptr[1:] = torch.cumsum(counts, 0)
    csr.shape = (n, n)
    csr.n = n
    csr.nnz = int(len(order))
    csr.indptr = indptr.to(torch.int32).contiguous()
    csr.indices = cols[order].to(torch.int32).contiguous()
    csr.data = vals[order].to(torch.float64).contiguous()
    csr.data32 = None
    csr.is_f32 = 0
    return csr


def _masked_sum_csr(st, J12: torch.Tensor, T: DeviceCSR, fm: torch.Tensor) -> DeviceCSR:
    """One CSR of the free rows of J12 + T (J12 on the store's pattern, T a
    CSR of the same size; rows with fm == 0 are empty, so its SpMV writes
    zero there).  Entries of one (row, column) from both stay separate, J12's
    first: the SpMV sums a row's entries in a fixed order."""
    n = st.n_nodes
    dev = J12.device

    def rows_of(indptr):
        return torch.repeat_interleave(torch.arange(n, device=dev),
                                       (indptr[1:] - indptr[:-1]).long())

    rows = torch.cat([rows_of(st.indptr), rows_of(T.indptr)])
    cols = torch.cat([st.indices, T.indices])
    vals = torch.cat([J12, T.data])
    keep = fm.bool()[rows]
    return _csr_from_coo(rows[keep], cols[keep], vals[keep], n)


def _wall_functional(nodes, phi, m_inf, alpha_deg, wf, flip, wx, wl, s_ref=1.0,
                     gamma=1.4):
    """torch restatement of the wall force (triangle tangential gradient,
    isentropic Cp, cf = -sum Cp A n_out / S_ref); returns wx*cf_x + wl*cl."""
    p0, p1, p2 = nodes[wf[:, 0]], nodes[wf[:, 1]], nodes[wf[:, 2]]
    ea, eb = p1 - p0, p2 - p0
    nv = torch.cross(ea, eb, dim=1)
    twice = torch.linalg.norm(nv, dim=1)
    nrm = nv / twice[:, None]
    t1 = ea / torch.linalg.norm(ea, dim=1)[:, None]
    t2 = torch.cross(nrm, t1, dim=1)
    a1, b1 = (ea * t1).sum(1), (ea * t2).sum(1)
    a2, b2 = (eb * t1).sum(1), (eb * t2).sum(1)
    f1 = phi[wf[:, 1]] - phi[wf[:, 0]]
    f2 = phi[wf[:, 2]] - phi[wf[:, 0]]
    det = a1 * b2 - b1 * a2
    g = ((b2 * f1 - b1 * f2) / det)[:, None] * t1 + ((a1 * f2 - a2 * f1) / det)[:, None] * t2
    q2 = (g * g).sum(1)
    base = 1.0 + 0.5 * (gamma - 1.0) * m_inf * m_inf * (1.0 - q2)
    rho = torch.clamp(base, min=1e-10) ** (1.0 / (gamma - 1.0))
    cp = 2.0 / (gamma * m_inf * m_inf) * (rho ** gamma - 1.0)
    n_out = nrm * flip[:, None]
    cf = -(cp * 0.5 * twice) @ n_out / s_ref
    a = alpha_deg * (np.pi / 180.0)
    cl = -torch.sin(a) * cf[0] + torch.cos(a) * cf[1]
    return wx * cf[0] + wl * cl, cf, cl


class ChannelForceProblem:
    """Bump-channel problem with Dirichlet phi = freestream_phi on the
    inflow/outflow nodes, slip walls, and the wall force on `wall_faces` as
    the output.  Parameters: y of `floor_nodes`, M_inf, alpha (deg)."""

    def __init__(self, nodes, elements, dirichlet_nodes, wall_faces, floor_nodes,
                 u_inf: float = 1.0, precision: str = "fp64",
                 newton_kw: Optional[dict] = None):
        self.base_nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.elements = np.ascontiguousarray(elements, dtype=np.int32)
        self.dn = np.asarray(dirichlet_nodes, dtype=np.int64)
        self.wall_faces = np.asarray(wall_faces, dtype=np.int64)
        self.floor_nodes = np.asarray(floor_nodes, dtype=np.int64)
        self.u_inf = float(u_inf)
        self.precision = precision
        self.newton_kw = dict(newton_kw or {})
        from ..post.surface import wall_outward_normals

        nv = np.cross(self.base_nodes[self.wall_faces[:, 1]] - self.base_nodes[self.wall_faces[:, 0]],
                      self.base_nodes[self.wall_faces[:, 2]] - self.base_nodes[self.wall_faces[:, 0]])
        n_out = wall_outward_normals(self.base_nodes, self.elements, self.wall_faces)
        self.flip = np.sign(np.einsum("ij,ij->i", nv, n_out))
        # the functional touches only the wall nodes: its autograd runs on
        # them (faces renumbered locally) and scatters back
        self.wall_nodes, wl = np.unique(self.wall_faces, return_inverse=True)
        self.wall_faces_local = wl.reshape(self.wall_faces.shape)
        self._ops = None

    def nodes_for(self, y_floor) -> np.ndarray:
        nodes = self.base_nodes.copy()
        if y_floor is not None:
            nodes[self.floor_nodes, 1] = np.asarray(y_floor, dtype=np.float64)
        return nodes

    # ----------------------------------------------------------- forward
    def solve(self, y_floor, m_inf: float, alpha_deg: float, phi_init=None,
              tol_residual: float = 1e-13):
        nodes = self.nodes_for(y_floor)
        dv = freestream_phi(nodes[self.dn], alpha_deg, self.u_inf)
        if phi_init is None:
            phi_init = freestream_phi(nodes, alpha_deg, self.u_inf)
        kw = dict(n_picard_seed=0, freeze_tol=1e-6, n_newton_max=40,
                  tol_residual=tol_residual)
        kw.update(self.newton_kw)
        r = solve_newton(nodes, self.elements, self.dn, dv, m_inf,
                         phi_init=phi_init, u_inf=self.u_inf,
                         precision=self.precision, **kw)
        r["nodes"] = nodes
        r["m_inf"] = float(m_inf)
        r["alpha_deg"] = float(alpha_deg)
        # the solve's last AMG hierarchy: the adjoint and tangent solves at the
        # converged state precondition with it instead of building their own
        ws = r.get("workspace")
        r["precond"] = getattr(ws, "precond", None)
        if r["converged"]:
            self._settle(r, tol_residual)
        return r

    @staticmethod
    def _preconditioner(sol, st, A, fm):
        """The converged solve's AMG hierarchy when it is on this mesh, else
        one of the masked Picard matrix A."""
        from .amg import AMGHierarchy
        from .common import csr_to_host, mask_dirichlet

        ml = sol.get("precond")
        if ml is not None and ml.shapes[0] == st.n_nodes:
            return ml
        A = A.clone()
        mask_dirichlet(st, fm, A)
        return AMGHierarchy(csr_to_host(st, A), precision="f64", max_coarse=1000)

    def _settle(self, r, tol_residual: float) -> None:
        """Put the accepted state on the root of the selection it is
        differentiated at.  An accepted state solves the system of the
        selection the Newton froze; where the live selection at that state
        differs (an accepted assignment cycle), the state is not a root of
        the system frozen at its own selection, and derivatives of that
        system taken at it are off by the residual's image.  So the
        selection is frozen at the accepted state and the state re-converged
        on it (full Newton steps, usually none or one); `frozen` keeps that
        selection for the gradients."""
        frozen = self.frozen_selection(r)
        pol = self.solve_frozen(r["nodes"][self.floor_nodes, 1], r["m_inf"],
                                r["alpha_deg"], frozen, r["phi"],
                                tol_residual=tol_residual)
        r["frozen_residual_at_accept"] = float(pol["residual_history"][0])
        r["frozen_polish_steps"] = len(pol["residual_history"]) - 1
        if not pol["converged"]:
            r["converged"] = False
            return
        r["phi"] = pol["phi"]
        r["frozen"] = frozen

    def frozen_selection(self, sol, upwind_c: float = 1.5, m_crit: float = 0.95,
                         rho_floor: float = 0.05):
        """(upstream, branch) the gradients of `sol` are taken at (device
        int32 / int8 tensors): the selection `sol` was settled on, else the
        one frozen at its state."""
        if sol.get("frozen") is not None:
            return sol["frozen"][0].clone(), sol["frozen"][1].clone()
        L = self._linearisation(sol, upwind_c, m_crit, rho_floor)
        return L["up"].clone(), L["br"].clone()

    def solve_frozen(self, y_floor, m_inf: float, alpha_deg: float, frozen,
                     phi_init, tol_residual: float = 1e-13, n_newton_max: int = 20,
                     upwind_c: float = 1.5, m_crit: float = 0.95,
                     rho_floor: float = 0.05):
        """Newton on the system with the upstream selection and branch held at
        `frozen` (frozen_selection of a converged state) -- the smooth piece of
        the solution map whose derivative gradients() returns, so central
        differences of it check the adjoint without the selection switches a
        free re-solve may make.  Full Newton steps from phi_init (a nearby
        converged state), GMRES to 1e-12 with the AMG of the clamped Jacobian."""
        from .linear import scale
        from .newton import NewtonWorkspace

        nodes = self.nodes_for(y_floor)
        dv = freestream_phi(nodes[self.dn], alpha_deg, self.u_inf)
        ws = NewtonWorkspace(nodes, self.elements, self.dn, dv, u_inf=self.u_inf)
        ws.set_mach(m_inf)
        s = ws.states[0]
        s.phi.copy_(to_dev(np.asarray(phi_init, dtype=np.float64)))
        s.phi.index_copy_(0, ws.dn, ws.dv)
        rhs = empty(ws.n)
        hist = []
        converged = False
        for _ in range(n_newton_max):
            ws.eval_residual(s, upwind_c, m_crit, rho_floor, frozen=frozen)
            hist.append(s.r_norm)
            if s.r_norm < tol_residual:
                converged = True
                break
            ws.assemble_coupled(s, upwind_c, m_crit, rho_floor, frozen)
            M = ws.build_preconditioner(s, "jacobian_spd", 0.9, se=ws.se_used)
            scale(-1.0, s.R, rhs)
            x, info, _ = gmres_dev(ws.jac.matvec, rhs, M.apply, rtol=1e-12,
                                   restart=200, maxiter=20)
            axpby(1.0, x, 1.0, s.phi)
        return {"phi": to_host(s.phi).copy(), "nodes": nodes, "m_inf": float(m_inf),
                "alpha_deg": float(alpha_deg), "converged": converged,
                "residual_history": hist, "n_limited": int(s.n_limited)}

    def functional(self, sol, weights=(1.0, 0.0)):
        t = lambda a: torch.tensor(a, dtype=torch.float64)  # noqa: E731
        val, cf, cl = _wall_functional(t(sol["nodes"]), t(sol["phi"]),
                                       t(sol["m_inf"]), t(sol["alpha_deg"]),
                                       torch.from_numpy(self.wall_faces),
                                       t(self.flip), weights[0], weights[1])
        return float(val), cf.numpy(), float(cl)

    def _functional_grads(self, sol, weights, device):
        """(value, cf, cl, dF/dnodes (n, 3), dF/dphi (n,), dF/dM_inf,
        dF/dalpha) by autograd over the wall nodes, on `device` (a few
        thousand faces: on the GPU the sixty-odd small operations cost
        about a millisecond; on the host each one wakes the intra-op
        thread pool)."""
        wn = self.wall_nodes
        n = len(sol["phi"])
        with torch.enable_grad():
            def t(a):
                return torch.tensor(a, dtype=torch.float64, device=device,
                                    requires_grad=True)

            nodes_t = t(np.asarray(sol["nodes"])[wn])
            phi_t = t(np.asarray(sol["phi"])[wn])
            m_t, a_t = t(sol["m_inf"]), t(sol["alpha_deg"])
            val, cf, cl = _wall_functional(
                nodes_t, phi_t, m_t, a_t,
                torch.as_tensor(self.wall_faces_local, device=device),
                torch.as_tensor(self.flip, dtype=torch.float64, device=device),
                weights[0], weights[1])
            gx, gphi, gm, ga = torch.autograd.grad(val, (nodes_t, phi_t, m_t, a_t))
        dFx = np.zeros((n, 3))
        dFx[wn] = gx.cpu().numpy()
        dFphi = np.zeros(n)
        dFphi[wn] = gphi.cpu().numpy()
        return (val.detach().cpu(), cf.detach().cpu(), cl.detach().cpu(), dFx, dFphi,
                float(gm), float(ga))

    # ----------------------------------------------------------- adjoint
    def _operators(self, nodes):
        """(PicardOperator, UpwindOperator, free mask) of these node
        coordinates, kept for the last coordinates asked for."""
        if self._ops is not None and np.array_equal(self._ops[0], nodes):
            return self._ops[1:]
        op = PicardOperator(nodes, self.elements)
        upw = UpwindOperator(nodes, self.elements)
        fm = dirichlet_setup(op.store.n_nodes, self.dn, np.zeros(len(self.dn)),
                             op.store.device)[0]
        self._ops = (np.array(nodes, copy=True), op, upw, fm)
        return op, upw, fm

    @staticmethod
    def _term3(L, transpose: bool) -> DeviceCSR:
        """Term 3 of the linearisation as a row-sorted (transpose: column-
        sorted) CSR, built on first use."""
        key = "T3t" if transpose else "T3"
        if key not in L:
            rows, cols, vals = L["t3_coo"]
            n = L["st"].n_nodes
            L[key] = _csr_from_coo(cols, rows, vals, n) if transpose else \
                _csr_from_coo(rows, cols, vals, n)
        return L[key]

    def _linearisation(self, sol, upwind_c=1.5, m_crit=0.95, rho_floor=0.05):
        """Device state, frozen (upstream, branch), sensitivities, Terms 1+2
        (unmasked) and Term 3's COO (CSRs of it through _term3)."""
        op, upw, fm = self._operators(sol["nodes"])
        st = op.store
        n, nt = st.n_nodes, st.n_tets
        m = sol["m_inf"]
        phi = to_dev(sol["phi"])
        grad = empty((nt, 3))
        q2n = empty(nt)
        q2l = empty(nt)
        rho = empty(nt)
        lim = empty(nt, torch.uint8)
        from ..physics.isentropic import q2_at_mach

        op.state_dev(phi, self.u_inf ** 2, m, 1.4, q2_at_mach(3.0, m), grad, q2n,
                     q2l, rho, lim)
        if sol.get("frozen") is not None:
            up, br = sol["frozen"]
        else:
            up, br = upw.freeze_dev(grad, q2l, rho, m, upwind_c, m_crit, 1.4, rho_floor)
        rt = upw.frozen_dev(q2l, rho, up, br, m, upwind_c, m_crit, 1.4, rho_floor,
                            monitors=False).clone()
        se, su = upw.frozen_sensitivities_dev(q2l, rho, up, br, m, upwind_c,
                                              m_crit, 1.4, rho_floor)
        se, su = se.clone(), su.clone()
        inv_u2 = 1.0 / self.u_inf ** 2
        J12 = empty(st.nnz)
        op.jacobian_dev(rt, grad, se, None, inv_u2, J12)
        active = torch.nonzero(su != 0.0).flatten().to(torch.int32)
        na = int(active.numel())
        rows = empty(16 * na, torch.int32)
        cols = empty(16 * na, torch.int32)
        vals = empty(16 * na)
        if na:
            _native.call("fp3d_fill_term3_coo", na, ptr(active), ptr(st.el),
                         ptr(st.B), ptr(st.V), ptr(grad), ptr(up), ptr(su),
                         float(inv_u2), ptr(rows), ptr(cols), ptr(vals),
                         stream_handle())
        A = empty(st.nnz)
        op.matrix_dev(rt, A)
        return dict(op=op, upw=upw, st=st, grad=grad, q2l=q2l, rho=rho, rt=rt,
                    up=up, br=br, se=se, su=su, J12=J12, t3_coo=(rows, cols, vals),
                    fm=fm, A=A, n_limited=int((lim == 0).sum().item()))

    def _jmv(self, L, x, y, transpose):
        """y = J x (or J^T x), unmasked."""
        st = L["st"]
        tmp = empty(st.n_nodes)
        _native.call("fp3d_csr_spmv", st.n_nodes, st.nnz, ptr(st.indptr),
                     ptr(st.indices), ptr(L["J12"]), ptr(x), ptr(y),
                     stream_handle())
        T = self._term3(L, transpose)
        if T.nnz:
            T.matvec(x, tmp)
            axpby(1.0, tmp, 1.0, y)
        return y

    def gradients(self, sol, weights=(1.0, 0.0), rtol: float = 1e-12,
                  upwind_c: float = 1.5, m_crit: float = 0.95) -> Dict[str, object]:
        """dF/d(y_floor, M_inf, alpha) for F = w0 cf_x + w1 cl at `sol`."""
        L = self._linearisation(sol, upwind_c, m_crit)
        if L["n_limited"]:
            raise RuntimeError("adjoint needs a state without speed-limited elements")
        st = L["st"]
        n, nt = st.n_nodes, st.n_tets
        fm = L["fm"]
        # functional derivatives (device autograd; also valid inside backward)
        val, cf, cl, dFx, dFphi, dFm, dFa = self._functional_grads(sol, weights,
                                                                   st.device)
        # adjoint solve J_ff^T lam = (dF/dphi)_free
        b = to_dev(dFphi)
        b.mul_(fm.to(torch.float64))
        ml = self._preconditioner(sol, st, L["A"], fm)
        # J_ff^T (J12 is symmetric) with the Dirichlet rows emptied
        JfT = _masked_sum_csr(st, L["J12"], self._term3(L, True), fm)
        lam, info, n_it = gmres_dev(JfT.matvec, b, ml.apply, rtol=rtol, restart=256,
                                    maxiter=12)
        red_norm = Reducer()
        # dR/dM_inf = scatter of d rho_tilde/dM
        drt = empty(nt)
        _native.call("fp3d_drho_tilde_dmach", nt, ptr(L["q2l"]), ptr(L["rho"]),
                     ptr(L["up"]), ptr(L["br"]), float(sol["m_inf"]), float(m_crit),
                     float(upwind_c), 1.4, ptr(drt), stream_handle())
        RM = L["op"].residual_dev(drt, L["grad"], out=empty(n))
        dF_dm = float(dFm) - dot(lam, RM)
        # J^T lam, unmasked: the Dirichlet-value chains of alpha and of the
        # Dirichlet nodes' coordinates (lam^T J v = (J^T lam) . v)
        JTl_h = to_host(self._jmv(L, lam, empty(n), transpose=True))
        # alpha: Dirichlet values phi_D = x cos a + y sin a (a in degrees)
        nodes = sol["nodes"]
        a = np.deg2rad(sol["alpha_deg"])
        vd = np.zeros(n)
        vd[self.dn] = (-nodes[self.dn, 0] * np.sin(a)
                       + nodes[self.dn, 1] * np.cos(a)) * (np.pi / 180.0)
        dF_da = float(dFa) + float(dFphi @ vd) - float(JTl_h @ vd)
        # node coordinates: lambda^T dR/dx (device) + Dirichlet-value chain
        work = empty(int(_native.lib().fp3d_adjoint_dx_work_doubles(nt)))
        up_l = L["up"].long()
        idx = torch.arange(nt, device=up_l.device)
        down = up_l != idx
        order = torch.argsort(up_l[down] * nt + idx[down], stable=True)
        dsrc = idx[down][order].to(torch.int32).contiguous()
        cnt = torch.bincount(up_l[down], minlength=nt)
        dptr = torch.zeros(nt + 1, dtype=torch.int64, device=up_l.device)
        dptr[1:] = torch.cumsum(cnt, 0)
        dptr = dptr.to(torch.int32).contiguous()
        dRx = empty((n, 3))
        _native.call("fp3d_adjoint_dx", nt, n, ptr(st.el), ptr(st.B), ptr(st.V),
                     ptr(L["grad"]), ptr(L["rt"]), ptr(L["se"]), ptr(L["su"]),
                     ptr(dptr), ptr(dsrc), ptr(st.inc_ptr), ptr(st.inc_ent),
                     ptr(lam), ptr(work), ptr(dRx), stream_handle())
        dRx_h = to_host(dRx)
        grad_x = dFx - dRx_h
        # Dirichlet nodes: phi_D = x cos a + y sin a
        coef = dFphi[self.dn] - JTl_h[self.dn]
        grad_x[self.dn, 0] += coef * np.cos(a)
        grad_x[self.dn, 1] += coef * np.sin(a)
        return {"d_y_floor": grad_x[self.floor_nodes, 1].copy(),
                "d_nodes": grad_x, "d_m_inf": dF_dm, "d_alpha_deg": dF_da,
                "value": float(val), "cf": cf.numpy(), "cl": float(cl),
                "adjoint_iters": n_it, "adjoint_info": info,
                "lambda": to_host(lam)}

    def jvp_params(self, sol, dm: float, dalpha_deg: float, weights=(1.0, 0.0),
                   rtol: float = 1e-12, upwind_c=1.5, m_crit=0.95) -> float:
        """Forward tangent dF for a perturbation of (M_inf, alpha):
        J_ff dphi_f = -(dR/dM dm + J_fD dphi_D), dF = F_p + F_phi . dphi."""
        L = self._linearisation(sol, upwind_c, m_crit)
        st = L["st"]
        n, nt = st.n_nodes, st.n_tets
        fmd = L["fm"].to(torch.float64)
        drt = empty(nt)
        _native.call("fp3d_drho_tilde_dmach", nt, ptr(L["q2l"]), ptr(L["rho"]),
                     ptr(L["up"]), ptr(L["br"]), float(sol["m_inf"]), float(m_crit),
                     float(upwind_c), 1.4, ptr(drt), stream_handle())
        RM = L["op"].residual_dev(drt, L["grad"], out=empty(n))
        nodes = sol["nodes"]
        a = np.deg2rad(sol["alpha_deg"])
        vd = np.zeros(n)
        vd[self.dn] = (-nodes[self.dn, 0] * np.sin(a)
                       + nodes[self.dn, 1] * np.cos(a)) * (np.pi / 180.0) * dalpha_deg
        Jv = self._jmv(L, to_dev(vd), empty(n), transpose=False)
        rhs = empty(n)
        rhs.copy_(RM)
        rhs.mul_(dm)
        axpby(1.0, Jv, 1.0, rhs)
        rhs.mul_(-fmd)
        ml = self._preconditioner(sol, st, L["A"], L["fm"])
        Jf = _masked_sum_csr(st, L["J12"], self._term3(L, False), L["fm"])
        dphi, info, _ = gmres_dev(Jf.matvec, rhs, ml.apply, rtol=rtol, restart=256,
                                  maxiter=12)
        dphi_h = to_host(dphi) + vd
        _, _, _, _, gphi, gm, ga = self._functional_grads(sol, weights, st.device)
        return gm * dm + ga * dalpha_deg + float(gphi @ dphi_h)


class ChannelForce(torch.autograd.Function):
    """(cf_x, cl) of the converged channel solve as a differentiable function
    of (y_floor, M_inf, alpha_deg); backward = one adjoint solve."""

    @staticmethod
    def forward(ctx, y_floor, m_inf, alpha_deg, problem: ChannelForceProblem,
                phi_init=None):
        sol = problem.solve(y_floor.detach().cpu().numpy(), float(m_inf),
                            float(alpha_deg), phi_init=phi_init)
        if not sol["converged"]:
            raise RuntimeError("ChannelForce: Newton solve did not converge")
        _, cf, cl = problem.functional(sol)
        ctx.problem = problem
        ctx.sol = sol
        ctx.dev = y_floor.device
        return torch.tensor([cf[0], cl], dtype=torch.float64, device=y_floor.device)

    @staticmethod
    def backward(ctx, gout):
        w = gout.detach().cpu().numpy().astype(np.float64)
        g = ctx.problem.gradients(ctx.sol, weights=(float(w[0]), float(w[1])))
        return (torch.tensor(g["d_y_floor"], dtype=torch.float64, device=ctx.dev),
                torch.tensor(g["d_m_inf"], dtype=torch.float64),
                torch.tensor(g["d_alpha_deg"], dtype=torch.float64), None, None)

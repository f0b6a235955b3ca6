"""Newton solve of the level-set (multivalued) lifting system
(pyfp3d/solve/newton_ls.py:115-987, LSNewtonSystem and
solve_multivalued_newton, restated for the vortex far field).

Residual at a given per-side upwind selection (live or frozen):

    R = E(rho_tilde) phi,  R[TE aux] = |q_u|^2 - |q_l|^2

with E the multivalued Picard matrix (mass conservation per side copy,
wake least-squares rows, frozen-mean Kutta rows) -- its plain-element part
and its sheet part (MultivaluedOperator.sheet_fill) both assembled on the
device.  Jacobian:

    J = D_keep E  +  exact Kutta Jacobian  +  Terms 2+3 per side

(D_keep drops the TE aux rows; Term-2/3 rows landing on non-TE aux rows are
dropped -- those rows are exactly linear -- and TE aux rows re-routed to the
TE main row, as the mass conservation is).  All of it is assembled on the
device: Terms 1+2 of the plain elements (node-gather CSR) and Term 3 of
plain elements whose upstream element is plain (the Term-3 gather
operator); the sheet part (D_keep C, the exact Kutta Jacobian and Term 2
of the special copies) filled into its fixed pattern; Term 3 of the
special copies and of plain elements with a special upstream element as
the SpecialTerm3 operator, rebuilt per linearisation because its columns
follow the upwind selection.  The linear step is device GMRES with the
block preconditioner of the Picard path (AMG of the single-valued Picard
matrix + exact aux-block solve, the aux block's sparse LU on the host).
"""

from __future__ import annotations

import time
from typing import Dict, Optional

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

from .. import _native
from ..device import empty, ptr, stream_handle, to_dev, to_host
from .amg import DeviceCSR
from .common import amg_from_device, mask_dirichlet
from .linear import NewtonOperator, gmres_dev
from .picard_ls import _farfield_main, solve_multivalued_lifting
from .timing import finalize, new_timings, phase


class LSNewtonSystem:
    def __init__(self, mvop, m_inf, upwind_c=1.5, m_crit=0.95, gamma_air=1.4,
                 u_inf=1.0, m_cap=3.0, rho_floor=0.05, external_rhs=None):
        self.mv = mvop
        self.m_inf, self.upwind_c, self.m_crit = float(m_inf), float(upwind_c), float(m_crit)
        self.gamma_air, self.u_inf = float(gamma_air), float(u_inf)
        self.m_cap, self.rho_floor = float(m_cap), float(rho_floor)
        self.external_rhs = external_rhs
        self.nm, self.nT = mvop.n_main, mvop.n_total
        cm = mvop.cm
        self.te_aux = cm.ext_dof_of_node[cm.te_nodes].astype(np.int64)
        self.row_map = np.arange(self.nT, dtype=np.int64)
        self.row_map[mvop._nonte_rows] = -1
        self.row_map[self.te_aux] = cm.te_nodes
        self.is_plain_dev = to_dev(((~mvop.is_cut) & (~mvop.is_tel)).astype(np.uint8),
                                   torch.uint8).bool()
        self.pp_dev = to_dev(mvop.plain_plus.astype(np.uint8), torch.uint8).bool()
        self.vals = empty(mvop.op.nnz)
        self.nop = None
        self._amg = None
        self._amg_age = 0
        self._jsheet = None
        self._t3s = None
        self._jc_host = None

    # ------------------------------------------------------------ residual
    def _sides(self, phi, frozen, sens):
        return self.mv.side_state_dev(phi, self.m_inf, self.upwind_c, self.m_crit,
                                      self.gamma_air, self.u_inf, self.m_cap,
                                      self.rho_floor, frozen=frozen,
                                      sensitivities=sens)

    def freeze(self, phi):
        return self.mv.freeze_side_state(phi, self.m_inf, self.upwind_c,
                                         self.m_crit, self.gamma_air, self.u_inf,
                                         self.m_cap, self.rho_floor)

    def te_residual(self, phi):
        qu, ql = self.mv.te_velocities(phi)
        return np.einsum("ij,ij->i", qu, qu) - np.einsum("ij,ij->i", ql, ql)

    def residual(self, phi, frozen=None, sens=False):
        """(R host (n_total,), side data, the Picard sheet terms C (device
        SheetMatrix), plain-part weights w (device))."""
        mv = self.mv
        op, st = mv.op, mv.op.store
        nm = self.nm
        sides = self._sides(phi, frozen, sens)
        w, special = mv.side_weights(sides[0]["rho_tilde"], sides[1]["rho_tilde"])
        op.matrix_dev(mv.plain_weight_dev(w), self.vals)
        C = mv.sheet_fill(special, "pressure", phi, user="newton")
        y = empty(self.nT)
        x = to_dev(np.asarray(phi, dtype=np.float64))
        _native.call("fp3d_csr_spmv", nm, op.nnz, ptr(st.indptr),
                     ptr(st.indices), ptr(self.vals), ptr(x[:nm]), ptr(y[:nm]),
                     stream_handle())
        if self.nT > nm:
            y[nm:].zero_()
        C.add_matvec(x, y)
        R = to_host(y)
        if self.external_rhs is not None:
            R -= self.external_rhs
        R[self.te_aux] = self.te_residual(phi)
        return R, sides, C, w

    # ------------------------------------------------------------ Jacobian
    def _kutta_jacobian(self, phi):
        """d(|q_u|^2 - |q_l|^2)/dphi on the TE aux rows
        (pyfp3d/kernels/cut_assembly.py:478-515)."""
        mv = self.mv
        qu, ql = mv.te_velocities(phi)
        rows, cols, data = [], [], []
        for i, t in enumerate(mv.cm.te_nodes):
            cv = mv._te_cv[i]
            r = mv.cm.ext_dof_of_node[t]
            for key, q in (("upper", 2.0 * qu[i]), ("lower", -2.0 * ql[i])):
                e, d = cv[key + "_elems"], cv[key + "_dofs"]
                if not len(e):
                    continue
                wv = mv.V[e] / mv.V[e].sum()
                c = wv[:, None] * np.einsum("ead,d->ea", mv.B[e], q)
                rows.append(np.full(d.size, r, dtype=np.int64))
                cols.append(d.reshape(-1))
                data.append(c.reshape(-1))
        if not rows:
            z = np.empty(0, dtype=np.int64)
            return z, z, np.empty(0)
        return np.concatenate(rows), np.concatenate(cols), np.concatenate(data)

    def _terms23_host(self, sides):
        """Terms 2+3 of the special copies and Term 3 of plain elements whose
        upstream element is special (pyfp3d/kernels/cut_assembly.py:401-475),
        rows re-mapped, as a host COO: the restatement the device sheet and
        SpecialTerm3 are checked against (the solver uses those)."""
        mv = self.mv
        cm = mv.cm
        inv_u2 = 1.0 / (self.u_inf ** 2)
        read = mv.read_dofs()
        special = np.zeros(len(mv.el), dtype=bool)
        special[mv.special] = True
        parts = []
        for k, (sd, own, dofs_sp, ids_sp) in enumerate((
                (sides[0], mv.plain_plus, cm.dofs_upper, cm.cut_elems),
                (sides[1], (~mv.is_cut) & (~mv.is_tel) & ~mv.plain_plus,
                 np.concatenate([cm.dofs_lower, mv.tel_dof]),
                 np.concatenate([cm.cut_elems, cm.te_lower_elems])))):
            rd = read[k]
            g = to_host(sd["grad"])
            lim = to_host(sd["lim"]).astype(np.float64)
            se = to_host(sd["s_e"]) * lim
            up = to_host(sd["upstream"]).astype(np.int64)
            su = to_host(sd["s_u"]) * lim[up]
            gB = np.einsum("ed,ead->ea", g, mv.B)
            dofvec = mv.el.copy()
            dofvec[ids_sp] = dofs_sp
            # Term 2 on the special copies
            e2 = ids_sp[se[ids_sp] != 0.0]
            if len(e2):
                w2 = 2.0 * inv_u2 * se[e2] * mv.V[e2]
                blk = w2[:, None, None] * gB[e2][:, :, None] * gB[e2][:, None, :]
                parts.append((np.repeat(dofvec[e2], 4, axis=1).reshape(-1),
                              np.tile(dofvec[e2], (1, 4)).reshape(-1), blk.reshape(-1)))
            # Term 3: special e (any u), or own-side plain e with special u
            cand = np.concatenate([ids_sp, np.flatnonzero(own & special[up])])
            e3 = cand[su[cand] != 0.0]
            if len(e3):
                u3 = up[e3]
                w3 = 2.0 * inv_u2 * su[e3] * mv.V[e3]
                blk = w3[:, None, None] * gB[e3][:, :, None] * gB[u3][:, None, :]
                parts.append((np.repeat(dofvec[e3], 4, axis=1).reshape(-1),
                              np.tile(rd[u3], (1, 4)).reshape(-1), blk.reshape(-1)))
        if not parts:
            z = np.empty(0, dtype=np.int64)
            return z, z, np.empty(0)
        r = np.concatenate([p[0] for p in parts])
        c = np.concatenate([p[1] for p in parts])
        d = np.concatenate([p[2] for p in parts])
        rr = self.row_map[r]
        keep = rr >= 0
        return rr[keep], c[keep], d[keep]

    def _jacobian_sheet(self):
        """Device pattern of the sheet part of J: D_keep C (the Picard sheet
        terms without the TE aux rows), the exact Kutta Jacobian on the TE
        aux rows (values per call) and Term 2 of the special copies (rows
        re-mapped, dropped rows removed)."""
        if self._jsheet is None:
            from ..wake.sheet import HOSTVAL, TERM2_LOWER, TERM2_UPPER, SheetMatrix

            mv = self.mv
            cm = mv.cm
            r, c, k, s, ab, coef = mv.sheet_contributions("pressure")
            is_te_aux = np.zeros(self.nT, dtype=bool)
            is_te_aux[self.te_aux] = True
            keep = ~is_te_aux[r]
            parts = [tuple(a[keep] for a in (r, c, k, s, ab, coef))]
            kr, kc, _ = self._kutta_jacobian(np.zeros(self.nT))
            parts.append((kr, kc, np.full(len(kr), HOSTVAL, dtype=np.uint8),
                          np.arange(len(kr)), np.zeros(len(kr), dtype=np.int64),
                          np.ones(len(kr))))
            inv_u2 = 1.0 / (self.u_inf ** 2)
            for kind, ids, dofs in (
                    (TERM2_UPPER, cm.cut_elems, cm.dofs_upper),
                    (TERM2_LOWER, np.concatenate([cm.cut_elems, cm.te_lower_elems]),
                     np.concatenate([cm.dofs_lower, mv.tel_dof]))):
                ids = np.asarray(ids, dtype=np.int64)
                if not len(ids):
                    continue
                dofs = np.asarray(dofs, dtype=np.int64).reshape(-1, 4)
                rows = self.row_map[np.repeat(dofs, 4, axis=1).reshape(-1)]
                ok = rows >= 0
                parts.append((rows[ok], np.tile(dofs, (1, 4)).reshape(-1)[ok],
                              np.full(int(ok.sum()), kind, dtype=np.uint8),
                              np.repeat(ids, 16)[ok], np.tile(np.arange(16), len(ids))[ok],
                              np.repeat(2.0 * inv_u2 * mv.V[ids], 16)[ok]))
            self._jsheet = SheetMatrix(self.nT, *(np.concatenate([p[i] for p in parts])
                                                  for i in range(6)))
        return self._jsheet

    def linearize(self, phi, sides, C, w, dir_dofs):
        """Device Jacobian operator at phi (Dirichlet rows as identity): the
        plain part (node-gather Terms 1+2, plain Term 3), the sheet part in
        its fixed pattern and the special Term 3, all filled on the device;
        the aux block is copied to the host for its sparse LU."""
        from ..wake.sheet import SpecialTerm3

        mv = self.mv
        op, st = mv.op, mv.op.store
        nm, nT = self.nm, self.nT
        fm = np.ones(nT, dtype=np.uint8)
        fm[dir_dofs] = 0
        fm_main = to_dev(fm[:nm], torch.uint8)
        fm_dev = to_dev(fm, torch.uint8)
        if self.nop is None:
            self.nop = NewtonOperator(st, fm_main)
        else:
            self.nop.free_mask = fm_main
        up_s, lo_s = sides
        pm = self.pp_dev
        plain = self.is_plain_dev.double()
        rt = torch.where(pm, up_s["rho_tilde"], lo_s["rho_tilde"]) * plain
        se = torch.where(pm, up_s["s_e"], lo_s["s_e"]) * plain
        lim = torch.where(pm, up_s["lim"], lo_s["lim"])
        grad = up_s["grad"]          # plain elements: both sides = main field
        inv_u2 = 1.0 / (self.u_inf ** 2)
        op.jacobian_dev(rt, grad, se, lim, inv_u2, self.nop.data)
        self.nop.finalize()
        up = torch.where(pm, up_s["upstream"], lo_s["upstream"])
        su = torch.where(pm, up_s["s_u"], lo_s["s_u"]) * plain
        su = su * self.is_plain_dev[up.long()].double()
        self.nop.build_term3(grad, up, su, lim, inv_u2)
        self.t3_inputs = (grad, up, su * lim[up.long()].double(), inv_u2)
        # sheet part: D_keep C + exact Kutta + Term 2 of the special copies
        J = self._jacobian_sheet()
        J.fill(wa=up_s["rho_tilde"], wb=lo_s["rho_tilde"],
               kv=self._kutta_jacobian(phi)[2],
               sa=up_s["s_e"] * up_s["lim"].double(), sb=lo_s["s_e"] * lo_s["lim"].double(),
               ga=up_s["grad"], gb=lo_s["grad"], B=st.B, fmask=fm_dev)
        if self._t3s is None:
            self._t3s = SpecialTerm3(mv, self.row_map)
        self._t3s.build(sides, inv_u2, fmask=fm_dev)
        self._jc_host = None
        self.aux_lu = spla.splu(J.block_csc(nm, nm)) if nT > nm else None
        self._t_aux = empty(max(nT - nm, 1))
        self._schur = None
        # AMG of the single-valued Picard matrix at the current densities
        if self._amg is None or self._amg_age >= 3:
            A_sv = empty(op.nnz)
            op.matrix_dev(w, A_sv)
            mask_dirichlet(st, fm_main, A_sv)
            self._amg = amg_from_device(st, A_sv, "f64")
            self._amg_age = 0
        self._amg_age += 1
        self.free = fm.astype(bool)

    @property
    def Jc_host(self) -> sp.csr_matrix:
        """The sheet part of J plus the special Term 3 as one host CSR (for
        the Schur mode and the transposed operator of the adjoint)."""
        if self._jc_host is None:
            tr, tc, tv = self._t3s.to_coo()
            J = self._jsheet.to_scipy()
            T = sp.coo_matrix((tv, (tr, tc)), shape=(self.nT, self.nT)).tocsr()
            self._jc_host = (J + T).tocsr()
        return self._jc_host

    def matvec(self, x, y):
        nm = self.nm
        self.nop.matvec(x[:nm], y[:nm])
        if self.nT > nm:
            y[nm:].zero_()
        self._jsheet.add_matvec(x, y)
        self._t3s.add_matvec(x, y)
        return y

    # ---------------------------------------------------- linear step
    def _schur_parts(self):
        if self._schur is None:
            nm = self.nm
            Jc = self.Jc_host
            self._schur = (DeviceCSR(Jc[:nm, :nm].tocsr()), DeviceCSR(Jc[:nm, nm:].tocsr()),
                           empty(nm), DeviceCSR(Jc[nm:, :nm].tocsr()))
        return self._schur

    def schur_matvec(self, x, y):
        """K x = J_mm x - J_ma J_aa^{-1} J_am x on main-length vectors
        (pyfp3d/solve/schur_ls.py:95-103)."""
        nm = self.nm
        cmm, cma, t, cam = self._schur_parts()
        self.nop.matvec(x, y)
        _native.call("fp3d_amg_prolong_add", cmm.n, cmm.nnz, ptr(cmm.indptr),
                     ptr(cmm.indices), ptr(cmm.data), 0, ptr(x), ptr(y),
                     stream_handle())
        if self.nT > nm:
            C = cam
            _native.call("fp3d_csr_spmv", C.n, C.nnz, ptr(C.indptr), ptr(C.indices),
                         ptr(C.data), ptr(x), ptr(self._t_aux), stream_handle())
            v = to_dev(self.aux_lu.solve(to_host(self._t_aux)))
            _native.call("fp3d_csr_spmv", cma.n, cma.nnz, ptr(cma.indptr),
                         ptr(cma.indices), ptr(cma.data), ptr(v), ptr(t),
                         stream_handle())
            y.sub_(t)
        return y

    def solve_step(self, rhs, rtol, restart, maxiter, mode="block"):
        """(d host (n_total,), info, n_inner) for J d = rhs.  mode 'block':
        GMRES on the whole system with the block preconditioner; 'schur':
        GMRES on the aux-eliminated main system, aux by back-substitution."""
        nm = self.nm
        if mode == "block":
            d, info, n_it = gmres_dev(self.matvec, to_dev(rhs), self.psolve,
                                      rtol=rtol, restart=restart, maxiter=maxiter)
            return to_host(d), info, n_it
        if mode != "schur":
            raise ValueError(f"linear mode {mode!r} unknown")
        b_a = rhs[nm:]
        r = rhs[:nm].copy()
        cmm, cma, t, _ = self._schur_parts()
        if self.nT > nm:
            v = to_dev(self.aux_lu.solve(b_a))
            _native.call("fp3d_csr_spmv", cma.n, cma.nnz, ptr(cma.indptr),
                         ptr(cma.indices), ptr(cma.data), ptr(v), ptr(t),
                         stream_handle())
            r -= to_host(t)

        def prec(a, b):
            return self._amg.apply(a.contiguous(), b)
        x_m, info, n_it = gmres_dev(self.schur_matvec, to_dev(r), prec, rtol=rtol,
                                    restart=restart, maxiter=maxiter)
        x_m = to_host(x_m)
        x_a = np.zeros(self.nT - nm)
        if self.nT > nm:
            x_a = self.aux_lu.solve(b_a - self.Jc_host[nm:, :nm] @ x_m)
        return np.concatenate([x_m, x_a]), info, n_it

    def psolve(self, r, z):
        nm = self.nm
        self._amg.apply(r[:nm].contiguous(), z[:nm])
        if self.nT > nm:
            # the aux rows of J (only the sheet part has aux rows) applied to
            # [z_main; 0]: the (aux, main) block
            z[nm:].zero_()
            self._jsheet.rows_matvec(nm, z, self._t_aux)
            rhs = to_host(r[nm:]) - to_host(self._t_aux)
            z[nm:].copy_(to_dev(self.aux_lu.solve(rhs)))
        return z


def solve_multivalued_newton(mvop, mesh, m_inf: float, alpha_deg: float = 1.25,
                             u_inf: float = 1.0, gamma_air: float = 1.4,
                             upwind_c: float = 1.5, m_crit: float = 0.95,
                             m_cap: float = 3.0, rho_floor: float = 0.05,
                             farfield: str = "vortex", lift_axis: int = 1,
                             vortex_center=(0.25, 0.0),
                             phi_init: Optional[np.ndarray] = None,
                             gamma_init: float = 0.0, n_seed: int = 40,
                             seed_damping_theta: Optional[float] = 0.2,
                             seed_omega_rho: float = 0.5, seed_amg_rebuild_every: int = 5,
                             n_newton_max: int = 40,
                             tol_residual: float = 1e-10, lam_min: float = 0.05,
                             freeze: bool = True, freeze_tol: Optional[float] = None,
                             freeze_refresh_max: int = 8,
                             freeze_max_reverts: int = 3, stall_window: int = 8,
                             linear_rtol: float = 1e-10, gmres_restart: int = 60,
                             gmres_maxiter: int = 50, linear: str = "block",
                             external_rhs: Optional[np.ndarray] = None,
                             verbose: bool = False) -> Dict[str, object]:
    """Level-set lifting Newton with a Picard seed, safety line search on
    ||R_free||_inf (halving down to lam_min, best-of on failure) and the
    freeze / honesty / refresh logic of the reference driver; the far-field
    vortex takes the current circulation at every step.  linear='block'
    solves each step on the whole extended system (AMG main block + exact
    aux block), 'schur' on the aux-eliminated main system.  external_rhs
    (n_total): a boundary load (R = E phi - f), e.g. a Neumann outflow.  Returns the
    reference's result keys (phi_ext, phi, gamma, cl_kj, te_jump, converged,
    residual_history, n_newton, n_limited, n_floored, nu_max, mach2_max,
    n_gmres_total, n_gmres_stalled, accept_reason, froze, n_freeze_refresh,
    n_freeze_reverts, residual_unfrozen, n_assignment_stale,
    assignment_cycle, timings)."""
    if farfield not in ("vortex", "freestream"):
        raise NotImplementedError(f"farfield={farfield!r}: 'vortex' or 'freestream'")
    t_wall0 = time.perf_counter()
    timings = new_timings()
    t0 = time.perf_counter()
    seed = solve_multivalued_lifting(
        mvop, mesh, m_inf, alpha_deg=alpha_deg, u_inf=u_inf, gamma_air=gamma_air,
        upwind_c=upwind_c, m_crit=m_crit, damping_theta=seed_damping_theta,
        omega_rho=seed_omega_rho, n_outer_max=n_seed, gamma_init=gamma_init,
        phi_init=phi_init, vortex_center=vortex_center, farfield=farfield,
        lift_axis=lift_axis, external_rhs=external_rhs,
        amg_rebuild_every=seed_amg_rebuild_every)
    phi = seed["phi_ext"].copy()
    gamma = float(seed["gamma"])
    timings["seed"] = time.perf_counter() - t0
    beta = float(np.sqrt(1.0 - m_inf ** 2))
    nodes = np.asarray(mesh.nodes, dtype=np.float64)
    with phase(timings, "assembly"):
        sysm = LSNewtonSystem(mvop, m_inf, upwind_c, m_crit, gamma_air, u_inf,
                              m_cap, rho_floor, external_rhs)
    residual_history = []
    converged = False
    n_gmres_total = n_gmres_stalled = 0
    frozen = None
    freeze_point = None
    freeze_cooldown = 0
    n_freeze_refresh = n_freeze_reverts = n_assignment_stale = 0
    residual_unfrozen = None
    prev_live = None
    assignment_cycle = False
    accept_reason = None
    r_best = np.inf
    armed = bool(freeze) and freeze_tol is not None
    ff = None
    for it in range(n_newton_max):
        with phase(timings, "assembly", sync=False):
            ff, vals = _farfield_main(nodes, mesh.boundary_faces["farfield"], alpha_deg,
                                      gamma, u_inf, vortex_center, beta, farfield,
                                      lift_axis)
            phi[ff] = vals
            free = np.ones(mvop.n_total, dtype=bool)
            free[ff] = False
        with phase(timings, "residual"):
            R, sides, C, w = sysm.residual(phi, frozen, sens=True)
        res = float(np.abs(R[free]).max())
        n_lim, n_flr = mvop.n_limited, mvop.n_floored
        residual_history.append(res)
        if verbose:
            print(f"    ls-newton {it}: |R|={res:.3e} gamma={gamma:.6f} "
                  f"lim/flr={n_lim}/{n_flr}" + (" [frozen]" if frozen else ""),
                  flush=True)
        if frozen is not None and freeze_point is not None and (
                res > 10.0 * freeze_point[1] or n_lim + n_flr > 0):
            phi = freeze_point[0].copy()
            frozen = None
            freeze_point = None
            freeze_cooldown = 5
            n_freeze_reverts += 1
            r_best = np.inf
            if n_freeze_reverts >= freeze_max_reverts:
                armed = False
            continue
        if freeze_cooldown > 0:
            freeze_cooldown -= 1
        r_best = min(r_best, res)
        if res > 100.0 * r_best and res > 1e-3:
            break
        live_stalled = False
        if len(residual_history) > stall_window and res < 1e-3:
            live_stalled = (min(residual_history[-stall_window:])
                            > 0.9 * min(residual_history[:-stall_window]))
        if n_lim == 0 and n_flr == 0 and res < tol_residual:
            if frozen is None:
                converged, accept_reason = True, "tol"
                break
            with phase(timings, "residual"):
                R_l, _, _, _ = sysm.residual(phi, None)
                residual_unfrozen = float(np.abs(R_l[free]).max())
                fz_new = sysm.freeze(phi)
            n_assignment_stale = int(sum(
                int((ua != ub).sum().item()) + int(((ba != bb) & (ua == ub)).sum().item())
                for (ua, ba), (ub, bb) in zip(frozen, fz_new)))
            if residual_unfrozen < tol_residual and mvop.n_limited == 0 \
                    and mvop.n_floored == 0:
                converged, accept_reason = True, "tol"
                break
            if prev_live is not None and residual_unfrozen > 0.5 * prev_live:
                converged, accept_reason, assignment_cycle = True, "assignment_cycle", True
                break
            if n_freeze_refresh >= freeze_refresh_max:
                converged, accept_reason = True, "refresh_budget"
                break
            prev_live = residual_unfrozen
            n_freeze_refresh += 1
            frozen = fz_new
            freeze_point = (phi.copy(), residual_unfrozen)
            r_best = np.inf
            continue
        if (armed and frozen is None and freeze_cooldown == 0 and m_inf > 0.0
                and n_lim + n_flr == 0 and (res < freeze_tol or live_stalled)):
            with phase(timings, "residual"):
                frozen = sysm.freeze(phi)
                freeze_point = (phi.copy(), res)
                r_best = np.inf
                R, sides, C, w = sysm.residual(phi, frozen, sens=True)
        with phase(timings, "assembly"):
            sysm.linearize(phi, sides, C, w, ff)
        with phase(timings, "linsolve"):
            rhs = np.where(free, -R, 0.0)
            d, info, n_it = sysm.solve_step(rhs, linear_rtol, gmres_restart,
                                            gmres_maxiter, linear)
            n_gmres_total += n_it
            n_gmres_stalled += int(info != 0)
        best = None
        lam = 1.0
        with phase(timings, "residual"):
            while True:
                trial = phi + lam * np.where(free, d, 0.0)
                R_t, _, _, _ = sysm.residual(trial, frozen)
                m_t = float(np.abs(R_t[free]).max())
                if np.isfinite(m_t) and (best is None or m_t < best[1]):
                    best = (trial, m_t)
                if np.isfinite(m_t) and m_t < res:
                    break
                lam *= 0.5
                if lam < lam_min:
                    break
        if best is None:
            raise RuntimeError("LS Newton line search: every trial state was non-finite")
        phi = best[0]
        with phase(timings, "kutta"):
            gamma = float(np.mean(mvop.te_jump(phi)))
    with phase(timings, "residual"):
        sysm.residual(phi, None)
        mach2_max = mvop.own_side_mach2_max(phi, m_inf, gamma_air, u_inf)
    finalize(timings, time.perf_counter() - t_wall0)
    return {
        "phi_ext": phi, "phi": mvop.main_potential(phi), "gamma": gamma,
        "cl_kj": 2.0 * gamma / u_inf, "te_jump": mvop.te_jump(phi),
        "converged": converged, "residual_history": residual_history,
        "n_newton": len(residual_history), "n_limited": mvop.n_limited,
        "n_floored": mvop.n_floored, "nu_max": mvop.nu_max,
        "mach2_max": mach2_max, "farfield": farfield,
        "n_gmres_total": n_gmres_total, "n_gmres_stalled": n_gmres_stalled,
        "accept_reason": accept_reason, "froze": frozen is not None,
        "n_freeze_refresh": n_freeze_refresh, "n_freeze_reverts": n_freeze_reverts,
        "residual_unfrozen": residual_unfrozen,
        "n_assignment_stale": n_assignment_stale,
        "assignment_cycle": assignment_cycle, "timings": timings,
        "seed": {"gamma": seed["gamma"], "n_outer": seed["n_outer"]},
    }


def solve_multivalued_newton_transonic(mvop, mesh, m_target: float,
                                       alpha_deg: float = 1.25, m_start: float = 0.60,
                                       dm: float = 0.05, dm_min: float = 0.01,
                                       freeze_tol: Optional[float] = 1e-6,
                                       upwind_c: float = 1.5, upwind_c_post=None,
                                       n_seed: int = 40, n_newton_max: int = 60,
                                       phi_init: Optional[np.ndarray] = None,
                                       gamma_init: float = 0.0, verbose: bool = False,
                                       **newton_kw) -> Dict[str, object]:
    """Mach ramp of the level-set Newton (pyfp3d/solve/newton_ls.py:1021-1192):
    upward schedule, each level warm-started from the last converged one
    (the first from a Picard seed); a level that fails halves the Mach step
    and retries below it (down to dm_min); optional decreasing upwind_c_post
    stages re-solved at the target.  Returns the last solve's dict plus
    `levels`, `mach_schedule`, `target_reached`, `m_last_converged`,
    `m_final`, `timings_total`, `wall_s`."""
    from .continuation import mach_schedule
    from .timing import sum_timings

    if m_target < m_start:
        raise ValueError(f"m_target={m_target} < m_start={m_start}: the ramp only goes up")
    sched = [float(m) for m in mach_schedule(m_target, m_start=m_start, dm=dm)]
    t0 = time.perf_counter()
    records = []
    good = None
    res = None

    def level(m, c, kw, tag):
        tl = time.perf_counter()
        r = solve_multivalued_newton(mvop, mesh, m, alpha_deg=alpha_deg,
                                     upwind_c=c, n_newton_max=n_newton_max,
                                     freeze_tol=freeze_tol, verbose=verbose, **kw)
        records.append({"m_inf": m, "tag": tag, "upwind_c": c,
                        "gamma": float(r["gamma"]), "cl_kj": float(r["cl_kj"]),
                        "mach_max": float(np.sqrt(r["mach2_max"])),
                        "n_newton": int(r["n_newton"]),
                        "converged": bool(r["converged"]),
                        "accept_reason": r["accept_reason"],
                        "residual_norm": float(r["residual_history"][-1]),
                        "n_limited": int(r["n_limited"]),
                        "n_floored": int(r["n_floored"]),
                        "froze": bool(r["froze"]),
                        "n_freeze_refresh": int(r["n_freeze_refresh"]),
                        "n_freeze_reverts": int(r["n_freeze_reverts"]),
                        "n_lin_iters": int(r["n_gmres_total"]),
                        "residual_history": list(r["residual_history"]),
                        "timings": r["timings"],
                        "wall_s": time.perf_counter() - tl})
        if verbose:
            print(f"  [ramp] M={m:.4f} C={c} conv={r['converged']} "
                  f"n={r['n_newton']} gamma={r['gamma']:.6f}", flush=True)
        return r

    i = 0
    while i < len(sched):
        m = sched[i]
        kw = dict(newton_kw)
        if good is None:
            kw.update(phi_init=phi_init, gamma_init=gamma_init, n_seed=n_seed)
        else:
            kw.update(phi_init=good[0], gamma_init=good[1], n_seed=0)
        res = level(m, upwind_c, kw, "strict")
        if res["converged"]:
            good = (res["phi_ext"], res["gamma"])
            i += 1
            continue
        m_prev = m_start if i == 0 else sched[i - 1]
        half = 0.5 * (m - m_prev)
        if i == 0 or half < dm_min:
            break
        sched.insert(i, m_prev + half)
    target_reached = bool(good is not None and i == len(sched)
                          and abs(sched[-1] - m_target) < 1e-12)
    if target_reached and upwind_c_post:
        for c in upwind_c_post:
            kw = dict(newton_kw)
            kw.update(phi_init=good[0], gamma_init=good[1], n_seed=0)
            r = level(float(m_target), float(c), kw, f"post_c{c}")
            if not r["converged"]:
                break
            res, good = r, (r["phi_ext"], r["gamma"])
    out = dict(res)
    out["levels"] = records
    out["mach_schedule"] = sched
    out["target_reached"] = target_reached
    out["m_last_converged"] = float(sched[i - 1]) if good is not None and i > 0 else None
    out["m_final"] = float(sched[min(i, len(sched) - 1)])
    out["timings_total"] = sum_timings([lv["timings"] for lv in records])
    out["wall_s"] = time.perf_counter() - t0
    return out

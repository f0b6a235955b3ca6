"""Multivalued (extended-DOF) operator of a level-set wake
(pyfp3d/wake/multivalued.py:43-907 and pyfp3d/kernels/cut_assembly.py,
restated).

The extended system has n_total = n_main + n_ext unknowns: the main DOF of
a node holds the potential on the node's own side of the sheet, the aux
DOF of a cut-element node the other side's.  Its mass-conservation part is

  * plain elements (neither cut nor in the below-TE fan): the ordinary
    single-valued P1 assembly on main DOFs, weighted by the element density
    of the main field (a mixed-side plain element reads the main field too,
    pyfp3d/wake/multivalued.py:543-628);
  * cut elements: the same element matrix twice, on dofs_upper with the
    upper density and on dofs_lower with the lower density;
  * below-TE fan elements: once, on main DOFs with the TE node's aux DOF,
    lower density.

Design: the plain part is the device Picard-matrix assembly
(PicardOperator.matrix_dev) with the special (cut + fan) elements given a
zero weight; everything that involves a special element or an aux row is
the sheet part C (O(sheet area) entries), whose extended CSR pattern the
cut census fixes: it is built once (`sheet`) and its values are filled on
the device per assembly (`sheet_fill`, up3d_amd/wake/sheet.py).  The
Krylov operator is y = [A_plain x_main ; 0] + C x on the device.
`correction` is the same C assembled on the host as a COO matrix: the
"continuity" weld closure and small-case checks use it.

Closures (aux rows):
  * "continuity": aux_k - main_j = 0 (the weld; the extended solve then
    equals the single-valued one), cut-element entries across the sheet
    redirected to aux columns;
  * "wake_ls": the non-TE aux rows carry the wake least-squares condition
    V_e [(n.B_a)(n.B_c) + (u.B_a)(u.B_c)] (delta_c) with delta_c =
    x[up_c] - x[lo_c]; with te_kutta="pressure" the TE aux rows carry the
    linearised pressure-equality Kutta s.(q_u - q_l) = 0 (s the frozen mean
    q_u + q_l over the wall-adjacent TE control volumes) and their lower
    mass conservation is re-routed onto the TE main row; te_kutta="mass"
    keeps lower-side mass conservation there.
"""

from __future__ import annotations

from typing import Optional

import numpy as np
import scipy.sparse as sp
import torch

from .. import _native
from ..device import empty, ptr, stream_handle, to_dev, to_host
from ..kernels.jacobian import PicardOperator
from ..physics.isentropic import q2_at_mach


def _elem_matrices(B, V, w):
    """(m, 4, 4) w_e V_e (B_a . B_b)."""
    return np.einsum("e,ead,ebd->eab", V * w, B, B)


def isentropic_density(q2, m_inf, gamma_air=1.4):
    """rho(q2) = [1 + (g-1)/2 M^2 (1 - q2)]^(1/(g-1)), base floored at 1e-10
    (pyfp3d/physics/isentropic.py:29-58); host form for the O(sheet)
    special elements."""
    base = 1.0 + 0.5 * (gamma_air - 1.0) * m_inf * m_inf * (1.0 - q2)
    return np.maximum(base, 1e-10) ** (1.0 / (gamma_air - 1.0))


class MultivaluedOperator:
    def __init__(self, nodes, elements, cm, levelset=None):
        self.op = PicardOperator(nodes, elements)
        st = self.op.store
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.el = np.asarray(st.elements_host, dtype=np.int64)
        self.B = st.B_host.reshape(-1, 4, 3)
        self.V = st.V_host
        self.cm = cm
        self.levelset = levelset
        self.n_main = cm.n_main
        self.n_ext = cm.n_ext_dofs
        self.n_total = cm.n_total_dofs
        if st.n_nodes != self.n_main:
            raise ValueError("CutElementMap was built on a different mesh")
        nt = len(self.el)
        self.is_cut = np.zeros(nt, dtype=bool)
        self.is_cut[cm.cut_elems] = True
        self.is_tel = np.zeros(nt, dtype=bool)
        self.is_tel[cm.te_lower_elems] = True
        self.special = np.flatnonzero(self.is_cut | self.is_tel)
        self.is_te = np.zeros(self.n_main, dtype=bool)
        self.is_te[cm.te_nodes] = True
        self.cut_nodes = np.flatnonzero(cm.ext_dof_of_node >= 0)
        self.cut_aux = cm.ext_dof_of_node[self.cut_nodes]
        # lower-copy DOF vector of the fan elements (TE node -> its aux)
        tel = cm.te_lower_elems
        self.tel_dof = np.where(self.is_te[self.el[tel]],
                                cm.ext_dof_of_node[self.el[tel]], self.el[tel])
        # device weight mask: 0 on special elements
        keep = np.ones(nt)
        keep[self.special] = 0.0
        self.d_plain_mask = to_dev(keep)
        self._nonte_rows = cm.ext_dof_of_node[
            self.cut_nodes[~self.is_te[self.cut_nodes]]]
        self._ls = None
        if levelset is not None and len(cm.cut_elems):
            cen = self.nodes[self.el[cm.cut_elems]].mean(axis=1)
            self._ls = self._wake_ls_coo(levelset.surface_normals(cen),
                                         levelset.direction)
        self._te_cv = self._te_control_volumes()
        # per-side element sets (pyfp3d/wake/multivalued.py:582-606): plain
        # elements with any "+" node and cut elements carry an upper copy;
        # plain all-"-", cut and fan elements a lower copy
        plain = ~self.is_cut & ~self.is_tel
        self.plain_plus = plain & (cm.node_side[self.el].max(axis=1) == 1)
        self.in_upper = self.plain_plus | self.is_cut
        self.in_lower = (plain & ~self.plain_plus) | self.is_cut | self.is_tel
        self._graphs = None
        self.nu_max = 0.0
        self.n_nu_active = 0
        self.n_floored = 0
        self.n_limited = 0
        self._nu_active = None

    # ------------------------------------------------------- COO pieces
    def _wake_ls_coo(self, n_hat, u_hat):
        """pyfp3d/kernels/cut_assembly.py:230-318 wake_ls_coo."""
        cm = self.cm
        cut = cm.cut_elems
        B, V = self.B[cut], self.V[cut]
        u = np.asarray(u_hat, dtype=np.float64)
        u = u / np.linalg.norm(u)
        Bn = np.einsum("ead,ed->ea", B, n_hat)
        Bu = np.einsum("ead,d->ea", B, u)
        P = V[:, None, None] * (Bn[:, :, None] * Bn[:, None, :]
                                + Bu[:, :, None] * Bu[:, None, :])
        conn = self.el[cut]
        ok = ~self.is_te[conn]
        e_i, a_i = np.nonzero(ok)
        rows = np.repeat(cm.ext_dof_of_node[conn[e_i, a_i]], 4)
        k = P[e_i, a_i, :].reshape(-1)
        cu = cm.dofs_upper[e_i].reshape(-1)
        cl = cm.dofs_lower[e_i].reshape(-1)
        return (np.concatenate([rows, rows]), np.concatenate([cu, cl]),
                np.concatenate([k, -k]))

    def _scatter(self, ids, dofs, w):
        K = _elem_matrices(self.B[ids], self.V[ids], w)
        return (np.repeat(dofs, 4, axis=1).reshape(-1),
                np.tile(dofs, (1, 4)).reshape(-1), K.reshape(-1))

    def _te_control_volumes(self):
        """pyfp3d/wake/multivalued.py:178-248: per TE node, the wall-adjacent
        (>= 3 wall nodes) elements of its fan, split into the copies that
        reference the TE through its main DOF (upper) and its aux DOF
        (lower)."""
        cm = self.cm
        el = self.el
        cut_index = np.full(len(el), -1, dtype=np.int64)
        cut_index[cm.cut_elems] = np.arange(len(cm.cut_elems))
        tel_index = np.full(len(el), -1, dtype=np.int64)
        tel_index[cm.te_lower_elems] = np.arange(len(cm.te_lower_elems))
        if cm.wall_nodes is None:
            on_wall = np.ones(len(el), dtype=bool)
        else:
            w = np.zeros(self.n_main, dtype=bool)
            w[cm.wall_nodes] = True
            on_wall = w[el].sum(axis=1) >= 3
        cvs = []
        is_te = np.zeros(self.n_main, dtype=bool)
        is_te[cm.te_nodes] = True
        near = np.flatnonzero(is_te[el].any(axis=1) & on_wall)
        el_near = el[near]
        for t in cm.te_nodes:
            fan = near[(el_near == t).any(axis=1)]
            ue, ud, le, ld = [], [], [], []
            for e in fan:
                k = cut_index[e]
                if k >= 0:
                    ue.append(e); ud.append(cm.dofs_upper[k])
                    le.append(e); ld.append(cm.dofs_lower[k])
                elif tel_index[e] >= 0:
                    le.append(e); ld.append(self.tel_dof[tel_index[e]])
                else:
                    ue.append(e); ud.append(el[e])
            cvs.append({"upper_elems": np.array(ue, dtype=np.int64),
                        "upper_dofs": np.array(ud, dtype=np.int64).reshape(-1, 4),
                        "lower_elems": np.array(le, dtype=np.int64),
                        "lower_dofs": np.array(ld, dtype=np.int64).reshape(-1, 4)})
        return cvs

    def te_velocities(self, phi_ext):
        """(q_upper, q_lower) (n_te, 3): volume-weighted gradients over each
        TE node's upper / lower wall-adjacent control volume."""
        x = np.asarray(phi_ext, dtype=np.float64)
        qu = np.zeros((len(self._te_cv), 3))
        ql = np.zeros((len(self._te_cv), 3))
        for i, cv in enumerate(self._te_cv):
            for key, out in (("upper", qu), ("lower", ql)):
                e, d = cv[key + "_elems"], cv[key + "_dofs"]
                if len(e):
                    g = np.einsum("ead,ea->ed", self.B[e], x[d])
                    out[i] = (self.V[e][:, None] * g).sum(0) / self.V[e].sum()
        return qu, ql

    def _te_kutta_coo(self, phi_ext):
        """pyfp3d/kernels/cut_assembly.py:321-398 te_kutta_coo: s.(q_u - q_l)
        with s = q_u + q_l frozen at phi_ext, on the TE aux rows."""
        qu, ql = self.te_velocities(phi_ext)
        s = qu + ql
        rows, cols, data = [], [], []
        for i, t in enumerate(self.cm.te_nodes):
            r = self.cm.ext_dof_of_node[t]
            for key, sign in (("upper", 1.0), ("lower", -1.0)):
                cv = self._te_cv[i]
                e, d = cv[key + "_elems"], cv[key + "_dofs"]
                if not len(e):
                    continue
                w = self.V[e] / self.V[e].sum()
                c = sign * w[:, None] * np.einsum("ead,d->ea", self.B[e], s[i])
                rows.append(np.full(d.size, r, dtype=np.int64))
                cols.append(d.reshape(-1))
                data.append(c.reshape(-1))
        if not rows:
            z = np.empty(0, dtype=np.int64)
            return z, z, np.empty(0)
        return np.concatenate(rows), np.concatenate(cols), np.concatenate(data)

    # ------------------------------------------------------- fields
    def side_potentials(self, phi_ext):
        """(phi_upper, phi_lower) nodal fields (main except at cut nodes of
        the opposite side, which read their aux DOF)."""
        x = np.asarray(phi_ext, dtype=np.float64)
        up = x[: self.n_main].copy()
        lo = x[: self.n_main].copy()
        plus = self.cm.node_side[self.cut_nodes] == 1
        lo[self.cut_nodes[plus]] = x[self.cut_aux[plus]]
        up[self.cut_nodes[~plus]] = x[self.cut_aux[~plus]]
        return up, lo

    def node_jump(self, phi_ext, node_ids):
        x = np.asarray(phi_ext, dtype=np.float64)
        ids = np.asarray(node_ids, dtype=np.int64)
        aux = self.cm.ext_dof_of_node[ids]
        jump = np.zeros(len(ids))
        h = aux >= 0
        jump[h] = self.cm.node_side[ids[h]] * (x[ids[h]] - x[aux[h]])
        return jump

    def te_jump(self, phi_ext):
        return self.node_jump(phi_ext, self.cm.te_nodes)

    def main_potential(self, phi_ext):
        return np.asarray(phi_ext)[: self.n_main]

    def special_densities(self, phi_ext, m_inf, gamma_air=1.4, u_inf=1.0):
        """(rho_up_cut, rho_lo_cut, rho_lo_tel): isentropic densities of the
        cut elements' two copies and of the fan elements' lower copy."""
        x = np.asarray(phi_ext, dtype=np.float64)
        cm = self.cm

        def rho(ids, dofs):
            g = np.einsum("ead,ea->ed", self.B[ids], x[dofs])
            return isentropic_density((g * g).sum(1) / u_inf ** 2, m_inf,
                                      gamma_air)
        return (rho(cm.cut_elems, cm.dofs_upper), rho(cm.cut_elems, cm.dofs_lower),
                rho(cm.te_lower_elems, self.tel_dof))

    # ------------------------------------------------- per-side upwinding
    def _side_graphs(self):
        """Face graphs restricted to each side's element set (the wake is a
        slip line: the upwind walk never crosses it), as device
        (face_neighbors, face displacement) pairs
        (pyfp3d/wake/multivalued.py:608-640), built on the device from the
        store's face table: a face is kept when both its elements are on the
        side, and its displacement is the store's centroid difference (the
        same float64 subtraction), 0 where the face is dropped."""
        if self._graphs is None:
            st = self.op.store
            fn = st.fn.long()
            d = st.dcen.view(-1, 4, 3)
            out = []
            for keep in (self.in_upper, self.in_lower):
                kd = to_dev(keep.astype(np.uint8), torch.uint8).bool()
                ok = (fn >= 0) & kd[:, None] & kd[fn.clamp(min=0)]
                f = torch.where(ok, fn, torch.full_like(fn, -1)).to(torch.int32)
                disp = torch.where(ok[:, :, None], d, torch.zeros_like(d))
                out.append((f.contiguous(), disp.reshape(-1, 12).contiguous()))
            self._graphs = tuple(out)
        return self._graphs

    def side_state_dev(self, phi_ext, m_inf, upwind_c, m_crit,
                       gamma_air=1.4, u_inf=1.0, m_cap=3.0, rho_floor=0.05,
                       frozen=None, sensitivities=False):
        """Per-side upwinded state (pyfp3d/wake/multivalued.py:642-780):
        for each side a dict of device (n_tets,) arrays -- grad, q2l, rho,
        lim (1 = speed limiter inactive), upstream, rho_tilde and, with
        sensitivities=True, s_e / s_u -- from that side's field (main field
        on plain elements, the side copy on cut / fan elements), walked on
        the side's restricted face graph, or at a frozen per-side (upstream,
        branch) selection.  Own-side monitors refreshed."""
        op, st = self.op, self.op.store
        cm = self.cm
        nt = st.n_tets
        x = np.asarray(phi_ext, dtype=np.float64)
        q2cap = float(q2_at_mach(m_cap, m_inf, gamma_air))
        grad = empty((nt, 3))
        q2n = empty(nt)
        q2l = empty(nt)
        rho = empty(nt)
        lim = empty(nt, torch.uint8)
        op.state_dev(to_dev(x[: self.n_main]), u_inf ** 2, m_inf, gamma_air,
                     q2cap, grad, q2n, q2l, rho, lim)
        (fu, du), (fl, dl) = self._side_graphs()
        sides = ((cm.cut_elems, cm.dofs_upper, fu, du, self.in_upper),
                 (np.concatenate([cm.cut_elems, cm.te_lower_elems]),
                  np.concatenate([cm.dofs_lower, self.tel_dof]), fl, dl,
                  self.in_lower))
        out = []
        limited = floored = 0
        for k, (ids, dofs, f, d, keep) in enumerate(sides):
            g_s, q2l_s, rho_s, lim_s = grad.clone(), q2l.clone(), rho.clone(), lim.clone()
            if len(ids):
                g = np.einsum("ead,ea->ed", self.B[ids], x[dofs])
                q2 = (g * g).sum(1) / u_inf ** 2
                ql = np.minimum(q2, q2cap)
                idd = to_dev(ids, torch.int64)
                g_s[idd] = to_dev(g)
                q2l_s[idd] = to_dev(ql)
                rho_s[idd] = to_dev(isentropic_density(ql, m_inf, gamma_air))
                lim_s[idd] = to_dev((q2 <= q2cap).astype(np.uint8), torch.uint8)
            kd = to_dev(keep.astype(np.uint8), torch.uint8).bool()
            limited += int(((lim_s == 0) & kd).sum().item())
            nu = empty(nt)
            rt = empty(nt)
            if frozen is None:
                up = empty(nt, torch.int32)
                _native.call("fp3d_upwind_rho_tilde_d", nt, ptr(f), ptr(d),
                             ptr(st.nodes), ptr(st.el), ptr(g_s), ptr(q2l_s),
                             ptr(rho_s), 4, float(m_inf), float(m_crit),
                             float(upwind_c), float(gamma_air), float(rho_floor),
                             ptr(up), ptr(nu), ptr(rt), stream_handle())
                floored += int(((rt == rho_floor) & kd).sum().item())
            else:
                up, br = frozen[k]
                _native.call("fp3d_rho_tilde_frozen_sweep", nt, ptr(q2l_s),
                             ptr(rho_s), ptr(up), ptr(br), float(m_inf),
                             float(m_crit), float(upwind_c), float(gamma_air),
                             float(rho_floor), ptr(nu), ptr(rt), stream_handle())
                floored += int(((br == 3) & kd).sum().item())
            side = {"grad": g_s, "q2l": q2l_s, "rho": rho_s, "lim": lim_s,
                    "upstream": up, "rho_tilde": rt, "nu": nu, "keep": kd}
            if sensitivities:
                se, su = empty(nt), empty(nt)
                if frozen is None:
                    _native.call("fp3d_rho_tilde_sensitivities", nt, ptr(q2l_s),
                                 ptr(rho_s), ptr(up), float(m_inf), float(m_crit),
                                 float(upwind_c), float(gamma_air), float(rho_floor),
                                 ptr(se), ptr(su), stream_handle())
                else:
                    _native.call("fp3d_rho_tilde_frozen_sensitivities", nt,
                                 ptr(q2l_s), ptr(rho_s), ptr(up), ptr(frozen[k][1]),
                                 float(m_inf), float(m_crit), float(upwind_c),
                                 float(gamma_air), float(rho_floor), ptr(se),
                                 ptr(su), stream_handle())
                side["s_e"], side["s_u"] = se, su
            out.append(side)
        nus = [torch.where(sd["keep"], sd["nu"], torch.zeros_like(sd["nu"])) for sd in out]
        self.nu_max = float(max(n.max().item() for n in nus))
        self.n_nu_active = int(((nus[0] > 0.0).sum() + (nus[1] > 0.0).sum()).item())
        self.n_floored = floored
        self.n_limited = limited
        self._nu_active = (nus[0] > 0.0) | (nus[1] > 0.0)
        return out

    def freeze_side_state(self, phi_ext, m_inf, upwind_c, m_crit,
                          gamma_air=1.4, u_inf=1.0, m_cap=3.0, rho_floor=0.05):
        """Per-side (upstream, branch) at phi_ext (device), for a frozen-
        selection Newton finish (pyfp3d/wake/multivalued.py:781-796)."""
        sides = self.side_state_dev(phi_ext, m_inf, upwind_c, m_crit,
                                    gamma_air, u_inf, m_cap, rho_floor)
        out = []
        for sd in sides:
            br = empty(self.op.store.n_tets, torch.int8)
            _native.call("fp3d_classify_upwind_branches", self.op.store.n_tets,
                         ptr(sd["q2l"]), ptr(sd["rho"]), ptr(sd["upstream"]),
                         float(m_inf), float(m_crit), float(upwind_c),
                         float(gamma_air), float(rho_floor), ptr(br),
                         stream_handle())
            out.append((sd["upstream"].clone(), br))
        return tuple(out)

    def side_rho_tilde_dev(self, phi_ext, m_inf, upwind_c, m_crit,
                           gamma_air=1.4, u_inf=1.0, m_cap=3.0,
                           rho_floor=0.05):
        """(rho_tilde_upper, rho_tilde_lower) device arrays (live walk)."""
        sides = self.side_state_dev(phi_ext, m_inf, upwind_c, m_crit, gamma_air,
                                    u_inf, m_cap, rho_floor)
        return sides[0]["rho_tilde"], sides[1]["rho_tilde"]

    def read_dofs(self):
        """(upper, lower) (n_tets, 4) DOF vectors each side's field reads:
        main DOFs on plain elements, the side copy on cut elements, the fan's
        lower copy (pyfp3d/wake/multivalued.py:606-660)."""
        if getattr(self, "_read", None) is None:
            cm = self.cm
            ru = self.el.copy()
            ru[cm.cut_elems] = cm.dofs_upper
            rl = self.el.copy()
            rl[cm.cut_elems] = cm.dofs_lower
            rl[cm.te_lower_elems] = self.tel_dof
            self._read = (ru, rl)
        return self._read

    def nu_active_elements(self) -> np.ndarray:
        if self._nu_active is None:
            return np.zeros(0, dtype=np.int64)
        return np.flatnonzero(to_host(self._nu_active))

    def side_weights(self, rt_up_dev, rt_lo_dev):
        """(element weights on the device for the plain part / the AMG, the
        special-copy weights as the device (upper, lower) pair)."""
        pp = getattr(self, "_d_plain_plus", None)
        if pp is None:
            pp = self._d_plain_plus = to_dev(self.plain_plus.astype(np.uint8),
                                             torch.uint8).bool()
        w = torch.where(pp, rt_up_dev, rt_lo_dev)
        return w, (rt_up_dev, rt_lo_dev)

    def special_weights_host(self, special):
        """The host triple (cut upper, cut lower, fan lower) of a device
        (upper, lower) density pair (a host triple passes through)."""
        if special is not None and len(special) == 2 and isinstance(special[0], torch.Tensor):
            cm = self.cm
            ru, rl = to_host(special[0]), to_host(special[1])
            return ru[cm.cut_elems], rl[cm.cut_elems], rl[cm.te_lower_elems]
        return special

    def own_side_mach2_max(self, phi_ext, m_inf, gamma_air=1.4, u_inf=1.0):
        """Largest own-side element M^2 (main field on plain elements, both
        copies on cut elements, the lower copy on fan elements)."""
        op, st = self.op, self.op.store
        nt = st.n_tets
        x = np.asarray(phi_ext, dtype=np.float64)
        grad, q2 = op.velocities_dev(to_dev(x[: self.n_main]), empty((nt, 3)), empty(nt))
        m2 = empty(nt)
        q2n = q2 / u_inf ** 2
        _native.call("fp3d_mach_squared_field", nt, ptr(q2n),
                     float(m_inf), float(gamma_air), ptr(m2), stream_handle())
        mask = torch.ones(nt, dtype=torch.bool, device=m2.device)
        if len(self.special):
            mask[to_dev(self.special, torch.int64)] = False
        best = float(m2[mask].max().item()) if bool(mask.any()) else 0.0
        cm = self.cm
        for ids, dofs in ((cm.cut_elems, cm.dofs_upper), (cm.cut_elems, cm.dofs_lower),
                          (cm.te_lower_elems, self.tel_dof)):
            if len(ids):
                g = np.einsum("ead,ea->ed", self.B[ids], x[dofs])
                q = (g * g).sum(1) / u_inf ** 2
                den = 1.0 + 0.5 * (gamma_air - 1.0) * m_inf ** 2 * (1.0 - q)
                best = max(best, float((q * m_inf ** 2 / den).max()))
        return best

    # ------------------------------------------------------- assembly
    def plain_weight_dev(self, rho_main_dev: Optional[torch.Tensor]):
        """Device element weights of the plain part (special elements 0)."""
        if rho_main_dev is None:
            return self.d_plain_mask.clone()
        return rho_main_dev * self.d_plain_mask

    def correction(self, special_rho=None, closure: str = "wake_ls",
                   te_kutta: str = "pressure", phi_ext=None) -> sp.csr_matrix:
        """Host (n_total x n_total) correction C: special-element mass
        conservation, aux rows (closure) and the TE Kutta rows."""
        cm = self.cm
        nT = self.n_total
        if special_rho is None:
            r_uc = np.ones(len(cm.cut_elems))
            r_lc = np.ones(len(cm.cut_elems))
            r_lt = np.ones(len(cm.te_lower_elems))
        else:
            r_uc, r_lc, r_lt = self.special_weights_host(special_rho)
        parts = []
        if closure == "continuity":
            # single-valued special elements on main DOFs with the upper
            # weight, then cross-sheet entries of cut elements moved to aux
            # columns (pyfp3d/kernels/cut_assembly.py:56-116), aux weld rows
            conn_c = self.el[cm.cut_elems]
            parts.append(self._scatter(cm.cut_elems, conn_c, r_uc))
            parts.append(self._scatter(cm.te_lower_elems, self.el[cm.te_lower_elems], r_lt))
            K = _elem_matrices(self.B[cm.cut_elems], self.V[cm.cut_elems], r_uc)
            side = cm.node_side[conn_c]
            aux = cm.ext_dof_of_node[conn_c]
            e_i, a_i, b_i = np.nonzero(side[:, :, None] != side[:, None, :])
            k = K[e_i, a_i, b_i]
            row = conn_c[e_i, a_i]
            parts.append((np.concatenate([row, row]),
                          np.concatenate([conn_c[e_i, b_i], aux[e_i, b_i]]),
                          np.concatenate([-k, k])))
            ones = np.ones(len(self.cut_aux))
            parts.append((np.concatenate([self.cut_aux, self.cut_aux]),
                          np.concatenate([self.cut_aux, self.cut_nodes]),
                          np.concatenate([ones, -ones])))
        elif closure == "wake_ls":
            if self._ls is None:
                raise ValueError("wake_ls closure needs a level set and cut elements")
            m_parts = [self._scatter(cm.cut_elems, cm.dofs_upper, r_uc),
                       self._scatter(cm.cut_elems, cm.dofs_lower, r_lc),
                       self._scatter(cm.te_lower_elems, self.tel_dof, r_lt)]
            r = np.concatenate([p[0] for p in m_parts])
            c = np.concatenate([p[1] for p in m_parts])
            d = np.concatenate([p[2] for p in m_parts])
            drop = np.zeros(nT, dtype=bool)
            drop[self._nonte_rows] = True
            keep = ~drop[r]
            r, c, d = r[keep], c[keep], d[keep]
            if te_kutta == "pressure":
                if phi_ext is None:
                    raise ValueError("te_kutta='pressure' needs phi_ext")
                reroute = np.full(nT, -1, dtype=np.int64)
                reroute[cm.ext_dof_of_node[cm.te_nodes]] = cm.te_nodes
                hit = reroute[r] >= 0
                r = r.copy()
                r[hit] = reroute[r[hit]]
                parts.append((r, c, d))
                parts.append(self._ls)
                parts.append(self._te_kutta_coo(phi_ext))
            elif te_kutta == "mass":
                parts.append((r, c, d))
                parts.append(self._ls)
            else:
                raise ValueError(f"te_kutta={te_kutta!r} unknown")
        else:
            raise ValueError(f"closure={closure!r} unknown")
        rows = np.concatenate([p[0] for p in parts])
        cols = np.concatenate([p[1] for p in parts])
        data = np.concatenate([p[2] for p in parts])
        C = sp.coo_matrix((data, (rows, cols)), shape=(nT, nT)).tocsr()
        C.sum_duplicates()
        return C

    # ------------------------------------------------ device sheet terms
    def special_copy_contributions(self):
        """Contributions (rows, cols, kind, src, ab, coef) of the special
        copies' mass conservation V_e B_a.B_b (times the copy's density):
        both copies of each cut element, the fan's lower copy; rows and
        columns are the copy's DOFs, before any row treatment."""
        from .sheet import LOWER, UPPER

        cm = self.cm
        parts = []
        for ids, dofs, kind in ((cm.cut_elems, cm.dofs_upper, UPPER),
                                (cm.cut_elems, cm.dofs_lower, LOWER),
                                (cm.te_lower_elems, self.tel_dof, LOWER)):
            ids = np.asarray(ids, dtype=np.int64)
            if not len(ids):
                continue
            dofs = np.asarray(dofs, dtype=np.int64).reshape(-1, 4)
            K = _elem_matrices(self.B[ids], self.V[ids], np.ones(len(ids)))
            parts.append((np.repeat(dofs, 4, axis=1).reshape(-1),
                          np.tile(dofs, (1, 4)).reshape(-1),
                          np.full(16 * len(ids), kind, dtype=np.uint8),
                          np.repeat(ids, 16), np.tile(np.arange(16), len(ids)),
                          K.reshape(-1)))
        if not parts:
            z = np.empty(0, dtype=np.int64)
            return z, z, z.astype(np.uint8), z, z, np.empty(0)
        return tuple(np.concatenate([p[i] for p in parts]) for i in range(6))

    def sheet_contributions(self, te_kutta: str = "pressure"):
        """Contributions of the correction C of closure 'wake_ls' (the
        special copies with non-TE aux rows dropped and, for the pressure
        Kutta, TE aux rows re-routed to the TE main row; the wake
        least-squares rows; the TE Kutta rows, whose values come per call
        from kutta_values), in the order `correction` builds them."""
        from .sheet import HOSTVAL, STATIC

        if self._ls is None:
            raise ValueError("wake_ls closure needs a level set and cut elements")
        cm = self.cm
        nT = self.n_total
        r, c, k, s, ab, coef = self.special_copy_contributions()
        drop = np.zeros(nT, dtype=bool)
        drop[self._nonte_rows] = True
        keep = ~drop[r]
        r, c, k, s, ab, coef = (a[keep] for a in (r, c, k, s, ab, coef))
        parts = []
        if te_kutta == "pressure":
            reroute = np.full(nT, -1, dtype=np.int64)
            reroute[cm.ext_dof_of_node[cm.te_nodes]] = cm.te_nodes
            hit = reroute[r] >= 0
            r = r.copy()
            r[hit] = reroute[r[hit]]
        elif te_kutta != "mass":
            raise ValueError(f"te_kutta={te_kutta!r} unknown")
        parts.append((r, c, k, s, ab, coef))
        lr, lc, ld = self._ls
        z = np.zeros(len(lr), dtype=np.int64)
        parts.append((lr, lc, np.full(len(lr), STATIC, dtype=np.uint8), z, z, ld))
        if te_kutta == "pressure":
            kr, kc, _ = self._te_kutta_coo(np.zeros(nT))
            parts.append((kr, kc, np.full(len(kr), HOSTVAL, dtype=np.uint8),
                          np.arange(len(kr)), np.zeros(len(kr), dtype=np.int64),
                          np.ones(len(kr))))
        return tuple(np.concatenate([p[i] for p in parts]) for i in range(6))

    def sheet(self, te_kutta: str = "pressure", user: str = "picard"):
        """The device pattern of C (closure 'wake_ls'), built once per
        (te_kutta, user): each user (the Picard operator, the Newton
        residual) fills its own values."""
        from .sheet import SheetMatrix

        cache = self.__dict__.setdefault("_sheets", {})
        key = (te_kutta, user)
        if key not in cache:
            cache[key] = SheetMatrix(self.n_total, *self.sheet_contributions(te_kutta))
        return cache[key]

    def kutta_values(self, phi_ext) -> np.ndarray:
        """Values of the TE pressure-Kutta rows at phi_ext, in the order of
        their contributions."""
        return self._te_kutta_coo(phi_ext)[2]

    def special_weights_dev(self, special):
        """(wa, wb) device (n_tets,) arrays of the upper / lower copy
        densities: `special` is a device pair (rho_tilde upper, lower), or
        the host triple (cut upper, cut lower, fan lower) of special
        densities, or None (all ones)."""
        nt = self.op.store.n_tets
        if special is None:
            one = torch.ones(nt, dtype=torch.float64, device=self.d_plain_mask.device)
            return one, one
        if len(special) == 2 and isinstance(special[0], torch.Tensor):
            return special[0], special[1]
        cm = self.cm
        wa = torch.zeros(nt, dtype=torch.float64, device=self.d_plain_mask.device)
        wb = torch.zeros_like(wa)
        if len(cm.cut_elems):
            ic = to_dev(np.asarray(cm.cut_elems, dtype=np.int64), torch.int64)
            wa[ic] = to_dev(np.asarray(special[0], dtype=np.float64))
            wb[ic] = to_dev(np.asarray(special[1], dtype=np.float64))
        if len(cm.te_lower_elems):
            it = to_dev(np.asarray(cm.te_lower_elems, dtype=np.int64), torch.int64)
            wb[it] = to_dev(np.asarray(special[2], dtype=np.float64))
        return wa, wb

    def sheet_fill(self, special, te_kutta: str = "pressure", phi_ext=None,
                   fmask=None, user: str = "picard"):
        """C of closure 'wake_ls' filled on the device (its SheetMatrix)."""
        S = self.sheet(te_kutta, user)
        wa, wb = self.special_weights_dev(special)
        kv = None
        if te_kutta == "pressure":
            if phi_ext is None:
                raise ValueError("te_kutta='pressure' needs phi_ext")
            kv = self.kutta_values(phi_ext)
        S.fill(wa=wa, wb=wb, kv=kv, fmask=fmask)
        return S

    def assemble_matrix(self, rho_main=None, special_rho=None,
                        closure: str = "wake_ls", te_kutta: str = "pressure",
                        phi_ext=None) -> sp.csr_matrix:
        """The whole extended matrix as a host scipy CSR (plain part from the
        device assembly) -- for tests and small cases; the solver keeps the
        plain part on the device."""
        op = self.op
        w = self.plain_weight_dev(None if rho_main is None else to_dev(rho_main))
        vals = empty(op.nnz)
        op.matrix_dev(w, vals)
        A = sp.csr_matrix((to_host(vals), op.store.indices_host,
                           op.store.indptr_host), shape=(self.n_main, self.n_main))
        E = sp.bmat([[A, None], [None, sp.csr_matrix((self.n_ext, self.n_ext))]],
                    format="csr")
        return (E + self.correction(special_rho, closure, te_kutta, phi_ext)).tocsr()

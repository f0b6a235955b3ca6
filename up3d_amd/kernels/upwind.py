"""Artificial-density upwinding on the device (pyfp3d/kernels/upwind.py,
walk flux): multi-hop upstream walk u(e), the nu switch and rho_tilde,
branch classification and freeze, and the exact sensitivities
(s_e, s_u) = (d rho_tilde_e / d q2_e, d rho_tilde_e / d q2_u(e)).

The walk is evaluated in uncontracted IEEE arithmetic in the reference's
order (strict `<`, first face wins a tie), so u(e) agrees exactly with an
IEEE restatement of the reference.
"""

from __future__ import annotations

import numpy as np
import torch

from .. import _native
from ..device import empty, ptr, stream_handle, to_host
from ..physics.isentropic import GAMMA
from ._io import Out, din
from .store import MeshStore


# ----------------------------------------------------- drop-in raw kernels
def upstream_elements(face_neighbors, centroids, nodes, elements, grad, out,
                      max_hops: int = 4) -> None:
    """u(e) into `out` (pyfp3d/kernels/upwind.py:68-125)."""
    fn = din(face_neighbors, torch.int32)
    cen = din(centroids, torch.float64)
    nd = din(nodes, torch.float64)
    el = din(elements, torch.int32)
    g = din(grad, torch.float64)
    o = Out(out, torch.int32)
    _native.call("fp3d_upstream_elements", el.shape[0], ptr(fn), ptr(cen),
                 ptr(nd), ptr(el), ptr(g), int(max_hops), ptr(o.t),
                 stream_handle())
    o.done()


def _sweep_in(q2, rho, upstream):
    return din(q2, torch.float64), din(rho, torch.float64), din(upstream, torch.int32)


def rho_tilde_sweep(q2, rho, upstream, m_inf, m_crit, C, gamma, rho_floor,
                    nu_out, rho_tilde_out) -> None:
    """(pyfp3d/kernels/upwind.py:456-500)"""
    q, r, u = _sweep_in(q2, rho, upstream)
    nu = Out(nu_out, torch.float64)
    rt = Out(rho_tilde_out, torch.float64)
    _native.call("fp3d_rho_tilde_sweep", q.numel(), ptr(q), ptr(r), ptr(u),
                 float(m_inf), float(m_crit), float(C), float(gamma),
                 float(rho_floor), ptr(nu.t), ptr(rt.t), stream_handle())
    nu.done()
    rt.done()


def classify_upwind_branches(q2, rho, upstream, m_inf, m_crit, C, gamma,
                             rho_floor, branch_out) -> None:
    """(pyfp3d/kernels/upwind.py:503-541)"""
    q, r, u = _sweep_in(q2, rho, upstream)
    b = Out(branch_out, torch.int8)
    _native.call("fp3d_classify_upwind_branches", q.numel(), ptr(q), ptr(r),
                 ptr(u), float(m_inf), float(m_crit), float(C), float(gamma),
                 float(rho_floor), ptr(b.t), stream_handle())
    b.done()


def rho_tilde_frozen_sweep(q2, rho, upstream, branch, m_inf, m_crit, C, gamma,
                           rho_floor, nu_out, rho_tilde_out) -> None:
    """(pyfp3d/kernels/upwind.py:544-593)"""
    q, r, u = _sweep_in(q2, rho, upstream)
    br = din(branch, torch.int8)
    nu = Out(nu_out, torch.float64)
    rt = Out(rho_tilde_out, torch.float64)
    _native.call("fp3d_rho_tilde_frozen_sweep", q.numel(), ptr(q), ptr(r),
                 ptr(u), ptr(br), float(m_inf), float(m_crit), float(C),
                 float(gamma), float(rho_floor), ptr(nu.t), ptr(rt.t),
                 stream_handle())
    nu.done()
    rt.done()


def rho_tilde_frozen_sensitivities_sweep(q2, rho, upstream, branch, m_inf,
                                         m_crit, C, gamma, rho_floor, se_out,
                                         su_out) -> None:
    """(pyfp3d/kernels/upwind.py:596-643)"""
    q, r, u = _sweep_in(q2, rho, upstream)
    br = din(branch, torch.int8)
    se = Out(se_out, torch.float64)
    su = Out(su_out, torch.float64)
    _native.call("fp3d_rho_tilde_frozen_sensitivities", q.numel(), ptr(q),
                 ptr(r), ptr(u), ptr(br), float(m_inf), float(m_crit),
                 float(C), float(gamma), float(rho_floor), ptr(se.t),
                 ptr(su.t), stream_handle())
    se.done()
    su.done()


def rho_tilde_sensitivities_sweep(q2, rho, upstream, m_inf, m_crit, C, gamma,
                                  rho_floor, se_out, su_out) -> None:
    """(pyfp3d/kernels/upwind.py:646-730)"""
    q, r, u = _sweep_in(q2, rho, upstream)
    se = Out(se_out, torch.float64)
    su = Out(su_out, torch.float64)
    _native.call("fp3d_rho_tilde_sensitivities", q.numel(), ptr(q), ptr(r),
                 ptr(u), float(m_inf), float(m_crit), float(C), float(gamma),
                 float(rho_floor), ptr(se.t), ptr(su.t), stream_handle())
    se.done()
    su.done()


# ------------------------------------------------------------ the operator
class UpwindOperator:
    """Per-mesh upwinding workspace (pyfp3d/kernels/upwind.py:733-920), walk
    mode.  Reference-compatible methods take numpy in the original element
    order; *_dev methods work on device tensors.

    weighted=True (the reference's blend / streamline-kernel fluxes) is not
    part of the Newton path (it runs weighted=False,
    pyfp3d/solve/newton.py:216-217) and raises NotImplementedError."""

    def __init__(self, nodes, elements, weighted: bool = False,
                 mode: str = "kernel", **_unused):
        if weighted:
            raise NotImplementedError(
                "weighted upwind fluxes are not provided; the Newton path uses "
                "the walk flux (weighted=False)")
        self.weighted = False
        self.mode = mode
        self.store = st = MeshStore.get(nodes, elements)
        n = st.n_tets
        self.face_neighbors = st.face_neighbors_host
        self.centroids = st.centroids_host
        self.d_up = empty(n, torch.int32)
        self.d_nu = empty(n)
        self.d_rt = empty(n)
        self.d_se = empty(n)
        self.d_su = empty(n)
        self.d_branch = empty(n, torch.int8)
        self._work = empty(_native.lib().fp3d_reduce_work_doubles())
        self._mon = empty(3)
        self.nu_max = 0.0
        self.n_supersonic = 0
        self.n_floored = 0

    # --------------------------------------------------------- device path
    def walk_dev(self, grad: torch.Tensor, out=None) -> torch.Tensor:
        st = self.store
        out = self.d_up if out is None else out
        _native.call("fp3d_upstream_elements_d", st.n_tets, ptr(st.fn),
                     ptr(st.dcen), ptr(st.nodes), ptr(st.el), ptr(grad), 4,
                     ptr(out), stream_handle())
        return out

    def rho_tilde_dev(self, grad, q2, rho, m_inf, C, m_crit, gamma=GAMMA,
                      rho_floor=0.05, monitors=True, up_out=None, rt_out=None,
                      se_out=None, su_out=None):
        """Fused walk + switch + rho_tilde; monitors (nu_max, n_supersonic,
        n_floored) refreshed by a deterministic device reduction.  With
        se_out / su_out the same sweep also writes the exact sensitivities
        of its selection (as sensitivities_at_dev on the walk's upstream)."""
        st = self.store
        up = self.d_up if up_out is None else up_out
        rt = self.d_rt if rt_out is None else rt_out
        if (se_out is None) != (su_out is None):
            raise ValueError("rho_tilde_dev: give both se_out and su_out or neither")
        if se_out is not None:
            _native.call("fp3d_upwind_rho_tilde_sens_d", st.n_tets, ptr(st.fn),
                         ptr(st.dcen), ptr(st.nodes), ptr(st.el), ptr(grad), ptr(q2),
                         ptr(rho), 4, float(m_inf), float(m_crit), float(C),
                         float(gamma), float(rho_floor), ptr(up), ptr(self.d_nu),
                         ptr(rt), ptr(se_out), ptr(su_out), stream_handle())
        else:
            _native.call("fp3d_upwind_rho_tilde_d", st.n_tets, ptr(st.fn),
                         ptr(st.dcen), ptr(st.nodes), ptr(st.el), ptr(grad), ptr(q2),
                         ptr(rho), 4, float(m_inf), float(m_crit), float(C),
                         float(gamma), float(rho_floor), ptr(up), ptr(self.d_nu),
                         ptr(rt), stream_handle())
        if monitors:
            self._monitors(self.d_nu, rt, rho_floor)
        return rt

    def sensitivities_at_dev(self, q2, rho, upstream, m_inf, C, m_crit,
                             gamma=GAMMA, rho_floor=0.05):
        """(s_e, s_u) for a given upstream selection (no walk)."""
        _native.call("fp3d_rho_tilde_sensitivities", self.store.n_tets, ptr(q2),
                     ptr(rho), ptr(upstream), float(m_inf), float(m_crit),
                     float(C), float(gamma), float(rho_floor), ptr(self.d_se),
                     ptr(self.d_su), stream_handle())
        return self.d_se, self.d_su

    def _monitors(self, nu, rt, rho_floor):
        _native.call("fp3d_upwind_monitors", self.store.n_tets, ptr(nu), ptr(rt),
                     float(rho_floor), ptr(self._work), ptr(self._mon),
                     stream_handle())
        m = to_host(self._mon)
        self.nu_max = float(m[0])
        self.n_supersonic = int(m[1])
        self.n_floored = int(m[2])

    def sensitivities_dev(self, grad, q2, rho, m_inf, C, m_crit, gamma=GAMMA,
                          rho_floor=0.05):
        """(s_e, s_u, upstream) at frozen selection (fresh walk on grad)."""
        self.walk_dev(grad)
        _native.call("fp3d_rho_tilde_sensitivities", self.store.n_tets, ptr(q2),
                     ptr(rho), ptr(self.d_up), float(m_inf), float(m_crit),
                     float(C), float(gamma), float(rho_floor), ptr(self.d_se),
                     ptr(self.d_su), stream_handle())
        return self.d_se, self.d_su, self.d_up

    def freeze_dev(self, grad, q2, rho, m_inf, C, m_crit, gamma=GAMMA,
                   rho_floor=0.05):
        """(upstream, branch) COPIES at the current state."""
        self.walk_dev(grad)
        _native.call("fp3d_classify_upwind_branches", self.store.n_tets, ptr(q2),
                     ptr(rho), ptr(self.d_up), float(m_inf), float(m_crit),
                     float(C), float(gamma), float(rho_floor),
                     ptr(self.d_branch), stream_handle())
        return self.d_up.clone(), self.d_branch.clone()

    def frozen_dev(self, q2, rho, upstream, branch, m_inf, C, m_crit,
                   gamma=GAMMA, rho_floor=0.05, monitors=True):
        _native.call("fp3d_rho_tilde_frozen_sweep", self.store.n_tets, ptr(q2),
                     ptr(rho), ptr(upstream), ptr(branch), float(m_inf),
                     float(m_crit), float(C), float(gamma), float(rho_floor),
                     ptr(self.d_nu), ptr(self.d_rt), stream_handle())
        if monitors:
            self._monitors(self.d_nu, self.d_rt, rho_floor)
            # the frozen sweep's floor count is the branch-3 count
            self.n_floored = int((branch == 3).sum().item())
        return self.d_rt

    def frozen_sensitivities_dev(self, q2, rho, upstream, branch, m_inf, C,
                                 m_crit, gamma=GAMMA, rho_floor=0.05):
        _native.call("fp3d_rho_tilde_frozen_sensitivities", self.store.n_tets,
                     ptr(q2), ptr(rho), ptr(upstream), ptr(branch),
                     float(m_inf), float(m_crit), float(C), float(gamma),
                     float(rho_floor), ptr(self.d_se), ptr(self.d_su),
                     stream_handle())
        return self.d_se, self.d_su

    # ---------------------------------------------- reference-compatible API
    def _in(self, a):
        return self.store.to_storage(a)

    def _out(self, t, host):
        t = self.store.to_original(t)
        return to_host(t) if host else t

    def rho_tilde(self, grad, q2, rho, m_inf, C, m_crit, gamma=GAMMA,
                  rho_floor=0.05):
        host = not isinstance(q2, torch.Tensor)
        rt = self.rho_tilde_dev(self._in(grad), self._in(q2), self._in(rho),
                                m_inf, C, m_crit, gamma, rho_floor)
        self.nu = self._out(self.d_nu, host)
        return self._out(rt, host)

    def rho_tilde_sensitivities(self, grad, q2, rho, m_inf, C, m_crit,
                                gamma=GAMMA, rho_floor=0.05):
        host = not isinstance(q2, torch.Tensor)
        se, su, up = self.sensitivities_dev(self._in(grad), self._in(q2),
                                            self._in(rho), m_inf, C, m_crit,
                                            gamma, rho_floor)
        upo = self.store.upstream_to_original(up)
        return (self._out(se, host), self._out(su, host),
                to_host(upo) if host else upo)

    def freeze_upwind_state(self, grad, q2, rho, m_inf, C, m_crit, gamma=GAMMA,
                            rho_floor=0.05):
        host = not isinstance(q2, torch.Tensor)
        up, br = self.freeze_dev(self._in(grad), self._in(q2), self._in(rho),
                                 m_inf, C, m_crit, gamma, rho_floor)
        upo = self.store.upstream_to_original(up)
        bro = self.store.to_original(br)
        return (to_host(upo), to_host(bro)) if host else (upo, bro)

    def rho_tilde_frozen(self, q2, rho, upstream, branch, m_inf, C, m_crit,
                         gamma=GAMMA, rho_floor=0.05):
        host = not isinstance(q2, torch.Tensor)
        up_s = self.store.upstream_to_storage(upstream)
        br_s = self.store.to_storage(branch, torch.int8)
        rt = self.frozen_dev(self._in(q2), self._in(rho), up_s, br_s, m_inf, C,
                             m_crit, gamma, rho_floor)
        self.nu = self._out(self.d_nu, host)
        return self._out(rt, host)

    def rho_tilde_frozen_sensitivities(self, q2, rho, upstream, branch, m_inf,
                                       C, m_crit, gamma=GAMMA, rho_floor=0.05):
        host = not isinstance(q2, torch.Tensor)
        up_s = self.store.upstream_to_storage(upstream)
        br_s = self.store.to_storage(branch, torch.int8)
        se, su = self.frozen_sensitivities_dev(self._in(q2), self._in(rho), up_s,
                                               br_s, m_inf, C, m_crit, gamma,
                                               rho_floor)
        return self._out(se, host), self._out(su, host)

"""Picard matrix and exact Newton Jacobian on the device
(pyfp3d/kernels/jacobian.py).

A_ij = sum_e rho_tilde_e V_e (B_i . B_j) and the Newton Terms 1+2 are
scattered colour by colour into the fixed CSR pattern through the
precomputed elem_to_csr slot map (no atomics, bit-deterministic).  Term 3
(the upstream coupling, outside the element pattern) is produced either as
the reference's fixed-slot COO (fill_term3_coo, for the scipy-returning
drop-in API) or, in the device solver, as a matrix-free row-gather operator
(up3d_amd.solve.linear.NewtonOperator) so the CSR pattern never changes.
"""

from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np
import scipy.sparse as sp
import torch

from .. import _native
from ..device import empty, ptr, stream_handle, to_dev, to_host
from ._io import Out, din
from .store import MeshStore, csr_pattern, elem_to_csr


# ----------------------------------------------------------- set-up helpers
def build_csr_pattern(elements, n_nodes: int) -> sp.csr_matrix:
    """Symbolic CSR of the P1 operator, canonical (sorted, deduplicated),
    zero data (pyfp3d/kernels/jacobian.py:60-74)."""
    indptr, indices = csr_pattern(elements, int(n_nodes))
    return sp.csr_matrix((np.zeros(len(indices)), indices, indptr),
                         shape=(n_nodes, n_nodes))


def build_elem_to_csr(elements, indptr, indices) -> np.ndarray:
    """elem_to_csr[e, i, j] = nnz slot of (elements[e,i], elements[e,j])
    (pyfp3d/kernels/jacobian.py:77-102); int32 here."""
    return elem_to_csr(elements, indptr, indices)


# ----------------------------------------------------- drop-in raw kernels
def _offsets(color_offsets):
    off = np.ascontiguousarray(np.asarray(color_offsets), dtype=np.int64)
    return off, off.ctypes.data_as(C.POINTER(C.c_int64))


def assemble_matrix_data_colored(color_offsets, color_elems, elem_to_csr_map,
                                 B, V, rho_tilde, data) -> None:
    """Scatter rho_tilde_e V_e (B_e B_e^T) into CSR data (the caller zeroes
    data, as in pyfp3d/kernels/jacobian.py:105-129)."""
    off, offp = _offsets(color_offsets)
    ce = din(color_elems, torch.int32)
    e2c = din(np.asarray(elem_to_csr_map).reshape(-1, 16)
              if not isinstance(elem_to_csr_map, torch.Tensor)
              else elem_to_csr_map.reshape(-1, 16), torch.int32)
    Bd = din(B, torch.float64)
    Vd = din(V, torch.float64)
    rt = din(rho_tilde, torch.float64)
    d = Out(data, torch.float64)
    _native.call("fp3d_assemble_matrix_colored", len(off) - 1, offp, ptr(ce),
                 ptr(e2c), ptr(Bd), ptr(Vd), ptr(rt), ptr(d.t), stream_handle())
    d.done()


def assemble_jacobian_data_colored(color_offsets, color_elems, elem_to_csr_map,
                                   B, V, rho_tilde, grad, s_e, inv_u2,
                                   data) -> None:
    """Newton Terms 1+2 into CSR data (caller zeroes data;
    pyfp3d/kernels/jacobian.py:132-180)."""
    off, offp = _offsets(color_offsets)
    ce = din(color_elems, torch.int32)
    e2c = din(np.asarray(elem_to_csr_map).reshape(-1, 16)
              if not isinstance(elem_to_csr_map, torch.Tensor)
              else elem_to_csr_map.reshape(-1, 16), torch.int32)
    Bd = din(B, torch.float64)
    Vd = din(V, torch.float64)
    rt = din(rho_tilde, torch.float64)
    gd = din(grad, torch.float64)
    se = din(s_e, torch.float64)
    d = Out(data, torch.float64)
    _native.call("fp3d_assemble_jacobian_colored", len(off) - 1, offp, ptr(ce),
                 ptr(e2c), ptr(Bd), ptr(Vd), ptr(rt), ptr(gd), ptr(se), None,
                 float(inv_u2), ptr(d.t), stream_handle())
    d.done()


def fill_term3_coo(active, elements, B, V, grad, upstream, s_u, inv_u2, rows,
                   cols, vals) -> None:
    """Term 3 as COO triplets, 16 per active element in the fixed slot slice
    [16a, 16a+16) (pyfp3d/kernels/jacobian.py:183-228)."""
    ac = din(active, torch.int32)
    el = din(elements, torch.int32)
    Bd = din(B, torch.float64)
    Vd = din(V, torch.float64)
    gd = din(grad, torch.float64)
    up = din(upstream, torch.int32)
    su = din(s_u, torch.float64)
    r = Out(rows, torch.int32)
    c = Out(cols, torch.int32)
    v = Out(vals, torch.float64)
    n = ac.numel()
    if r.t.numel() < 16 * n or c.t.numel() < 16 * n or v.t.numel() < 16 * n:
        raise ValueError("fill_term3_coo: triplet arrays shorter than 16*n_active")
    _native.call("fp3d_fill_term3_coo", n, ptr(ac), ptr(el), ptr(Bd), ptr(Vd),
                 ptr(gd), ptr(up), ptr(su), float(inv_u2), ptr(r.t), ptr(c.t),
                 ptr(v.t), stream_handle())
    r.done()
    c.done()
    v.done()


# ------------------------------------------------------------ the operator
class PicardOperator:
    """Per-mesh assembly workspace (pyfp3d/kernels/jacobian.py:231-375),
    resident in HBM (kernels/store.py).

    Reference-compatible methods take/return numpy arrays
    (velocities, assemble_matrix, assemble_residual,
    assemble_newton_jacobian).  The *_dev methods are the device hot path
    used by the solvers: device tensors in and out, nothing allocated per
    call; the residual and matrices are node-row gathers (one launch each,
    ascending-element accumulation, no colour passes)."""

    def __init__(self, nodes: np.ndarray, elements: np.ndarray):
        self.store = st = MeshStore.get(nodes, elements)
        self.n_nodes = st.n_nodes
        self.n_tets = st.n_tets
        self.elements = st.elements_host
        self.B = st.B_host
        self.V = st.V_host
        self.color_offsets = st.color_offsets
        self.color_elems = st.color_elems
        self.n_colors = st.n_colors
        self.nnz = st.nnz
        self._pattern_indptr = st.indptr_host
        self._pattern_indices = st.indices_host
        n = self.n_tets
        # preallocated hot-loop buffers
        self.d_grad = empty((n, 3))
        self.d_q2 = empty(n)
        self.d_R = empty(self.n_nodes)
        self.d_ones = torch.ones(n, dtype=torch.float64, device=st.device)
        self.d_terms = empty(4 * n)      # element residual terms
        self.d_K = None                  # element matrices, on first use

    # --------------------------------------------------------- device path
    def velocities_dev(self, phi: torch.Tensor, grad=None, q2=None):
        st = self.store
        grad = self.d_grad if grad is None else grad
        q2 = self.d_q2 if q2 is None else q2
        _native.call("fp3d_element_velocity_q2", self.n_tets, ptr(st.el),
                     ptr(st.B), ptr(phi), ptr(grad), ptr(q2), stream_handle())
        return grad, q2

    def state_dev(self, phi, u2, m_inf, gamma, q2cap, grad, q2n, q2l, rho, lim,
                  gdotb=None):
        """Fused grad / q2 / limiter / density sweep; gdotb (n_tets*4,
        optional) receives grad_e . B_e,a for residual_dev / Term 3 of the
        same state."""
        st = self.store
        _native.call("fp3d_element_state", self.n_tets, ptr(st.el), ptr(st.B),
                     ptr(phi), float(u2), float(m_inf), float(gamma),
                     float(q2cap), ptr(grad), ptr(q2n), ptr(q2l), ptr(rho),
                     ptr(lim), ptr(gdotb) if gdotb is not None else None,
                     stream_handle())

    def residual_dev(self, rho_tilde: torch.Tensor, grad: torch.Tensor,
                     out: Optional[torch.Tensor] = None,
                     gdotb: Optional[torch.Tensor] = None) -> torch.Tensor:
        """R_v = sum over incident (e, i) of rho_tilde_e V_e (grad_e . B_e,i):
        element terms, then the node gather in ascending element order.
        gdotb: grad . B of the same state from state_dev (then read instead
        of recomputed; the same bits)."""
        st = self.store
        R = self.d_R if out is None else out
        # terms written in incidence order (slot inc_pos[4e + a]), so the
        # node pass streams them
        _native.call("fp3d_element_residual_terms", self.n_tets, ptr(st.B),
                     ptr(st.V), ptr(rho_tilde), ptr(grad),
                     ptr(gdotb) if gdotb is not None else None, ptr(st.inc_pos),
                     ptr(self.d_terms), stream_handle())
        _native.call("fp3d_gather_node_terms", self.n_nodes, ptr(st.inc_ptr),
                     None, ptr(self.d_terms), ptr(R), stream_handle())
        return R

    def _gather_matrix(self, rho_tilde, grad, s_e, lim, inv_u2, clamp, out):
        st = self.store
        if not st.gather_ok:
            return self._scatter_matrix(rho_tilde, grad, s_e, lim, inv_u2, clamp, out)
        if self.d_K is None:
            self.d_K = empty(16 * self.n_tets)
        _native.call("fp3d_element_matrix_terms", self.n_tets, ptr(st.B),
                     ptr(st.V), ptr(rho_tilde),
                     ptr(grad) if grad is not None else None,
                     ptr(s_e) if s_e is not None else None,
                     ptr(lim) if lim is not None else None, float(inv_u2),
                     float(clamp), ptr(st.inc_pos), ptr(self.d_K), stream_handle())
        _native.call("fp3d_gather_csr_terms", self.n_nodes, st.gather_max_seg,
                     ptr(st.indptr), ptr(st.inc_ptr), None,
                     ptr(st.rel8), ptr(self.d_K), ptr(out), stream_handle())
        return out

    def _scatter_matrix(self, rho_tilde, grad, s_e, lim, inv_u2, clamp, out):
        """The same matrices by the coloured scatter (the reference's
        colour-class order; no LDS or row-length bound), for meshes the
        node gather cannot take (a CSR row longer than 256 entries, or a
        row block longer than 8,192)."""
        st = self.store
        out.zero_()
        common = (st.n_colors, st.offsets_ct, ptr(st.celems), ptr(st.e2c), ptr(st.B),
                  ptr(st.V), ptr(rho_tilde))
        p = lambda t: None if t is None else ptr(t)  # noqa: E731
        if s_e is None:
            _native.call("fp3d_assemble_matrix_colored", *common, ptr(out),
                         stream_handle())
        elif clamp == 0.0:
            _native.call("fp3d_assemble_jacobian_colored", *common, ptr(grad), ptr(s_e),
                         p(lim), float(inv_u2), ptr(out), stream_handle())
        else:
            _native.call("fp3d_assemble_jacobian_spd_colored", *common, ptr(grad),
                         ptr(s_e), p(lim), float(inv_u2), float(clamp), ptr(out),
                         stream_handle())
        return out

    def matrix_dev(self, rho_tilde: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        """Picard matrix values A = sum rho_tilde V B_i.B_j (every CSR value
        written)."""
        return self._gather_matrix(rho_tilde, None, None, None, 0.0, 0.0, out)

    def jacobian_dev(self, rho_tilde, grad, s_e, lim, inv_u2, out):
        """Terms 1+2 (s_e masked by lim when lim is not None)."""
        return self._gather_matrix(rho_tilde, grad, s_e, lim, inv_u2, 0.0, out)

    def jacobian_spd_dev(self, rho_tilde, grad, s_e, lim, inv_u2, clamp, out):
        """SPD preconditioner matrix: Terms 1+2 with the Term-2 weight
        clamped at -clamp * rho_tilde V / q2."""
        return self._gather_matrix(rho_tilde, grad, s_e, lim, inv_u2, clamp, out)

    # ---------------------------------------------- reference-compatible API
    def _csr(self, data: np.ndarray) -> sp.csr_matrix:
        return sp.csr_matrix((data, self._pattern_indices, self._pattern_indptr),
                             shape=(self.n_nodes, self.n_nodes))

    def velocities(self, phi):
        """(grad (n_tets,3), q2 (n_tets,)) in the original element order."""
        host = not isinstance(phi, torch.Tensor)
        g, q = self.velocities_dev(din(phi, torch.float64))
        g = self.store.to_original(g)
        q = self.store.to_original(q)
        return (to_host(g), to_host(q)) if host else (g, q)

    def assemble_matrix(self, rho_tilde=None) -> sp.csr_matrix:
        """Weighted stiffness matrix (6.2) as a scipy CSR sharing the pattern;
        rho_tilde=None is the Laplace limit."""
        rt = self.d_ones if rho_tilde is None else self.store.to_storage(rho_tilde)
        data = empty(self.nnz)
        self.matrix_dev(rt, data)
        return self._csr(to_host(data))

    def assemble_residual(self, phi, rho_tilde=None):
        host = not isinstance(phi, torch.Tensor)
        rt = self.d_ones if rho_tilde is None else self.store.to_storage(rho_tilde)
        g, _ = self.velocities_dev(din(phi, torch.float64))
        R = self.residual_dev(rt, g)
        return to_host(R) if host else R

    def assemble_newton_jacobian(self, phi, rho_tilde, s_e, s_u, upstream,
                                 u_inf: float = 1.0, lim_mask=None) -> sp.csr_matrix:
        """Exact Newton Jacobian at frozen upstream selection
        (pyfp3d/kernels/jacobian.py:288-359): Terms 1+2 gathered into the CSR
        on the device, Term 3 as device COO added with scipy (the
        reference's own merge).  Records newton_nnz and n_term3_active."""
        st = self.store
        up_o = np.asarray(upstream if not isinstance(upstream, torch.Tensor)
                          else upstream.cpu().numpy()).astype(np.int64)
        s_e = np.asarray(s_e, dtype=np.float64)
        s_u = np.asarray(s_u, dtype=np.float64)
        if lim_mask is not None:
            lm = np.asarray(lim_mask)
            s_e = s_e * lm
            s_u = s_u * lm[up_o]
        inv_u2 = 1.0 / (u_inf * u_inf)
        g, _ = self.velocities_dev(din(phi, torch.float64), grad=empty((self.n_tets, 3)),
                                   q2=empty(self.n_tets))
        rt = st.to_storage(rho_tilde)
        se = st.to_storage(s_e)
        data = empty(self.nnz)
        self.jacobian_dev(rt, g, se, None, inv_u2, data)
        J = self._csr(to_host(data))
        active = np.nonzero(s_u)[0].astype(np.int64)
        self.n_term3_active = int(len(active))
        if len(active):
            # Term 3 on the device from the resident arrays
            # (rows/cols are node ids, identical in both numberings)
            n3 = 16 * len(active)
            rows = empty(n3, torch.int32)
            cols = empty(n3, torch.int32)
            vals = empty(n3)
            fill_term3_coo(to_dev(active.astype(np.int32), torch.int32),
                           st.el, st.B, st.V, g,
                           st.upstream_to_storage(up_o), st.to_storage(s_u),
                           inv_u2, rows, cols, vals)
            coo = sp.coo_matrix((to_host(vals), (to_host(rows).astype(np.int64),
                                                 to_host(cols).astype(np.int64))),
                                shape=J.shape)
            J = (J + coo.tocsr()).tocsr()
        self.newton_nnz = int(J.nnz)
        return J

"""Galerkin P1 residual assembly on the device
(pyfp3d/kernels/residual.py:43-103).

R_i = sum_e rho_tilde_e V_e (grad(phi)_e . grad(N_i)).  The drop-in
assemble_residual_colored launches colour classes in order and elements
within a class in parallel; a class shares no node, so each R_i update is a
plain read-modify-write and the per-node accumulation order is the colour
sequence -- bit-deterministic, no atomics.  The operators' hot path
(PicardOperator.residual_dev) is the node gather in ascending element order.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _native
from ..device import ptr, stream_handle
from ._io import Out, din


def assemble_residual_colored(color_offsets, color_elems, elements, B, V,
                              rho_tilde, grad_elem, R) -> None:
    """Drop-in for the reference kernel (R overwritten, zeroed first)."""
    off = np.ascontiguousarray(np.asarray(color_offsets), dtype=np.int64)
    ce = din(color_elems, torch.int32)
    el = din(elements, torch.int32)
    Bd = din(B, torch.float64)
    Vd = din(V, torch.float64)
    rt = din(rho_tilde, torch.float64)
    gd = din(grad_elem, torch.float64)
    Ro = Out(R, torch.float64)
    if len(off) < 1 or off[-1] > ce.numel() or ce.numel() > el.shape[0]:
        raise ValueError("assemble_residual_colored: bad colour partition")
    _native.call("fp3d_assemble_residual_colored", len(off) - 1,
                 off.ctypes.data_as(C.POINTER(C.c_int64)), ptr(ce), ptr(el),
                 ptr(Bd), ptr(Vd), ptr(rt), ptr(gd), Ro.t.numel(), ptr(Ro.t),
                 stream_handle())
    Ro.done()


def assemble_stiffness_matrix(nodes, elements):
    """SPD Laplace stiffness matrix (rho == 1) through the device assembly,
    as a scipy CSR (pyfp3d/kernels/residual.py:79-99)."""
    from .jacobian import PicardOperator

    return PicardOperator(nodes, elements).assemble_matrix()

"""Element-constant P1 velocity sweep on the device
(pyfp3d/kernels/gradient.py:32-64 element_velocity_q2)."""

from __future__ import annotations

import torch

from .. import _native
from ..device import ptr, stream_handle
from ._io import Out, din


def element_velocity_q2(elements, B, phi, grad_out, q2_out) -> None:
    """grad(phi)_e = sum_i phi_i B_e[i, :] and q2_e = |grad(phi)_e|^2.

    Same signature as the reference: elements (n_tets, 4), B (n_tets, 4, 3),
    phi (n_nodes,), grad_out (n_tets, 3) and q2_out (n_tets,) overwritten.
    Arguments may be numpy arrays or device tensors."""
    el = din(elements, torch.int32)
    Bd = din(B, torch.float64)
    ph = din(phi, torch.float64)
    g = Out(grad_out, torch.float64)
    q = Out(q2_out, torch.float64)
    n = el.shape[0]
    if Bd.numel() != 12 * n or g.t.numel() != 3 * n or q.t.numel() != n:
        raise ValueError("element_velocity_q2: inconsistent array sizes")
    _native.call("fp3d_element_velocity_q2", n, ptr(el), ptr(Bd), ptr(ph),
                 ptr(g.t), ptr(q.t), stream_handle())
    g.done()
    q.done()

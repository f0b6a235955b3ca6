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

"""Exact elimination of the aux (sheet) block of the level-set system
(pyfp3d/solve/schur_ls.py:46-146, SchurReducedSystem and
main_block_preconditioner):

    K x_m = J_mm x_m - J_ma J_aa^{-1} J_am x_m
    r     = b_m - J_ma J_aa^{-1} b_a
    x_a   = J_aa^{-1} (b_a - J_am x_m)

J_mm, J_ma and J_am live in HBM as CSR; J_aa -- one row per aux DOF, the
thin strip along the sheet -- is factorised on the host (SuperLU) and its
solves are the only host step inside a matvec (n_aux doubles each way).
GMRES runs on the device over the main block; the aux rows hold exactly by
back-substitution, so the reduced residual is the full-system residual.
"""

from __future__ import annotations

from typing import Optional, Tuple

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

from .. import _native
from ..device import empty, ptr, stream_handle, to_dev, to_host
from .amg import DeviceCSR
from .linear import build_amg_preconditioner, solve_gmres


class SchurReducedSystem:
    """Aux-tail elimination of a free-reduced fused matrix A_free (main-free
    DOFs first, then the free aux DOFs: `free` sorted ascending).
    n_aux_expected, when given, must equal the number of free aux DOFs
    (ValueError otherwise, as the reference does)."""

    def __init__(self, A_free: sp.spmatrix, free: np.ndarray, n_main: int,
                 n_aux_expected: Optional[int] = None) -> None:
        free = np.asarray(free, dtype=np.int64)
        if free.size > 1 and np.any(np.diff(free) <= 0):
            raise ValueError("`free` must be sorted ascending without repeats")
        n_mf = int(np.searchsorted(free, n_main))
        n_aux = int(free.size - n_mf)
        if n_aux_expected is not None and n_aux != int(n_aux_expected):
            raise ValueError(f"Schur split expected {n_aux_expected} free aux DOFs, "
                             f"found {n_aux}")
        A = sp.csr_matrix(A_free)
        self.n_mf, self.n_aux = n_mf, n_aux
        self.main_free = free[:n_mf]
        self.J_mm = A[:n_mf, :n_mf].tocsr()
        self.J_ma = A[:n_mf, n_mf:].tocsr()
        self.J_am = A[n_mf:, :n_mf].tocsr()
        self.J_aa = A[n_mf:, n_mf:].tocsc()
        self.lu_aa = spla.splu(self.J_aa)
        self._d_mm = DeviceCSR(self.J_mm)
        self._d_ma = DeviceCSR(self.J_ma)
        self._d_am = DeviceCSR(self.J_am)
        self._t_a = empty(max(n_aux, 1))
        self._t_m = empty(max(n_mf, 1))

    def _ma_times(self, v_aux: np.ndarray, y: torch.Tensor) -> None:
        """y <- J_ma v_aux (device)."""
        M = self._d_ma
        _native.call("fp3d_csr_spmv", M.n, M.nnz, ptr(M.indptr), ptr(M.indices),
                     ptr(M.data), ptr(to_dev(v_aux)), ptr(y), stream_handle())

    def matvec_dev(self, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
        self._d_mm.matvec(x, y)
        if self.n_aux:
            self._d_am.matvec(x, self._t_a[: self.n_aux])
            self._ma_times(self.lu_aa.solve(to_host(self._t_a[: self.n_aux])), self._t_m)
            y.sub_(self._t_m)
        return y

    def operator(self) -> spla.LinearOperator:
        """K as a host LinearOperator (matrix-free, one strip solve each)."""
        def mv(x):
            return self.J_mm @ x - self.J_ma @ self.lu_aa.solve(self.J_am @ x)
        return spla.LinearOperator((self.n_mf, self.n_mf), matvec=mv, dtype=np.float64)

    def reduce_rhs(self, b_free: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
        b_free = np.asarray(b_free, dtype=np.float64)
        b_m, b_a = b_free[: self.n_mf], b_free[self.n_mf:]
        return b_m - self.J_ma @ self.lu_aa.solve(b_a), b_a

    def back_substitute(self, x_m: np.ndarray, b_a: np.ndarray) -> np.ndarray:
        return self.lu_aa.solve(np.asarray(b_a) - self.J_am @ np.asarray(x_m))

    def solve(self, b_free: np.ndarray, M=None, rtol: float = 1e-10,
              x0_main: Optional[np.ndarray] = None, restart: int = 60,
              maxiter: int = 50) -> Tuple[np.ndarray, int, int]:
        """(x_full, n_inner, info): device GMRES on K, aux by
        back-substitution (exact aux rows also when GMRES stalls)."""
        r, b_a = self.reduce_rhs(b_free)
        x_m, n_it, info = solve_gmres(self.matvec_dev, to_dev(r), M=M, rtol=rtol,
                                      x0=x0_main, restart=restart, maxiter=maxiter,
                                      on_fail="return")
        x_m = to_host(x_m)
        return np.concatenate([x_m, self.back_substitute(x_m, b_a)]), n_it, info


def main_block_preconditioner(mvop, rho_pair, main_free: np.ndarray):
    """SA-AMG (device V-cycle) of the single-valued Picard matrix at the
    own-side densities, restricted to the free main DOFs."""
    if rho_pair is None:
        w = None
    else:
        w, _ = mvop.side_weights(*rho_pair)
    A = mvop.op.assemble_matrix(None if w is None else to_host(w)).tocsr()
    mf = np.asarray(main_free, dtype=np.int64)
    return build_amg_preconditioner(A[mf][:, mf].tocsr())[1]

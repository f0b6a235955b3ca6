"""Explicit-circulation wake constraint on a cut mesh
(pyfp3d/constraints/wake.py:32-212, WakeConstraint and kutta_targets).

Reduced unknowns are the original nodes; a slave ("+" copy of a sheet
node) is eliminated as phi_slave = phi_master + Gamma_station:

    phi_cut = T phi_red + G Gamma,   T = [I; E_master],  G = slave indicator
    A_red = T^T A T,   rhs_red = T^T b - Gamma . h,   h_j = T^T A g_j

The reductions are host sparse products on the cut mesh's CSR (assembled
on the device); the solves run on the device."""

from __future__ import annotations

from typing import Tuple

import numpy as np
import scipy.sparse as sp


class WakeConstraint:
    def __init__(self, A: sp.spmatrix, wc):
        n_cut = A.shape[0]
        n_red = wc.n_nodes_orig
        if n_cut != n_red + len(wc.slave_nodes):
            raise ValueError(f"A is {n_cut}x{n_cut} but the cut has {n_red} + "
                             f"{len(wc.slave_nodes)} nodes")
        rows = np.concatenate([np.arange(n_red), wc.slave_nodes])
        cols = np.concatenate([np.arange(n_red), wc.master_nodes])
        self.T = sp.csr_matrix((np.ones(n_cut), (rows, cols)), shape=(n_cut, n_red))
        self.G = sp.csr_matrix((np.ones(len(wc.slave_nodes)),
                                (wc.slave_nodes, wc.node_station)),
                               shape=(n_cut, wc.n_stations))
        self.wc = wc
        self.n_reduced = n_red
        self.update_matrix(A)

    def reduce_operator(self, A: sp.spmatrix) -> Tuple[sp.csr_matrix, sp.csr_matrix]:
        """(T^T A T, T^T A G): the reduced operator and d(reduced residual)
        / d(Gamma) through the jump map."""
        A = sp.csr_matrix(A)
        AT = A @ self.T
        return (self.T.T @ AT).tocsr(), (self.T.T @ (A @ self.G)).tocsr()

    def update_matrix(self, A: sp.spmatrix) -> None:
        self.A_reduced, H = self.reduce_operator(A)
        self._h = H.T.toarray()

    def reduced_rhs(self, b: np.ndarray, gamma) -> np.ndarray:
        gamma = np.atleast_1d(np.asarray(gamma, dtype=np.float64))
        return self.T.T @ np.asarray(b, dtype=np.float64) - gamma @ self._h

    def expand(self, phi_red: np.ndarray, gamma) -> np.ndarray:
        gamma = np.atleast_1d(np.asarray(gamma, dtype=np.float64))
        phi = np.empty(self.n_reduced + len(self.wc.slave_nodes))
        phi[: self.n_reduced] = phi_red
        phi[self.wc.slave_nodes] = (np.asarray(phi_red)[self.wc.master_nodes]
                                    + self.wc.gamma_at_nodes(gamma))
        return phi

    def to_reduced_dirichlet(self, dirichlet_nodes, dirichlet_values):
        """Drop slave entries (they follow their master + Gamma)."""
        d = np.asarray(dirichlet_nodes, dtype=np.int64)
        keep = d < self.n_reduced
        red, idx = np.unique(d[keep], return_index=True)
        return red, np.asarray(dirichlet_values, dtype=np.float64)[keep][idx]


def kutta_targets(phi: np.ndarray, wc) -> np.ndarray:
    """Per station: mean over its TE nodes of phi[upper probe] - phi[lower
    probe] (wall nodes one edge off the TE on either side)."""
    jumps = np.asarray(phi)[wc.kutta_upper] - np.asarray(phi)[wc.kutta_lower]
    n_st = wc.n_stations
    return (np.bincount(wc.te_station, weights=jumps, minlength=n_st)
            / np.bincount(wc.te_station, minlength=n_st))

"""Shared device helpers of the nonlinear drivers."""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import torch

from .. import _native
from ..device import empty, ptr, stream_handle, to_dev, to_host
from .amg import AMGHierarchy


class Reducer:
    """Deterministic device norms (2-norm and inf-norm) with one host read."""

    def __init__(self):
        self.work = empty(_native.lib().fp3d_reduce_work_doubles())
        self.out = empty(2)

    def norms(self, x: torch.Tensor):
        _native.call("fp3d_norms", x.numel(), ptr(x), ptr(self.work),
                     ptr(self.out), stream_handle())
        h = to_host(self.out)
        return float(h[0]), float(h[1])


def dirichlet_setup(n_nodes: int, dirichlet_nodes, dirichlet_values, device):
    """(free_mask uint8 dev, dir idx int64 dev, vals dev, phi_d dev)."""
    dn = np.asarray(dirichlet_nodes, dtype=np.int64)
    dv = np.asarray(dirichlet_values, dtype=np.float64)
    if dn.shape != dv.shape:
        raise ValueError("dirichlet_nodes and dirichlet_values differ in length")
    if len(dn) and (dn.min() < 0 or dn.max() >= n_nodes):
        raise ValueError("Dirichlet node id out of range")
    fm = np.ones(n_nodes, dtype=np.uint8)
    fm[dn] = 0
    phi_d = np.zeros(n_nodes)
    phi_d[dn] = dv
    return (to_dev(fm, torch.uint8), to_dev(dn, torch.int64),
            to_dev(dv, torch.float64), to_dev(phi_d, torch.float64))


def csr_to_host(store, data: torch.Tensor) -> sp.csr_matrix:
    return sp.csr_matrix((to_host(data), store.indices_host, store.indptr_host),
                         shape=(store.n_nodes, store.n_nodes))


def mask_dirichlet(store, free_mask, data, data32=None):
    _native.call("fp3d_csr_mask_dirichlet", store.n_nodes, ptr(store.indptr),
                 ptr(store.indices), ptr(free_mask), ptr(data), ptr(data32),
                 stream_handle())


def masked(free_mask, x, out):
    _native.call("fp3d_mask_copy", x.numel(), ptr(free_mask), ptr(x), ptr(out),
                 stream_handle())
    return out


def amg_from_device(store, data, precision="f64", **amg_kw) -> AMGHierarchy:
    """SA hierarchy (host set-up) of a masked device CSR matrix on the
    store's pattern; the finest level keeps the device values and pattern.
    amg_kw go to AMGHierarchy (degree, max_coarse, ...)."""
    return AMGHierarchy(csr_to_host(store, data), precision=precision, seed=0,
                        device_pattern=store, device_values=data, **amg_kw)


def amg_plan_from_device(store, data, precision="f64", **amg_kw):
    """amg_from_device in parts (solve.amg.AMGPlan): the host set-up can
    run on another thread."""
    from .amg import AMGPlan

    return AMGPlan(csr_to_host(store, data), precision=precision, seed=0,
                   device_pattern=store, device_values=data, **amg_kw)

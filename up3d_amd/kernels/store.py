"""Per-mesh device store: all element data resident in HBM in COLOUR-MAJOR
order.

Set-up (host, once per mesh, as pyfp3d/kernels/jacobian.py:246-262 does):
geometry B_e/V_e, greedy colouring + stable colour partition, canonical CSR
pattern, elem_to_csr, face adjacency, centroids.  The elements are then
renumbered so that colour class c is the contiguous range
[color_offsets[c], color_offsets[c+1]) and, inside it, ascending original
element id -- exactly the reference's colour_elems order.  Every element
kernel therefore reads coalesced rows, the coloured scatters keep the
reference's accumulation order (colour by colour, one element per node
per colour), and element-indexed results convert to the original order
only at the API boundary (`perm`: storage slot -> original id).

Device layout (storage order, k = storage slot):
  el       int32 (n_tets, 4)     node ids
  B        f64   (n_tets, 12)    shape gradients, row-major (i, d)
  V        f64   (n_tets,)
  e2c      int32 (n_tets, 16)    CSR slot of (elements[k,i], elements[k,j])
  fn       int32 (n_tets, 4)     face neighbours as storage slots, -1 boundary
  cen      f64   (n_tets, 3)     centroids
  dcen     f64   (n_tets, 12)    centroid(fn[k,f]) - centroid(k), per face
  nodes    f64   (n_nodes, 3)
  indptr   int32 (n_nodes+1,), indices int32 (nnz,)
  inc_ptr  int32 (n_nodes+1,), inc_ent int32 (4 n_tets,)  node -> 4k+i
"""

from __future__ import annotations

import ctypes as C
import hashlib
from typing import Dict

import numpy as np
import torch

from .. import _native
from ..device import require_gpu, to_dev
from ..mesh.coloring import color_partition_csr
from ..mesh.metrics import build_face_adjacency, precompute_element_geometry


def csr_pattern(elements: np.ndarray, n_nodes: int):
    """Canonical sorted, deduplicated node-node CSR pattern (host C++;
    pyfp3d/kernels/jacobian.py:60-74).  Returns (indptr, indices) int32."""
    el = np.ascontiguousarray(elements, dtype=np.int32)
    indptr = np.empty(n_nodes + 1, dtype=np.int32)
    nnz = C.c_int64(0)
    L = _native.lib()
    _native.check(L.fp3d_host_csr_pattern(len(el), el.ctypes.data, n_nodes,
                                          indptr.ctypes.data, None,
                                          C.byref(nnz)), "build_csr_pattern")
    indices = np.empty(nnz.value, dtype=np.int32)
    _native.check(L.fp3d_host_csr_pattern(len(el), el.ctypes.data, n_nodes,
                                          indptr.ctypes.data,
                                          indices.ctypes.data, C.byref(nnz)),
                  "build_csr_pattern")
    return indptr, indices


def elem_to_csr(elements: np.ndarray, indptr: np.ndarray, indices: np.ndarray):
    """(n_tets, 4, 4) int32 CSR slot map (pyfp3d/kernels/jacobian.py:77-102)."""
    el = np.ascontiguousarray(elements, dtype=np.int32)
    ip = np.ascontiguousarray(indptr, dtype=np.int32)
    ix = np.ascontiguousarray(indices, dtype=np.int32)
    out = np.empty((len(el), 4, 4), dtype=np.int32)
    _native.check(_native.lib().fp3d_host_elem_to_csr(
        len(el), el.ctypes.data, ip.ctypes.data, ix.ctypes.data,
        out.ctypes.data), "build_elem_to_csr")
    return out


class MeshStore:
    """Device-resident mesh in colour-major order (module docstring)."""

    _cache: Dict[str, "MeshStore"] = {}

    @classmethod
    def get(cls, nodes: np.ndarray, elements: np.ndarray) -> "MeshStore":
        nd = np.ascontiguousarray(nodes, dtype=np.float64)
        el = np.ascontiguousarray(elements, dtype=np.int32)
        h = hashlib.blake2b(digest_size=16)
        h.update(np.int64(len(nd)).tobytes())
        h.update(nd.tobytes())
        h.update(el.tobytes())
        key = h.hexdigest()
        st = cls._cache.get(key)
        if st is None:
            st = cls(nd, el)
            cls._cache.clear()   # keep one mesh resident at a time
            cls._cache[key] = st
        return st

    def __init__(self, nodes: np.ndarray, elements: np.ndarray):
        require_gpu()
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        el = np.ascontiguousarray(elements, dtype=np.int32)
        self.n_nodes = len(nodes)
        self.n_tets = len(el)
        self.nodes_host = nodes
        self.elements_host = el
        # ---- host set-up, original element order
        B, V = precompute_element_geometry(nodes, el)
        off, celems, n_colors = color_partition_csr(el)
        indptr, indices = csr_pattern(el, self.n_nodes)
        e2c = elem_to_csr(el, indptr, indices)
        fn, _ = build_face_adjacency(el)
        cen = nodes[el].mean(axis=1)
        self.B_host, self.V_host = B, V
        self.color_offsets = np.ascontiguousarray(off, dtype=np.int64)
        self.color_elems = celems
        self.n_colors = int(n_colors)
        self.indptr_host, self.indices_host = indptr, indices
        self.nnz = int(len(indices))
        self.face_neighbors_host = fn
        self.centroids_host = cen
        self.e2c_host = e2c
        # ---- colour-major renumbering
        perm = celems.astype(np.int64)                 # slot -> original
        inv = np.empty_like(perm)
        inv[perm] = np.arange(self.n_tets)
        self.perm_host, self.inv_perm_host = perm, inv
        el_s = el[perm]
        fn_s = fn[perm]
        fn_s = np.where(fn_s >= 0, inv[np.maximum(fn_s, 0)], -1).astype(np.int32)
        # node -> (4 k + i) incidence in storage order (ascending k)
        flat = el_s.reshape(-1).astype(np.int64)
        order = np.argsort(flat, kind="stable")
        inc_ptr = np.zeros(self.n_nodes + 1, dtype=np.int64)
        np.cumsum(np.bincount(flat, minlength=self.n_nodes), out=inc_ptr[1:])
        # ---- upload
        self.el = to_dev(el_s, torch.int32)
        self.B = to_dev(B[perm].reshape(self.n_tets, 12), torch.float64)
        self.V = to_dev(V[perm], torch.float64)
        self.e2c = to_dev(e2c[perm].reshape(self.n_tets, 16), torch.int32)
        self.fn = to_dev(fn_s, torch.int32)
        self.cen = to_dev(cen[perm], torch.float64)
        # per-face centroid displacement, storage order (0 on the boundary)
        cs = cen[perm]
        dcen = np.where((fn_s >= 0)[:, :, None],
                        cs[np.maximum(fn_s, 0)] - cs[:, None, :], 0.0)
        self.dcen = to_dev(np.ascontiguousarray(dcen).reshape(self.n_tets, 12),
                           torch.float64)
        self.nodes = to_dev(nodes, torch.float64)
        self.indptr = to_dev(indptr, torch.int32)
        self.indices = to_dev(indices, torch.int32)
        self.inc_ptr = to_dev(inc_ptr.astype(np.int32), torch.int32)
        self.inc_ent = to_dev(order.astype(np.int32), torch.int32)
        self.perm = to_dev(perm, torch.int64)
        self.inv_perm = to_dev(inv, torch.int64)
        self.offsets_ct = self.color_offsets.ctypes.data_as(
            C.POINTER(C.c_int64))

    # element arrays: original order <-> storage order
    def to_storage(self, a, dtype=torch.float64) -> torch.Tensor:
        t = to_dev(a, dtype) if not (isinstance(a, torch.Tensor) and a.is_cuda) \
            else a.to(dtype)
        return t.index_select(0, self.perm).contiguous()

    def to_original(self, t: torch.Tensor) -> torch.Tensor:
        return t.index_select(0, self.inv_perm).contiguous()

    def upstream_to_original(self, up_s: torch.Tensor) -> torch.Tensor:
        """storage-slot upstream ids -> original element ids, original order"""
        return self.perm.index_select(0, up_s.long()).index_select(0, self.inv_perm)

    def upstream_to_storage(self, up_o) -> torch.Tensor:
        t = up_o if isinstance(up_o, torch.Tensor) and up_o.is_cuda \
            else to_dev(np.asarray(up_o), torch.int64)
        return self.inv_perm.index_select(0, t.long().index_select(0, self.perm)
                                          ).to(torch.int32).contiguous()

    @property
    def device(self):
        return self.B.device

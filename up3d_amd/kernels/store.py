"""Per-mesh device store: all element data resident in HBM in the mesh's
own element order.

Set-up (host, once per mesh, as pyfp3d/kernels/jacobian.py:246-262 does):
geometry B_e/V_e, greedy colouring + stable colour partition, canonical CSR
pattern, elem_to_csr, face adjacency, centroids.  Element arrays stay in
the caller's element order (a mesh whose elements are sorted by node ids
after RCM, as the reference's meshes are, keeps face neighbours and the
elements around a node close in memory), so element ids are the same in
the API and on the device.  The node-gather assembly walks the
node -> (4 e + i) incidence in ascending e; the coloured drop-in kernels
visit colour class c through the index list color_elems.

Device layout (element e, node v):
  el       int32 (n_tets, 4)     node ids
  B        f64   (n_tets, 12)    shape gradients, row-major (i, d)
  V        f64   (n_tets,)
  e2c      int32 (n_tets, 16)    CSR slot of (elements[e,i], elements[e,j])
  rel8     uint8 (4 n_tets, 4)   rel8[q, j]: offset inside row v of the slot
                                 of (v, elements[e, j]) for incidence
                                 q = (v, 4 e + i) of inc_ent
  fn       int32 (n_tets, 4)     face neighbours, -1 on the boundary
  cen      f64   (n_tets, 3)     centroids
  dcen     f64   (n_tets, 12)    centroid(fn[e,f]) - centroid(e), per face
  celems   int32 (n_tets,)       colour-partitioned element ids
  nodes    f64   (n_nodes, 3)
  indptr   int32 (n_nodes+1,), indices int32 (nnz,)
  inc_ptr  int32 (n_nodes+1,), inc_ent int32 (4 n_tets,)  node -> 4e+i
  inc_pos  int32 (4 n_tets,)   inverse of inc_ent: 4e+i -> incidence slot
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


def array_digest(*arrays) -> str:
    """128-bit content key of numpy arrays (their bytes in order, read in
    place): xxh3 where the xxhash module is present (about 15x faster than
    blake2b on mesh-sized arrays), else blake2b.  A process-local cache key,
    never stored."""
    try:
        import xxhash

        h = xxhash.xxh3_128()
    except ImportError:
        h = hashlib.blake2b(digest_size=16)
    for a in arrays:
        a = np.ascontiguousarray(a)
        h.update(memoryview(a).cast("B"))
    return h.hexdigest()


def csr_offsets16(indptr: np.ndarray, indices: np.ndarray):
    """int16 device array of col - row for a square CSR pattern whose every
    |col - row| fits in int16, else None."""
    indptr = np.asarray(indptr, dtype=np.int64)
    rows = np.repeat(np.arange(len(indptr) - 1, dtype=np.int64), np.diff(indptr))
    off = np.asarray(indices, dtype=np.int64) - rows
    if len(off) and (off.min() < -32768 or off.max() > 32767):
        return None
    return to_dev(off.astype(np.int16), torch.int16)


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
    """Device-resident mesh (module docstring)."""

    _cache: Dict[str, "MeshStore"] = {}
    max_resident = 4      # meshes kept set up on the device (LRU)

    @classmethod
    def get(cls, nodes: np.ndarray, elements: np.ndarray) -> "MeshStore":
        nd = np.ascontiguousarray(nodes, dtype=np.float64)
        el = np.ascontiguousarray(elements, dtype=np.int32)
        key = array_digest(np.int64(len(nd)), nd, el)
        st = cls._cache.pop(key, None)
        if st is None:
            st = cls(nd, el)
        cls._cache[key] = st          # most recently used last
        while len(cls._cache) > cls.max_resident:
            cls._cache.pop(next(iter(cls._cache)))
        return st

    def __init__(self, nodes: np.ndarray, elements: np.ndarray):
        require_gpu()
        nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        el = np.ascontiguousarray(elements, dtype=np.int32)
        self.n_nodes = len(nodes)
        self.n_tets = len(el)
        self.nodes_host = nodes
        self.elements_host = el
        # ---- host set-up, original element order.  The independent pieces
        # (each a library call or a numpy pass that releases the GIL) run on
        # worker threads beside the CSR pattern; each gives the values it
        # gives alone
        from concurrent.futures import ThreadPoolExecutor

        with ThreadPoolExecutor(max_workers=4) as pool:
            f_geo = pool.submit(precompute_element_geometry, nodes, el)
            f_col = pool.submit(color_partition_csr, el)
            f_fn = pool.submit(build_face_adjacency, el)
            f_cen = pool.submit(lambda: nodes[el].mean(axis=1))
            indptr, indices = csr_pattern(el, self.n_nodes)
            e2c = elem_to_csr(el, indptr, indices)
            B, V = f_geo.result()
            off, celems, n_colors = f_col.result()
            fn, _ = f_fn.result()
            cen = f_cen.result()
            # per-face centroid displacement (0 on the boundary), beside the
            # incidence sort and the uploads below
            f_dcen = pool.submit(lambda: np.ascontiguousarray(
                np.where((fn >= 0)[:, :, None], cen[np.maximum(fn, 0)] - cen[:, None, :],
                         0.0)).reshape(self.n_tets, 12))
            self._build_rest(nodes, el, B, V, off, celems, n_colors, indptr, indices,
                             e2c, fn, cen, f_dcen)

    def _build_rest(self, nodes, el, B, V, off, celems, n_colors, indptr, indices, e2c,
                    fn, cen, f_dcen):
        self.B_host, self.V_host = B, V
        self.color_offsets = np.ascontiguousarray(off, dtype=np.int64)
        self.color_elems = celems
        self.n_colors = int(n_colors)
        self.indptr_host, self.indices_host = indptr, indices
        self.nnz = int(len(indices))
        self.face_neighbors_host = fn
        self.centroids_host = cen
        self.e2c_host = e2c
        # ---- node -> (4 e + i) incidence, ascending e
        flat = el.reshape(-1).astype(np.int64)
        order = np.argsort(flat, kind="stable")
        inc_ptr = np.zeros(self.n_nodes + 1, dtype=np.int64)
        np.cumsum(np.bincount(flat, minlength=self.n_nodes), out=inc_ptr[1:])
        # LDS bound of the node-gather matrix assembly: the longest CSR slice
        # owned by one block of consecutive rows
        rows = _native.lib().fp3d_gather_rows_per_block()
        starts = np.arange(0, self.n_nodes, rows)
        ends = np.minimum(starts + rows, self.n_nodes)
        self.gather_max_seg = int((indptr[ends].astype(np.int64)
                                   - indptr[starts]).max()) if self.n_nodes else 1
        # the node-gather matrix assembly stages a row block's CSR slice in
        # LDS (at most 8,192 entries) and addresses a slot by its byte offset
        # inside the row (rows of at most 256 entries); a mesh beyond either
        # bound assembles its matrices with the coloured scatter instead
        self.gather_ok = self.gather_max_seg * 8 <= 65536
        # ---- upload
        self.el = to_dev(el, torch.int32)
        self.B = to_dev(B.reshape(self.n_tets, 12), torch.float64)
        self.V = to_dev(V, torch.float64)
        self.e2c = to_dev(e2c.reshape(self.n_tets, 16), torch.int32)
        # slot offset inside its row, one byte per (e, a, b)
        rel = e2c.reshape(self.n_tets, 4, 4) - indptr[el][:, :, None]
        if self.n_tets and (rel.min() < 0 or rel.max() > 255):
            self.gather_ok = False
        self.rel8 = None
        if self.gather_ok:
            # stored in node-incidence order (entry q holds the four offsets
            # of incidence inc_ent[q]), so a row's gather reads them
            # contiguously
            rel_inc = np.ascontiguousarray(rel.astype(np.uint8)).reshape(-1, 4)[order]
            self.rel8 = to_dev(np.ascontiguousarray(rel_inc).reshape(-1), torch.uint8)
        self.fn = to_dev(fn.astype(np.int32), torch.int32)
        self.cen = to_dev(cen, torch.float64)
        self.celems = to_dev(celems.astype(np.int32), torch.int32)
        self.nodes = to_dev(nodes, torch.float64)
        self.indptr = to_dev(indptr, torch.int32)
        self.indices = to_dev(indices, torch.int32)
        # the same columns as int16 offsets from the row when the bandwidth
        # allows (RCM orderings: 5,314 on the 4.7M-tet channel), for the
        # Krylov matvecs (half the index bytes); None otherwise
        self.indices16 = csr_offsets16(indptr, indices)
        self.inc_ptr = to_dev(inc_ptr.astype(np.int32), torch.int32)
        self.inc_ent = to_dev(order.astype(np.int32), torch.int32)
        # inverse of inc_ent: slot of (e, a) in the incidence list, where the
        # element passes write their terms so the node passes read in order
        inc_pos = np.empty(len(order), dtype=np.int32)
        inc_pos[order] = np.arange(len(order), dtype=np.int32)
        self.inc_pos = to_dev(inc_pos, torch.int32)
        self.dcen = to_dev(f_dcen.result(), torch.float64)
        self.offsets_ct = self.color_offsets.ctypes.data_as(
            C.POINTER(C.c_int64))

    # element arrays: API order == device order
    def to_storage(self, a, dtype=torch.float64) -> torch.Tensor:
        if isinstance(a, torch.Tensor) and a.is_cuda:
            return a.to(dtype).contiguous()
        return to_dev(a, dtype)

    def to_original(self, t: torch.Tensor) -> torch.Tensor:
        return t

    def upstream_to_original(self, up_s: torch.Tensor) -> torch.Tensor:
        return up_s.long()

    def upstream_to_storage(self, up_o) -> torch.Tensor:
        t = up_o if isinstance(up_o, torch.Tensor) and up_o.is_cuda \
            else to_dev(np.asarray(up_o), torch.int64)
        return t.to(torch.int32).contiguous()

    @property
    def device(self):
        return self.B.device

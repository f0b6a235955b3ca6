"""Per-mesh geometry set-up (host, once): shape gradients B_e, volumes V_e
and face adjacency, computed by the native host code of libfp3d_hip.so.

Mirrors pyfp3d/mesh/metrics.py:176-204 (precompute_element_geometry) and
:207-278 (build_face_adjacency): same outputs, same degenerate-tet
ValueError, face f opposite local node f, -1 on the boundary.
"""

from __future__ import annotations

import ctypes as C
from typing import Tuple

import numpy as np

from .. import _native


def _el32(elements) -> np.ndarray:
    el = np.ascontiguousarray(elements, dtype=np.int32)
    if el.ndim != 2 or el.shape[1] != 4:
        raise ValueError(f"elements must be (n_tets, 4), got {el.shape}")
    return el


def precompute_element_geometry(nodes: np.ndarray, elements: np.ndarray
                                ) -> Tuple[np.ndarray, np.ndarray]:
    """(B (n_tets,4,3), V (n_tets,)); raises ValueError on a degenerate tet
    (|det J| < 1e-12 x product of edge lengths, pyfp3d/mesh/metrics.py:155-164)."""
    nd = np.ascontiguousarray(nodes, dtype=np.float64)
    el = _el32(elements)
    if len(el) and (el.min() < 0 or el.max() >= len(nd)):
        raise ValueError("element node id out of range")
    n = len(el)
    B = np.empty((n, 4, 3), dtype=np.float64)
    V = np.empty(n, dtype=np.float64)
    bad = C.c_int64(-1)
    rc = _native.lib().fp3d_host_element_geometry(
        n, el.ctypes.data, nd.ctypes.data, B.ctypes.data, V.ctypes.data,
        C.byref(bad))
    _native.check(rc, "precompute_element_geometry")
    return B, V


def build_face_adjacency(elements: np.ndarray):
    """(face_neighbors int32 (n,4), face_orientations int32 (n,4)): the
    neighbour across face f (opposite local node f), -1 on the boundary;
    orientation +1 on the first owner of a shared face, -1 on the second
    (pyfp3d/mesh/metrics.py:207-278)."""
    el = _el32(elements)
    fn = np.empty(el.shape, dtype=np.int32)
    rc = _native.lib().fp3d_host_face_adjacency(len(el), el.ctypes.data,
                                                fn.ctypes.data)
    _native.check(rc, "build_face_adjacency")
    ids = np.arange(len(el))[:, None]
    orient = np.where(fn < 0, 0, np.where(fn > ids, 1, -1)).astype(np.int32)
    return fn, orient


def element_centroids(nodes: np.ndarray, elements: np.ndarray) -> np.ndarray:
    """nodes[elements].mean(axis=1) (pyfp3d/kernels/upwind.py:750)."""
    return np.asarray(nodes, dtype=np.float64)[np.asarray(elements)].mean(axis=1)

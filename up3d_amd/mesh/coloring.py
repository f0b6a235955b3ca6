"""Element colouring for atomic-free coloured assembly (host, once per mesh).

Same greedy first-fit as pyfp3d/mesh/coloring.py:65-126 (elements visited
in id order, smallest colour not used by any node-sharing element, at
most 256 colours) and the same partition as :129-145 (stable sort by
colour, so ascending element id within a colour).  The colouring is
bit-identical to the reference's for the same element order.
"""

from __future__ import annotations

import ctypes as C
from typing import Tuple

import numpy as np

from .. import _native


def greedy_coloring(elements: np.ndarray) -> Tuple[np.ndarray, int]:
    """(colors int32 (n_tets,), n_colors); ValueError above 256 colours."""
    el = np.ascontiguousarray(elements, dtype=np.int32)
    n_nodes = int(el.max()) + 1 if len(el) else 0
    colors = np.empty(len(el), dtype=np.int32)
    nc = C.c_int32(0)
    rc = _native.lib().fp3d_host_greedy_coloring(
        len(el), el.ctypes.data, n_nodes, colors.ctypes.data, C.byref(nc))
    _native.check(rc, "greedy_coloring")
    return colors, int(nc.value)


def color_partition_csr(elements: np.ndarray) -> Tuple[np.ndarray, np.ndarray, int]:
    """(color_offsets int64 (n_colors+1,), color_elems int64 (n_tets,),
    n_colors): elements of colour c are color_elems[off[c]:off[c+1]] in
    ascending id order."""
    colors, n_colors = greedy_coloring(elements)
    order = np.argsort(colors, kind="stable")
    counts = np.bincount(colors, minlength=n_colors)
    off = np.zeros(n_colors + 1, dtype=np.int64)
    np.cumsum(counts, out=off[1:])
    return off, order.astype(np.int64), n_colors


def validate_coloring(elements: np.ndarray, colors: np.ndarray) -> bool:
    """True when no two elements sharing a node have the same colour; raises
    AssertionError otherwise (pyfp3d/mesh/coloring.py:148-177)."""
    el = np.asarray(elements, dtype=np.int64)
    c = np.asarray(colors, dtype=np.int64)
    node = el.reshape(-1)
    col = np.repeat(c, 4)
    key = node * 512 + col
    uniq, counts = np.unique(key, return_counts=True)
    if np.any(counts > 1):
        k = uniq[np.argmax(counts > 1)]
        raise AssertionError(
            f"Same-color adjacency at node {k // 512}: colour {k % 512}")
    return True

"""Transfer of nodal P1 fields between tetrahedral meshes of one domain.

(New functionality: the reference has no mesh-to-mesh transfer.)  Used for
grid sequencing: a potential converged on a coarse mesh, interpolated
linearly onto the nodes of a finer mesh, seeds the fine-mesh Newton solve.
"""

from __future__ import annotations

import numpy as np


def locate_points(nodes, elements, points, face_neighbors=None):
    """(element id (m,), barycentric coordinates (m, 4)) of each point in
    the tet mesh (nodes, elements): a native visibility walk over the face
    graph from the previous point's element; points outside the mesh get
    the boundary tet where the walk stopped."""
    from .. import _native
    from .metrics import build_face_adjacency

    nd = np.ascontiguousarray(nodes, dtype=np.float64)
    el = np.ascontiguousarray(elements, dtype=np.int32)
    pts = np.ascontiguousarray(points, dtype=np.float64)
    if face_neighbors is None:
        face_neighbors, _ = build_face_adjacency(el)
    fn = np.ascontiguousarray(face_neighbors, dtype=np.int32)
    elem = np.empty(len(pts), dtype=np.int32)
    bary = np.empty((len(pts), 4))
    _native.check(_native.lib().fp3d_host_locate_points(
        len(el), el.ctypes.data, nd.ctypes.data, fn.ctypes.data, len(pts),
        pts.ctypes.data, elem.ctypes.data, bary.ctypes.data), "locate_points")
    return elem, bary


def interpolate_p1(nodes, elements, values, points) -> np.ndarray:
    """Piecewise-linear interpolant of the nodal `values` on the tet mesh
    (nodes, elements) evaluated at `points` (points just outside the mesh
    extrapolate linearly from the nearest boundary tet on the walk)."""
    elem, bary = locate_points(nodes, elements, points)
    el = np.asarray(elements, dtype=np.int64)
    v = np.asarray(values, dtype=np.float64)
    return np.einsum("mi,mi->m", bary, v[el[elem]])

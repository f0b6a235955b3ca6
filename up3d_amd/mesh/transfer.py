"""Transfer of nodal P1 fields between tetrahedral meshes of one domain.

(New functionality: the reference has no mesh-to-mesh transfer.)  Used for
grid sequencing: a potential converged on a coarse mesh, interpolated
linearly onto the nodes of a finer mesh, seeds the fine-mesh Newton solve.
"""

from __future__ import annotations

from collections import OrderedDict

import numpy as np

# point locations per (source mesh, points) content, most recent last: a
# sweep transfers every condition from one companion mesh onto the same
# target nodes
_LOCATIONS: "OrderedDict[str, tuple]" = OrderedDict()
_MAX_LOCATIONS = 4


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


def locate_points_cached(nodes, elements, points):
    """locate_points, kept for the last few (mesh, points) contents (the
    same arrays for the same inputs; callers must not modify them)."""
    from ..kernels.store import array_digest

    nd = np.ascontiguousarray(nodes, dtype=np.float64)
    el = np.ascontiguousarray(elements, dtype=np.int32)
    pts = np.ascontiguousarray(points, dtype=np.float64)
    key = array_digest(np.int64(len(nd)), np.int64(len(el)), nd, el, pts)
    hit = _LOCATIONS.pop(key, None)
    if hit is None:
        hit = locate_points(nd, el, pts)
        hit[0].setflags(write=False)
        hit[1].setflags(write=False)
    _LOCATIONS[key] = hit
    while len(_LOCATIONS) > _MAX_LOCATIONS:
        _LOCATIONS.popitem(last=False)
    return hit


def interpolate_p1(nodes, elements, values, points) -> np.ndarray:
    """Piecewise-linear interpolant of the nodal `values` on the tet mesh
    (nodes, elements) evaluated at `points` (points just outside the mesh
    extrapolate linearly from the nearest boundary tet on the walk)."""
    elem, bary = locate_points_cached(nodes, elements, points)
    el = np.asarray(elements, dtype=np.int64)
    v = np.asarray(values, dtype=np.float64)
    return np.einsum("mi,mi->m", bary, v[el[elem]])

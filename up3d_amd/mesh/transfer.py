"""Transfer of nodal P1 fields between tetrahedral meshes of one domain.

(New functionality: the reference has no mesh-to-mesh transfer.)  Used for
grid sequencing: a potential converged on a coarse mesh, interpolated
linearly onto the nodes of a finer mesh, seeds the fine-mesh Newton solve.
"""

from __future__ import annotations

import numpy as np
from scipy.spatial import cKDTree


def _barycentric(nodes, elements, cand, pts):
    """Barycentric coordinates of pts (m, 3) in candidate tets cand (m, k)."""
    p0 = nodes[elements[cand, 0]]                       # (m, k, 3)
    T = np.stack([nodes[elements[cand, i]] - p0 for i in (1, 2, 3)], axis=-1)
    rhs = pts[:, None, :] - p0
    lam123 = np.linalg.solve(T, rhs[..., None])[..., 0]  # (m, k, 3)
    lam0 = 1.0 - lam123.sum(-1, keepdims=True)
    return np.concatenate([lam0, lam123], axis=-1)      # (m, k, 4)


def interpolate_p1(nodes, elements, values, points, k: int = 16,
                   chunk: int = 65536) -> np.ndarray:
    """Piecewise-linear interpolant of the nodal `values` on the tet mesh
    (nodes, elements) evaluated at `points`.  Each point is located among
    the k tets with the nearest centroids; a point outside all of them
    (mesh-boundary jitter) takes the candidate whose smallest barycentric
    coordinate is largest, i.e. a short linear extrapolation."""
    nodes = np.asarray(nodes, dtype=np.float64)
    elements = np.asarray(elements, dtype=np.int64)
    values = np.asarray(values, dtype=np.float64)
    points = np.asarray(points, dtype=np.float64)
    cen = nodes[elements].mean(axis=1)
    tree = cKDTree(cen)
    out = np.empty(len(points))
    for s in range(0, len(points), chunk):
        pts = points[s:s + chunk]
        _, cand = tree.query(pts, k=min(k, len(elements)))
        cand = cand.reshape(len(pts), -1)
        lam = _barycentric(nodes, elements, cand, pts)
        best = np.argmax(lam.min(axis=-1), axis=1)
        rows = np.arange(len(pts))
        e = cand[rows, best]
        lb = lam[rows, best]
        out[s:s + chunk] = np.einsum("mi,mi->m", lb, values[elements[e]])
    return out

"""Level-set description of a straight ruled wake sheet
(pyfp3d/wake/levelset.py:50-225, WakeLevelSet).

The sheet is swept from a trailing-edge polyline along a ruling direction g.
For a point x and each TE panel (segment a -> a + v) the point is written in
the panel's oblique frame

    x - a = u v + d g + s n,     n = unit(v x g)

(a 2x2 solve in the (v, g) plane; n is orthogonal to both).  The panel whose
foot is nearest wins, measured by s^2 + (spanwise excess past the panel)^2 +
(downstream shortfall min(0, d))^2; the point then gets

    s  signed offset from the sheet, "+" = the n side,
    d  downstream distance from the TE along g,
    q  spanwise arclength along the TE polyline, NOT clamped (q < 0 or
       q > span_length: off the ends of the sheet).

`direction` is the physical wake (convection) direction; `sheet_direction`
optionally decouples the geometric ruling from it.  Host numpy, evaluated
once per mesh (not a hot kernel).
"""

from __future__ import annotations

from typing import Optional, Tuple

import numpy as np


class WakeLevelSet:
    def __init__(self, te_points, direction, span_hint=(0.0, 0.0, 1.0),
                 extent: Optional[float] = None,
                 sheet_direction: Optional[np.ndarray] = None):
        pts = np.atleast_2d(np.asarray(te_points, dtype=np.float64))
        if pts.shape[1] != 3:
            raise ValueError("te_points must be (m, 3)")
        if len(pts) == 1:
            # a single TE point: a long synthetic span segment through it
            t = np.asarray(span_hint, dtype=np.float64)
            t = t / np.linalg.norm(t)
            pts = np.vstack([pts[0] - 1e3 * t, pts[0] + 1e3 * t])
        seg = pts[1:] - pts[:-1]
        seg_len2 = np.einsum("ij,ij->i", seg, seg)
        if np.any(seg_len2 <= 0.0):
            raise ValueError("degenerate TE segment (repeated points)")
        if sheet_direction is not None:
            sheet_direction = np.asarray(sheet_direction, dtype=np.float64)
            if sheet_direction.shape != (3,) or not np.any(sheet_direction):
                raise ValueError("sheet_direction must be a nonzero (3,) vector")
        self.te_points = pts
        self.extent = extent
        self._sheet_dir = sheet_direction
        self._a = pts[:-1]
        self._v = seg
        self._v2 = seg_len2
        self._vlen = np.sqrt(seg_len2)
        self._q0 = np.concatenate([[0.0], np.cumsum(self._vlen)[:-1]])
        self.span_length = float(self._vlen.sum())
        self.update_direction(direction)

    @property
    def direction(self) -> np.ndarray:
        return self._dir.copy()

    @property
    def sheet_direction(self):
        return None if self._sheet_dir is None else self._sheet_dir.copy()

    def update_direction(self, new_direction) -> None:
        """Re-aim the sheet (an alpha change); the mesh is untouched."""
        d = np.asarray(new_direction, dtype=np.float64)
        nd = np.linalg.norm(d)
        if nd == 0.0:
            raise ValueError("direction must be nonzero")
        self._dir = d / nd
        g = self._dir if self._sheet_dir is None else \
            self._sheet_dir / np.linalg.norm(self._sheet_dir)
        self._g = g
        n = np.cross(self._v, g[None, :])
        nn = np.linalg.norm(n, axis=1)
        if np.any(nn < 1e-12 * self._vlen):
            raise ValueError("wake direction (near-)parallel to a TE segment")
        self._n = n / nn[:, None]
        self._vg = self._v @ g                  # v . g
        self._det = self._v2 - self._vg ** 2    # Gram determinant of (v, g)

    def _frame(self, x):
        rel = x[:, None, :] - self._a[None, :, :]            # (p, nseg, 3)
        rv = np.einsum("pnk,nk->pn", rel, self._v)
        rg = rel @ self._g
        u = (rv - self._vg * rg) / self._det
        d = (self._v2 * rg - self._vg * rv) / self._det
        s = np.einsum("pnk,nk->pn", rel, self._n)
        out = np.maximum(0.0, np.maximum(-u, u - 1.0)) * self._vlen
        behind = np.minimum(0.0, d)
        best = np.argmin(s * s + out * out + behind * behind, axis=1)
        return u, d, s, best

    def evaluate(self, points) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(s, d, q) at points (n, 3)."""
        x = np.atleast_2d(np.asarray(points, dtype=np.float64))
        u, d, s, best = self._frame(x)
        i = np.arange(len(x))
        q = self._q0[best] + u[i, best] * self._vlen[best]
        return s[i, best], d[i, best], q

    def surface_normals(self, points) -> np.ndarray:
        """Unit "+" normal of the panel `evaluate` selects for each point."""
        x = np.atleast_2d(np.asarray(points, dtype=np.float64))
        _, _, _, best = self._frame(x)
        return self._n[best]

"""Spanwise sections (pyfp3d/post/section_cut.py:26-430): SectionData,
section_cut (a plane x_s = const through the volume mesh: the symmetry-
plane triangulation when the plane is one, else marching tetrahedra with
P1 interpolation along cut edges), and section_cp_curve (the sectional wall
Cp(x/c) from the wall triangles the plane crosses, each contributing its
own constant Cp at the midpoint of its crossing segment).

span_axis (no reference counterpart; default 2 = the reference's z) names
the spanwise coordinate, so the embedded plate (span along y) is cut the
same way.  The marching and crossing loops are vectorised over elements."""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np

from .surface import cp_from_q2, triangle_tangential_gradients

GAMMA = 1.4
_TET_EDGES = np.array([(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)])
_TRI_EDGES = np.array([(0, 1), (0, 2), (1, 2)])


@dataclass
class SectionData:
    """A spanwise section: in-plane 2D points, triangles, restricted fields."""
    points2d: np.ndarray
    triangles: np.ndarray
    fields: Dict[str, np.ndarray] = field(default_factory=dict)
    z: float = 0.0


def _nudged(coord, z):
    s = coord - z
    extent = float(np.ptp(coord))
    eps = 1e-12 * (extent if extent > 0 else 1.0)
    return np.where(np.abs(s) < eps, eps, s)


def _inplane(span_axis):
    return [a for a in range(3) if a != span_axis]


def section_cut(mesh, point_fields: Dict[str, np.ndarray], z: float,
                atol: Optional[float] = None, span_axis: int = 2) -> SectionData:
    """The section x_span = z of nodal fields (exact for P1 fields)."""
    nodes = np.asarray(mesh.nodes, dtype=np.float64)
    ax = _inplane(span_axis)
    c = nodes[:, span_axis]
    if atol is None:
        extent = float(np.ptp(c))
        atol = 1e-9 * extent if extent > 0 else 1e-12
    on_plane = np.abs(c - z) <= atol
    sym = getattr(mesh, "boundary_faces", {}).get("symmetry")
    if np.any(on_plane) and sym is not None:
        tri = np.asarray(sym, dtype=np.int64)
        tri = tri[np.all(on_plane[tri], axis=1)]
        if len(tri):
            keep = np.unique(tri)
            remap = np.full(len(nodes), -1, dtype=np.int64)
            remap[keep] = np.arange(len(keep))
            return SectionData(points2d=nodes[keep][:, ax].copy(), triangles=remap[tri],
                               fields={k: np.asarray(f)[keep].copy()
                                       for k, f in point_fields.items()}, z=float(z))
    return _marching(mesh, point_fields, z, span_axis)


def _marching(mesh, point_fields, z, span_axis):
    nodes = np.asarray(mesh.nodes, dtype=np.float64)
    el = np.asarray(mesh.elements, dtype=np.int64)
    ax = _inplane(span_axis)
    s = _nudged(nodes[:, span_axis], z)
    se = s[el]
    npos = (se > 0).sum(axis=1)
    cut = np.flatnonzero((npos > 0) & (npos < 4))
    if not len(cut):
        raise ValueError(f"plane {z} does not intersect the mesh")
    names = list(point_fields)
    F = (np.column_stack([np.asarray(point_fields[n], dtype=np.float64) for n in names])
         if names else np.zeros((len(nodes), 0)))
    tets = el[cut]
    sa = se[cut][:, _TET_EDGES[:, 0]]
    sb = se[cut][:, _TET_EDGES[:, 1]]
    crossed = (sa > 0) != (sb > 0)                      # (n, 6)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(crossed, sa / (sa - sb), 0.0)
    na, nb = tets[:, _TET_EDGES[:, 0]], tets[:, _TET_EDGES[:, 1]]
    P = nodes[na][:, :, ax] * (1 - t)[..., None] + nodes[nb][:, :, ax] * t[..., None]
    Fv = F[na] * (1 - t)[..., None] + F[nb] * t[..., None]
    ncut = crossed.sum(axis=1)
    pts, vals, tris = [], [], []
    base = 0
    for k in (3, 4):
        sel = np.flatnonzero(ncut == k)
        if not len(sel):
            continue
        cr = crossed[sel]
        order = np.argsort(~cr, axis=1, kind="stable")[:, :k]       # crossed edges first
        p = np.take_along_axis(P[sel], order[..., None], axis=1)
        v = np.take_along_axis(Fv[sel], order[..., None], axis=1)
        if k == 4:                                   # order the quad around its centroid
            d = p - p.mean(axis=1, keepdims=True)
            ang = np.arctan2(d[..., 1], d[..., 0])
            o2 = np.argsort(ang, axis=1)
            p = np.take_along_axis(p, o2[..., None], axis=1)
            v = np.take_along_axis(v, o2[..., None], axis=1)
            local = np.array([(0, 1, 2), (0, 2, 3)])
        else:
            local = np.array([(0, 1, 2)])
        m = len(sel)
        idx = base + k * np.arange(m)[:, None, None] + local[None]
        tris.append(idx.reshape(-1, 3))
        pts.append(p.reshape(-1, 2))
        vals.append(v.reshape(-1, F.shape[1]))
        base += k * m
    pts = np.concatenate(pts)
    vals = np.concatenate(vals)
    return SectionData(points2d=pts, triangles=np.concatenate(tris),
                       fields={n: vals[:, j].copy() for j, n in enumerate(names)},
                       z=float(z))


def wall_plane_crossings(nodes, wall_faces, z, span_axis: int = 2):
    """(indices into wall_faces, (n, 3) midpoints of the crossing segments)
    of the wall triangles the plane x_span = z crosses."""
    nodes = np.asarray(nodes, dtype=np.float64)
    wf = np.asarray(wall_faces, dtype=np.int64)
    s = _nudged(nodes[:, span_axis], z)[wf]
    npos = (s > 0).sum(axis=1)
    hit = np.flatnonzero((npos > 0) & (npos < 3))
    if not len(hit):
        return np.zeros(0, dtype=np.int64), np.zeros((0, 3))
    sa = s[hit][:, _TRI_EDGES[:, 0]]
    sb = s[hit][:, _TRI_EDGES[:, 1]]
    crossed = (sa > 0) != (sb > 0)
    t = np.where(crossed, sa / np.where(crossed, sa - sb, 1.0), 0.0)
    tri = wf[hit]
    pa = nodes[tri[:, _TRI_EDGES[:, 0]]]
    pb = nodes[tri[:, _TRI_EDGES[:, 1]]]
    p = pa * (1 - t)[..., None] + pb * t[..., None]
    mid = (p * crossed[..., None]).sum(axis=1) / 2.0
    return hit, mid


def section_cp_curve(mesh, phi, *, eta: Optional[float] = None,
                     z: Optional[float] = None, b_semi: Optional[float] = None,
                     u_inf: float = 1.0, m_inf: float = 0.0,
                     wall_tag: str = "wall", upper_hint=(0.0, 1.0, 0.0),
                     min_points_per_side: int = 5, smooth_passes: int = 0,
                     gamma: float = GAMMA, span_axis: int = 2,
                     chord_axis: int = 0, side_by: str = "position"
                     ) -> Dict[str, np.ndarray]:
    """Sectional wall Cp(x/c) at z = eta * b_semi (or z): x_upper, cp_upper,
    x_lower, cp_lower, chord, x_le, z, eta; upper = midpoints with
    dot(mid, upper_hint) > 0 (side_by="position", the reference's rule), or
    wall triangles whose fluid-outward normal points against upper_hint
    (side_by="normal", for a zero-thickness wall whose two faces share
    positions).  ValueError as the reference raises."""
    if smooth_passes:
        raise NotImplementedError("smooth_passes > 0 (recovered wall gradients) is not built")
    if (eta is None) == (z is None):
        raise ValueError("pass exactly one of eta or z")
    if eta is not None:
        if b_semi is None:
            raise ValueError("b_semi is required when eta is given")
        z = float(eta) * float(b_semi)
    z = float(z)
    wall = np.asarray(mesh.boundary_faces[wall_tag], dtype=np.int64)
    g, _, _ = triangle_tangential_gradients(mesh.nodes, wall, np.asarray(phi))
    hit, mid = wall_plane_crossings(mesh.nodes, wall, z, span_axis)
    q2 = np.einsum("ij,ij->i", g[hit], g[hit]) / u_inf ** 2
    cps = cp_from_q2(q2, m_inf, gamma)
    hint = np.asarray(upper_hint, dtype=np.float64)
    if side_by == "position":
        sides = mid @ hint > 0.0
    elif side_by == "normal":
        p = np.asarray(mesh.nodes, dtype=np.float64)[wall[hit]]
        nrm = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
        sides = nrm @ hint < 0.0
    else:
        raise ValueError(f"side_by={side_by!r}: 'position' or 'normal'")
    n_up, n_lo = int(sides.sum()), int((~sides).sum())
    if n_up < min_points_per_side or n_lo < min_points_per_side:
        raise ValueError(f"section at z={z:.4f} too sparse (upper={n_up}, lower={n_lo}); "
                         "the plane likely missed the wing or hit the flat tip cap")
    xs = mid[:, chord_axis]
    x_le = float(xs.min())
    chord = float(xs.max() - xs.min())
    xn = (xs - x_le) / chord
    iu = np.argsort(xn[sides], kind="stable")
    il = np.argsort(xn[~sides], kind="stable")
    return {"x_upper": xn[sides][iu], "cp_upper": cps[sides][iu],
            "x_lower": xn[~sides][il], "cp_lower": cps[~sides][il],
            "chord": chord, "x_le": x_le, "z": z,
            "eta": (z / float(b_semi)) if b_semi else None}

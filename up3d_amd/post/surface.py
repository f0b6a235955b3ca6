"""Wall pressure and integrated force (pyfp3d/post/surface.py:80-118,
192-231, 265-363): the output functional of the solve.

Per wall triangle: the in-plane (tangential) P1 gradient of phi -- the full
wall velocity under the slip condition -- then q2 = |g|^2 / u_inf^2, the
exact isentropic Cp (Bernoulli at M_inf = 0) and
cf = -sum(Cp A n_out) / S_ref with n_out pointing out of the body (into
the fluid, found from the tet that owns the face).  Host numpy: the wall
set is a 2-D surface (24,576 triangles on the 4.7M-tet channel).
"""

from __future__ import annotations

from collections import OrderedDict

import numpy as np

from ..physics.isentropic import GAMMA


def triangle_tangential_gradients(nodes, wall_faces, phi):
    """(grad_tri (n,3), area (n,), unit normal of the stored winding)."""
    nodes = np.asarray(nodes, dtype=np.float64)
    wf = np.asarray(wall_faces, dtype=np.int64)
    phi = np.asarray(phi, dtype=np.float64)
    p0, p1, p2 = nodes[wf[:, 0]], nodes[wf[:, 1]], nodes[wf[:, 2]]
    e_a = p1 - p0
    e_b = p2 - p0
    nvec = np.cross(e_a, e_b)
    twice = np.linalg.norm(nvec, axis=1)
    nrm = nvec / twice[:, None]
    t1 = e_a / np.linalg.norm(e_a, axis=1)[:, None]
    t2 = np.cross(nrm, t1)
    a1, b1 = np.einsum("ij,ij->i", e_a, t1), np.einsum("ij,ij->i", e_a, t2)
    a2, b2 = np.einsum("ij,ij->i", e_b, t1), np.einsum("ij,ij->i", e_b, t2)
    f1 = phi[wf[:, 1]] - phi[wf[:, 0]]
    f2 = phi[wf[:, 2]] - phi[wf[:, 0]]
    det = a1 * b2 - b1 * a2
    g1 = (b2 * f1 - b1 * f2) / det
    g2 = (a1 * f2 - a2 * f1) / det
    return g1[:, None] * t1 + g2[:, None] * t2, 0.5 * twice, nrm


def wall_outward_normals(nodes, elements, wall_faces):
    """Unit wall normals pointing into the fluid (the owning tet's side);
    ValueError if a face is not owned by exactly one tet."""
    nodes = np.asarray(nodes, dtype=np.float64)
    el_all = np.asarray(elements)
    wf = np.asarray(wall_faces, dtype=np.int64)
    n = int(max(el_all.max(), wf.max())) + 1

    def keys(tri):
        s = np.sort(tri, axis=1)
        return (s[:, 0] * n + s[:, 1]) * n + s[:, 2]

    # only a tet with three wall nodes can own (or share) a wall face: key
    # the faces of those tets, not all 4 N_t faces
    on_wall = np.zeros(n, dtype=bool)
    on_wall[wf.ravel()] = True
    cand = np.flatnonzero(on_wall[el_all].sum(axis=1) >= 3)
    el = np.asarray(el_all[cand], dtype=np.int64)
    faces = np.concatenate([el[:, [1, 2, 3]], el[:, [0, 2, 3]],
                            el[:, [0, 1, 3]], el[:, [0, 1, 2]]])
    owner = np.tile(np.arange(len(el)), 4)
    k = keys(faces)
    order = np.argsort(k, kind="stable")
    ks, os_ = k[order], owner[order]
    q = keys(wf)
    lo = np.searchsorted(ks, q, side="left")
    hi = np.searchsorted(ks, q, side="right")
    if np.any(hi - lo != 1):
        raise ValueError(f"{int(np.sum(hi - lo != 1))} wall face(s) not owned "
                         "by exactly one tet")
    own = os_[lo]
    p0 = nodes[wf[:, 0]]
    nv = np.cross(nodes[wf[:, 1]] - p0, nodes[wf[:, 2]] - p0)
    nrm = nv / np.linalg.norm(nv, axis=1)[:, None]
    fc = nodes[wf].mean(axis=1)
    oc = nodes[el[own]].mean(axis=1)
    flip = np.einsum("ij,ij->i", nrm, oc - fc) < 0.0
    nrm[flip] *= -1.0
    return nrm


# outward normals per (nodes, elements, wall faces) content, most recent
# last: a sweep evaluates the force of every condition on a few meshes
_NORMALS: "OrderedDict[str, np.ndarray]" = OrderedDict()
_MAX_NORMALS = 8


def _outward_normals_cached(nodes, elements, wall_faces) -> np.ndarray:
    from ..kernels.store import array_digest

    nd = np.ascontiguousarray(nodes, dtype=np.float64)
    el = np.ascontiguousarray(elements)
    wf = np.ascontiguousarray(wall_faces, dtype=np.int64)
    key = array_digest(np.int64(len(nd)), np.int64(len(el)), np.int64(el.dtype.itemsize),
                       nd, el, wf)
    nrm = _NORMALS.pop(key, None)
    if nrm is None:
        nrm = wall_outward_normals(nd, el, wf)
        nrm.setflags(write=False)
    _NORMALS[key] = nrm
    while len(_NORMALS) > _MAX_NORMALS:
        _NORMALS.popitem(last=False)
    return nrm


def cp_from_q2(q2, m_inf, gamma=GAMMA):
    """Exact isentropic Cp (incompressible 1 - q2 at M_inf = 0)."""
    q2 = np.asarray(q2, dtype=np.float64)
    if m_inf > 0.0:
        base = 1.0 + 0.5 * (gamma - 1.0) * (m_inf * m_inf) * (1.0 - q2)
        base = np.maximum(base, 1e-10)
        rho = base ** (1.0 / (gamma - 1.0))
        return 2.0 / (gamma * m_inf ** 2) * (rho ** gamma - 1.0)
    return 1.0 - q2


def wall_force_coefficients(nodes, elements, wall_faces, phi, alpha_deg=0.0,
                            u_inf=1.0, s_ref=1.0, m_inf=0.0):
    """dict(cl, cd_pressure, cf (3,), cp_tri (n,))."""
    g, area, _ = triangle_tangential_gradients(nodes, wall_faces, phi)
    n_out = _outward_normals_cached(nodes, elements, wall_faces)
    q2 = np.einsum("ij,ij->i", g, g) / u_inf ** 2
    cp = cp_from_q2(q2, m_inf)
    cf = -(cp * area) @ n_out / s_ref
    a = np.deg2rad(alpha_deg)
    lift = np.array([-np.sin(a), np.cos(a), 0.0])
    drag = np.array([np.cos(a), np.sin(a), 0.0])
    return {"cl": float(cf @ lift), "cd_pressure": float(cf @ drag), "cf": cf,
            "cp_tri": cp}

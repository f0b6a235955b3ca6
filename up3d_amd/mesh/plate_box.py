"""Box mesh around an embedded zero-thickness swept flat-plate wing.

Half model: the symmetry plane y = 0 is a slip wall (natural boundary
condition), the plate has chord c, half-span b (full span 2b) and leading-
edge sweep angle L, and lies in the mid plane z = 0:

    plate = { (x, y, 0) : y t <= x <= y t + c,  0 <= y <= b },  t = tan L.

Grid: n_xi x n_y x n_z hexes (6 Kuhn tets each, node numbering of
`channel.kuhn_box`) in the sheared coordinates (xi, y, z) with
x = xi + t min(y, b), so the leading edge (xi = 0), the trailing edge
(xi = c) and the tip (y = b) are grid lines and the plate is a set of
grid faces.  Spacing: uniform over the chord and the half-span, and
geometrically stretched toward the outer boundaries; in z the first cell
off the plate is as tall as a chordwise cell and the cells stretch to
the top and bottom.

The plate is a slip wall of zero thickness: every plate node strictly
inside the planform (not on the leading edge, the trailing edge or the
tip) is duplicated, and the tets below the plate read the duplicate, so
the two faces of the plate carry independent potentials and the face
graph does not connect them.  (One triangle of the leading-edge / tip
corner cell has all three nodes on the planform's edges and so stays
shared by the tets above and below it.)  The trailing edge stays single-valued: the
potential jump there is the level-set wake's (its TE nodes and their aux
DOFs).

Boundary tags: "inflow" (xi = xi_min), "outflow" (xi = xi_max), "side"
(y = y_max), "bottom" / "top" (z = -+z_max), "farfield" (inflow, side,
bottom and top: the Dirichlet part; the wake sheet leaves through the
outflow, which takes a Neumann freestream flux), "symmetry" (y = 0),
"wall" (both faces of the plate, outward from the fluid).  `lower_nodes` lists the duplicated (lower-face) node ids; they
sit on the plate at z = 0, so a wake level set must be told they belong
to the "-" side.
"""

from __future__ import annotations

from types import SimpleNamespace

import numpy as np

from .channel import _FACE_DEFS, _boundary_faces, kuhn_box


def _stretched(length: float, n: int, h0: float) -> np.ndarray:
    """n cell widths summing to `length`, the first h0, growing
    geometrically (uniform when n h0 >= length)."""
    if n <= 0:
        raise ValueError("a stretched segment needs at least one cell")
    if n * h0 >= length:
        return np.full(n, length / n)
    lo, hi = 1.0, 2.0
    while h0 * (hi ** n - 1.0) / (hi - 1.0) < length:
        hi *= 2.0
    for _ in range(200):
        r = 0.5 * (lo + hi)
        if h0 * (r ** n - 1.0) / (r - 1.0) < length:
            lo = r
        else:
            hi = r
    r = 0.5 * (lo + hi)
    w = h0 * r ** np.arange(n)
    return w * (length / w.sum())


def _axis(segments):
    """Concatenate (length, n, h_fine, fine_at_start) segments into node
    coordinates starting at 0."""
    widths = []
    for L, n, h, fine_first in segments:
        w = _stretched(L, n, h)
        widths.append(w if fine_first else w[::-1])
    return np.concatenate([[0.0], np.cumsum(np.concatenate(widths))])


def plate_wing_box(n_xi: int = 128, n_y: int = 96, n_z: int = 96, *,
                   chord: float = 1.0, half_span: float = 1.0,
                   sweep_deg: float = 30.0, upstream: float = 4.0,
                   downstream: float = 5.0, y_max: float = 4.0,
                   z_max: float = 4.0, chord_cells=None, span_cells=None):
    """Kuhn-split box around the embedded plate (module doc).  chord_cells /
    span_cells default to 5/16 of n_xi and 1/2 of n_y; n_z must be even."""
    if n_z % 2:
        raise ValueError("n_z must be even (the plate lies on the mid plane)")
    nc = int(chord_cells) if chord_cells is not None else max(2, round(n_xi * 5 / 16))
    ns = int(span_cells) if span_cells is not None else max(2, n_y // 2)
    n_up = (n_xi - nc) * 5 // 11
    n_dn = n_xi - nc - n_up
    if min(n_up, n_dn, n_y - ns) < 1:
        raise ValueError("grid too small for the plate box")
    hc = chord / nc
    hs = half_span / ns
    xi = _axis([(upstream, n_up, hc, False), (chord, nc, hc, True),
                (downstream, n_dn, hc, True)]) - upstream
    yy = _axis([(half_span, ns, hs, True), (y_max - half_span, n_y - ns, hs, True)])
    zh = n_z // 2
    zz = _axis([(z_max, zh, hc, False), (z_max, zh, hc, True)]) - z_max
    i_le, i_te, j_tip, k0 = n_up, n_up + nc, ns, zh
    xi[i_le], xi[i_te], yy[j_tip], zz[k0] = 0.0, chord, half_span, 0.0

    ijk, el = kuhn_box(n_xi, n_y, n_z)
    t = np.tan(np.deg2rad(sweep_deg))
    X, Y, Z = xi[ijk[:, 0]], yy[ijk[:, 1]], zz[ijk[:, 2]]
    nodes = np.stack([X + t * np.minimum(Y, half_span), Y, Z], axis=1)
    n0 = len(nodes)
    i, j, k = ijk[:, 0], ijk[:, 1], ijk[:, 2]
    on_plate = (k == k0) & (i >= i_le) & (i <= i_te) & (j <= j_tip)
    interior = on_plate & (i > i_le) & (i < i_te) & (j < j_tip)
    crack = np.flatnonzero(interior)
    lower = n0 + np.arange(len(crack))
    remap = np.arange(n0 + len(crack))
    remap[crack] = lower
    below = nodes[el, 2].mean(axis=1) < 0.0
    el = el.copy()
    el[below] = remap[el[below]]
    nodes = np.vstack([nodes, nodes[crack]])
    el = el.astype(np.int32)

    def plane(mask0):
        return np.concatenate([mask0, mask0[crack]])

    wall_mask = plane(on_plate)
    faces = {
        "inflow": _boundary_faces(el, nodes, plane(i == 0)),
        "outflow": _boundary_faces(el, nodes, plane(i == n_xi)),
        "side": _boundary_faces(el, nodes, plane(j == n_y)),
        "bottom": _boundary_faces(el, nodes, plane(k == 0)),
        "top": _boundary_faces(el, nodes, plane(k == n_z)),
        "symmetry": _boundary_faces(el, nodes, plane(j == 0)),
        "wall": _plate_faces(el, nodes, wall_mask,
                             np.concatenate([interior, np.ones(len(crack), dtype=bool)])),
    }
    faces["farfield"] = np.concatenate([faces[t] for t in ("inflow", "side", "bottom", "top")])
    te_nodes = np.flatnonzero(plane((k == k0) & (i == i_te) & (j <= j_tip))[:n0])
    te_tip = np.array([t * half_span + chord, half_span, 0.0])
    te_root = np.array([chord, 0.0, 0.0])
    return SimpleNamespace(
        nodes=np.ascontiguousarray(nodes), elements=el, boundary_faces=faces,
        lower_nodes=lower, upper_nodes=crack, wall_nodes=np.flatnonzero(wall_mask),
        te_nodes=te_nodes, te_line=np.vstack([te_tip, te_root]),
        shape=(n_xi, n_y, n_z),
        meta={"chord": chord, "half_span": half_span, "sweep_deg": sweep_deg,
              "area_full": 2.0 * chord * half_span, "i_le": i_le, "i_te": i_te,
              "j_tip": j_tip, "k0": k0})


def _plate_faces(el, nodes, on_wall, cracked):
    """Tet faces lying on the plate (all three nodes plate nodes, at least
    one of them duplicated), outward from their tet.  A face whose three
    nodes all lie on the planform's edges is shared by the tets above and
    below (the one such triangle sits in the leading-edge / tip corner
    cell): it is not a wall face."""
    out = []
    for f, fd in enumerate(_FACE_DEFS):
        tri = el[:, fd]
        m = on_wall[tri].all(axis=1) & cracked[tri].any(axis=1)
        if not m.any():
            continue
        tri = tri[m]
        apex = el[m, f]
        p0, p1, p2 = nodes[tri[:, 0]], nodes[tri[:, 1]], nodes[tri[:, 2]]
        nrm = np.cross(p1 - p0, p2 - p0)
        inward = np.einsum("ij,ij->i", nrm, nodes[apex] - p0) > 0.0
        tri = tri.copy()
        tri[inward] = tri[inward][:, [0, 2, 1]]
        out.append(tri)
    if not out:
        return np.empty((0, 3), dtype=np.int32)
    return np.concatenate(out).astype(np.int32)

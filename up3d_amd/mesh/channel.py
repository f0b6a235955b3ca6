"""Synthetic bump-channel meshes for the full-potential solve.

Channel [0, Lx] x [0, 1] x [0, Lz] (x streamwise, y vertical, z spanwise,
the reference's axis convention: alpha rotates x-y,
pyfp3d/constraints/dirichlet.py:29-34).  A circular-arc bump of height
`bump_height` and chord 1 sits on the floor y = 0 over x in [1, 2]:

    R = (1/4 + t^2) / (2 t),   y_b(x) = sqrt(R^2 - (x - 1.5)^2) - (R - t)

and grid nodes are mapped with the linear vertical blend
y' = y + y_b(x) (1 - y).  Each hex is split into 6 tets with the Kuhn
(Freudenthal) permutation order of tests/mesh_utils.py:16-56 (axis order
x, y, z; node id i*(ny+1)*(nz+1) + j*(nz+1) + k).  Optional: interior
nodes jittered by U(-j, j)*h per axis from a seed, then the nodes are
renumbered by reverse Cuthill-McKee and the elements sorted by their
sorted new node tuples.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import reverse_cuthill_mckee

_PERMS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))
_FACE_DEFS = ((1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2))


@dataclass
class ChannelMesh:
    nodes: np.ndarray                      # (n_nodes, 3) float64
    elements: np.ndarray                   # (n_tets, 4) int32
    boundary_faces: Dict[str, np.ndarray]  # tag -> (n_f, 3) int32, outward
    shape: tuple
    node_order: Optional[np.ndarray] = None  # new id -> grid id (RCM)
    meta: dict = field(default_factory=dict)

    @property
    def n_nodes(self):
        return len(self.nodes)

    @property
    def n_tets(self):
        return len(self.elements)


def kuhn_box(nx: int, ny: int, nz: int):
    """Grid indices (i, j, k) per node and Kuhn-split elements, in the
    reference's element order (hexes i-major, then the 6 permutations)."""
    ii, jj, kk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1),
                             np.arange(nz + 1), indexing="ij")
    ijk = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1)

    def idx(a):
        return a[..., 0] * (ny + 1) * (nz + 1) + a[..., 1] * (nz + 1) + a[..., 2]

    hi, hj, hk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz),
                             indexing="ij")
    base = np.stack([hi.ravel(), hj.ravel(), hk.ravel()], axis=1)  # (nh, 3)
    unit = np.eye(3, dtype=np.int64)
    tets = np.empty((len(base), 6, 4), dtype=np.int64)
    for p, perm in enumerate(_PERMS):
        v0 = base
        v1 = v0 + unit[perm[0]]
        v2 = v1 + unit[perm[1]]
        v3 = v2 + unit[perm[2]]
        tets[:, p, 0] = idx(v0)
        tets[:, p, 1] = idx(v1)
        tets[:, p, 2] = idx(v2)
        tets[:, p, 3] = idx(v3)
    return ijk, tets.reshape(-1, 4)


def bump_profile(x: np.ndarray, height: float = 0.1, x0: float = 1.0,
                 chord: float = 1.0) -> np.ndarray:
    """Circular-arc bump ordinate y_b(x) (0 outside the chord)."""
    if height == 0.0:
        return np.zeros_like(x)
    half = 0.5 * chord
    R = (half * half + height * height) / (2.0 * height)
    xc = x0 + half
    yb = np.sqrt(np.maximum(R * R - (x - xc) ** 2, 0.0)) - (R - height)
    return np.where((x >= x0) & (x <= x0 + chord), np.maximum(yb, 0.0), 0.0)


def _boundary_faces(elements: np.ndarray, nodes: np.ndarray, on_plane: np.ndarray):
    """Faces whose three nodes all lie on the plane flagged by on_plane
    (bool per node), oriented with outward normals."""
    out = []
    for f, fd in enumerate(_FACE_DEFS):
        tri = elements[:, fd]
        m = on_plane[tri].all(axis=1)
        if not m.any():
            continue
        tri = tri[m]
        apex = elements[m, f]
        p0, p1, p2 = nodes[tri[:, 0]], nodes[tri[:, 1]], nodes[tri[:, 2]]
        nrm = np.cross(p1 - p0, p2 - p0)
        inward = np.einsum("ij,ij->i", nrm, nodes[apex] - p0) > 0.0
        tri = tri.copy()
        tri[inward] = tri[inward][:, [0, 2, 1]]
        out.append(tri)
    if not out:
        return np.empty((0, 3), dtype=np.int32)
    return np.concatenate(out).astype(np.int32)


def bump_channel(nx: int, ny: int, nz: int, *, length: float = 3.0,
                 width: float = 1.0, bump_height: float = 0.1,
                 jitter: float = 0.0, seed: Optional[int] = None,
                 rcm: bool = False) -> ChannelMesh:
    """Kuhn-split bump channel with nx x ny x nz hexes (see module doc).

    Boundary tags: inflow (x=0), outflow (x=length), floor (y=0, carries the
    bump), ceiling (y=1), side_lo (z=0), side_hi (z=width)."""
    ijk, el = kuhn_box(nx, ny, nz)
    h = np.array([length / nx, 1.0 / ny, width / nz])
    xyz = ijk * h
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        interior = ((ijk[:, 0] > 0) & (ijk[:, 0] < nx) & (ijk[:, 1] > 0)
                    & (ijk[:, 1] < ny) & (ijk[:, 2] > 0) & (ijk[:, 2] < nz))
        d = rng.uniform(-jitter, jitter, size=(int(interior.sum()), 3)) * h
        xyz[interior] += d
    nodes = xyz.copy()
    nodes[:, 1] = xyz[:, 1] + bump_profile(xyz[:, 0], bump_height) * (1.0 - xyz[:, 1])
    planes = {
        "inflow": ijk[:, 0] == 0, "outflow": ijk[:, 0] == nx,
        "floor": ijk[:, 1] == 0, "ceiling": ijk[:, 1] == ny,
        "side_lo": ijk[:, 2] == 0, "side_hi": ijk[:, 2] == nz,
    }
    node_order = None
    if rcm:
        n = len(nodes)
        rows = np.repeat(el, 4, axis=1).reshape(-1)
        cols = np.tile(el, (1, 4)).reshape(-1)
        adj = sp.coo_matrix((np.ones(len(rows), dtype=np.int8), (rows, cols)),
                            shape=(n, n)).tocsr()
        order = reverse_cuthill_mckee(adj, symmetric_mode=True)  # new -> old
        inv = np.empty(n, dtype=np.int64)
        inv[order] = np.arange(n)
        nodes = nodes[order]
        el = inv[el]
        planes = {k: v[order] for k, v in planes.items()}
        srt = np.sort(el, axis=1)
        perm = np.lexsort((srt[:, 3], srt[:, 2], srt[:, 1], srt[:, 0]))
        el = el[perm]
        node_order = order
    el = el.astype(np.int32)
    bfaces = {tag: _boundary_faces(el, nodes, m) for tag, m in planes.items()}
    return ChannelMesh(nodes=np.ascontiguousarray(nodes), elements=el,
                       boundary_faces=bfaces, shape=(nx, ny, nz),
                       node_order=node_order,
                       meta={"length": length, "width": width,
                             "bump_height": bump_height, "jitter": jitter,
                             "seed": seed, "rcm": rcm})


def boundary_nodes(mesh: ChannelMesh, tags) -> np.ndarray:
    faces = [mesh.boundary_faces[t] for t in tags]
    return np.unique(np.concatenate(faces).ravel()) if faces else np.empty(0, np.int64)


def neumann_flux_rhs(mesh: ChannelMesh, tags=("inflow", "outflow"),
                     alpha_deg: float = 0.0, u_inf: float = 1.0,
                     rho_inf: float = 1.0) -> np.ndarray:
    """Consistent P1 load of the freestream flux rho_inf (u . n) on the
    given faces: each face adds rho_inf (u . n) area / 3 to its three nodes
    (the Galerkin boundary term of pyfp3d/solve/picard_ls.py:268-294)."""
    a = np.deg2rad(alpha_deg)
    u = np.array([u_inf * np.cos(a), u_inf * np.sin(a), 0.0])
    b = np.zeros(mesh.n_nodes)
    for t in tags:
        f = mesh.boundary_faces[t].astype(np.int64)
        if len(f) == 0:
            continue
        p = mesh.nodes[f]
        cr = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
        area = 0.5 * np.linalg.norm(cr, axis=1)
        nhat = cr / (2.0 * area)[:, None]
        contrib = rho_inf * (nhat @ u) * area / 3.0
        np.add.at(b, f.ravel(), np.repeat(contrib, 3))
    return b


def config_mesh(name: str, **kw) -> ChannelMesh:
    """Named workloads: 'c1' 48x16x16 (73,728 tets), 'c2' 192x64x64 jittered
    +-15% + RCM (4,718,592 tets), 'c3' 96x32x32 jittered + RCM."""
    if name == "c1":
        return bump_channel(48, 16, 16, **kw)
    if name == "c2":
        kw.setdefault("jitter", 0.15)
        kw.setdefault("seed", 20260101)
        kw.setdefault("rcm", True)
        return bump_channel(192, 64, 64, **kw)
    if name == "c3":
        kw.setdefault("jitter", 0.15)
        kw.setdefault("seed", 20260102)
        kw.setdefault("rcm", True)
        return bump_channel(96, 32, 32, **kw)
    raise ValueError(f"unknown workload {name!r}")

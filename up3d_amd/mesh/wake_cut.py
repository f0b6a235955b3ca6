"""Conforming wake cut: duplicate the nodes of a tagged interior wake sheet
so the potential can jump across it (pyfp3d/mesh/wake_cut.py:86-594,
WakeCut and cut_wake).

Conventions (the reference's):
  * the side of the sheet that `upper_hint` points to is "+": its elements
    read NEW node ids appended after the original ones (slaves); the "-"
    side keeps the original ids (masters); the jump phi(+) - phi(-) is the
    station circulation Gamma;
  * trailing-edge nodes (on the wall and the sheet) are duplicated like the
    rest of the sheet, and so are sheet nodes on other boundaries (their
    boundary faces owned by a "+" element are re-pointed to the slave);
  * sheet edges that end inside the domain (a wing-tip edge) are free
    edges: their nodes stay single-valued and leave the stations;
  * stations group the TE nodes by (x, y) (one station on a quasi-2D
    extrusion); every duplicated node takes the station nearest in z;
  * Kutta probes: per TE node, the nearest wall-edge neighbour on each side
    of the sheet (same z-plane first), not on the sheet.

Each duplicated node's element star is split into its two sides by a
union-find over the star's faces through that node that are not wake
faces; the "+" component is the one holding the "+" owner of the node's
wake faces.  Host numpy, once per mesh.
"""

from __future__ import annotations

from types import SimpleNamespace
from typing import Tuple

import numpy as np

_FACES = np.array([(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)], dtype=np.int64)


class WakeCut:
    """Duplication map, stations and Kutta probes of a cut (attribute
    names and meaning as pyfp3d/mesh/wake_cut.py:86-137)."""

    def __init__(self):
        self.n_nodes_orig = 0
        self.master_nodes = np.empty(0, dtype=np.int64)
        self.slave_nodes = np.empty(0, dtype=np.int64)
        self.free_nodes = np.empty(0, dtype=np.int64)
        self.te_nodes = np.empty(0, dtype=np.int64)
        self.station_z = np.empty(0, dtype=np.float64)
        self.node_station = np.empty(0, dtype=np.int64)
        self.te_station = np.empty(0, dtype=np.int64)
        self.kutta_upper = np.empty(0, dtype=np.int64)
        self.kutta_lower = np.empty(0, dtype=np.int64)
        self.wake_faces_minus = np.empty((0, 3), dtype=np.int32)
        self.wake_faces_plus = np.empty((0, 3), dtype=np.int32)

    @property
    def n_stations(self) -> int:
        return len(self.station_z)

    def gamma_at_nodes(self, gamma_stations) -> np.ndarray:
        return np.asarray(gamma_stations, dtype=np.float64)[self.node_station]


def _keys(tris: np.ndarray) -> np.ndarray:
    """One int64-triple row per face, sorted within the row, as a
    structured view usable with np.unique / searchsorted."""
    t = np.sort(np.asarray(tris, dtype=np.int64), axis=1)
    return np.ascontiguousarray(t).view([("a", np.int64), ("b", np.int64), ("c", np.int64)]).ravel()


def _face_owners(elements: np.ndarray):
    """(sorted face keys, owner element per key) of every tet face."""
    tris = elements[:, _FACES].reshape(-1, 3)
    owner = np.repeat(np.arange(len(elements), dtype=np.int64), 4)
    k = _keys(tris)
    order = np.argsort(k)
    return k[order], owner[order]


def _owners(sorted_keys, sorted_owner, tris):
    k = _keys(tris)
    lo = np.searchsorted(sorted_keys, k, side="left")
    hi = np.searchsorted(sorted_keys, k, side="right")
    return lo, hi, sorted_owner


def _free_edge_nodes(wake_faces: np.ndarray, boundary_faces: dict, wake_tag: str) -> np.ndarray:
    """Nodes of sheet edges used by one wake face only and by no other
    boundary face (edges where the sheet ends inside the domain)."""
    e = np.sort(wake_faces[:, [(0, 1), (1, 2), (0, 2)]].reshape(-1, 2), axis=1)
    uniq, cnt = np.unique(e, axis=0, return_counts=True)
    rim = uniq[cnt == 1]
    if not len(rim):
        return np.empty(0, dtype=np.int64)
    other = [np.asarray(t, dtype=np.int64) for tag, t in boundary_faces.items()
             if tag != wake_tag and len(t)]
    if other:
        ob = np.concatenate(other)
        oe = np.unique(np.sort(ob[:, [(0, 1), (1, 2), (0, 2)]].reshape(-1, 2), axis=1), axis=0)
        on_boundary = np.isin(rim[:, 0] * (1 << 32) + rim[:, 1], oe[:, 0] * (1 << 32) + oe[:, 1])
        rim = rim[~on_boundary]
    return np.unique(rim)


def _split_star(d, star, elements, wake_set):
    """(component label per star element, number of components) of node d's
    element star, connected through faces that contain d and are not wake
    faces."""
    parent = np.arange(len(star))

    def find(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i

    seen = {}
    for li, e in enumerate(star):
        for f in _FACES:
            tri = elements[e, f]
            if d not in tri:
                continue
            key = tuple(sorted(int(x) for x in tri))
            if key in wake_set:
                continue
            if key in seen:
                a, b = find(seen[key]), find(li)
                if a != b:
                    parent[a] = b
            else:
                seen[key] = li
    roots = np.array([find(i) for i in range(len(star))])
    _, comp = np.unique(roots, return_inverse=True)
    return comp, int(comp.max()) + 1 if len(comp) else 0


def cut_wake(mesh, wake_tag: str = "wake", wall_tag: str = "wall",
             upper_hint=(0.0, 1.0, 0.0)) -> Tuple[SimpleNamespace, WakeCut]:
    """(mesh_cut, wake_cut) for a mesh with an interior face group
    `wake_tag` (each face shared by two tets) and a wall group.  mesh_cut
    has nodes, elements (int32) and boundary_faces with the wake group
    replaced by "wake_minus" / "wake_plus".  Raises AssertionError on a
    sheet that does not cut a node star in two, as the reference does."""
    bf = mesh.boundary_faces
    if wake_tag not in bf:
        raise ValueError(f"mesh has no '{wake_tag}' face group")
    if wall_tag not in bf:
        raise ValueError(f"mesh has no '{wall_tag}' face group")
    nodes = np.asarray(mesh.nodes, dtype=np.float64)
    el = np.asarray(mesh.elements, dtype=np.int64)
    n = len(nodes)
    hint = np.asarray(upper_hint, dtype=np.float64)
    hint = hint / np.linalg.norm(hint)
    wake = np.asarray(bf[wake_tag], dtype=np.int64)
    sk, so = _face_owners(el)
    lo, hi, _ = _owners(sk, so, wake)
    if np.any(hi - lo != 2):
        raise AssertionError(f"{int(np.sum(hi - lo != 2))} wake face(s) not shared "
                             "by exactly two tets")
    e0, e1 = so[lo], so[lo + 1]
    cen = nodes[el].mean(axis=1)
    fc = nodes[wake].mean(axis=1)
    s0 = (cen[e0] - fc) @ hint
    s1 = (cen[e1] - fc) @ hint
    if np.any(s0 == s1):
        raise AssertionError("a wake face's owners cannot be told apart along upper_hint")
    plus_owner = np.where(s0 > s1, e0, e1)

    wake_nodes = np.unique(wake)
    wall_nodes = np.unique(np.asarray(bf[wall_tag], dtype=np.int64))
    is_wall = np.zeros(n, dtype=bool)
    is_wall[wall_nodes] = True
    free = _free_edge_nodes(wake, bf, wake_tag)
    is_free = np.zeros(n, dtype=bool)
    is_free[free] = True
    te = wake_nodes[is_wall[wake_nodes] & ~is_free[wake_nodes]]
    dup = wake_nodes[~is_free[wake_nodes]]
    if not len(te):
        raise AssertionError("wake sheet does not touch the wall (no TE nodes)")

    # stations: TE nodes grouped by (x, y)
    zext = float(np.ptp(nodes[:, 2]))
    z_tol = 1e-9 * zext if zext > 0 else 1e-12
    bbox = float(np.max(np.ptp(nodes, axis=0)))
    xy = np.round(nodes[te, :2] / (1e-6 * bbox)).astype(np.int64)
    _, te_station = np.unique(xy, axis=0, return_inverse=True)
    te_station = te_station.astype(np.int64).ravel()
    n_st = int(te_station.max()) + 1
    station_z = np.array([nodes[te[te_station == j], 2].mean() for j in range(n_st)])
    if n_st == 1:
        node_station = np.zeros(len(dup), dtype=np.int64)
    else:
        order = np.argsort(station_z)
        node_station = order[np.argmin(np.abs(nodes[dup, 2][:, None]
                                              - station_z[order][None, :]), axis=1)]

    # split each duplicated node's star
    flat = el.reshape(-1)
    srt = np.argsort(flat, kind="stable")
    ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(flat, minlength=n), out=ptr[1:])
    wake_set = {tuple(sorted(int(x) for x in t)) for t in wake}
    node_faces = {}
    for i, tri in enumerate(wake):
        for d in tri:
            node_faces.setdefault(int(d), []).append(i)
    el_cut = el.copy()
    slave = n + np.arange(len(dup), dtype=np.int64)
    slave_of = np.full(n, -1, dtype=np.int64)
    slave_of[dup] = slave
    plus_elems = {}
    for k, d in enumerate(dup):
        d = int(d)
        star = srt[ptr[d]:ptr[d + 1]] // 4
        comp, ncomp = _split_star(d, star, el, wake_set)
        if ncomp != 2:
            raise AssertionError(f"wake node {d}: element star splits into {ncomp} "
                                 "components (expected 2)")
        pc = {int(comp[np.flatnonzero(star == plus_owner[i])[0]]) for i in node_faces[d]}
        if len(pc) != 1:
            raise AssertionError(f"wake node {d}: inconsistent side classification")
        pe = star[comp == pc.pop()]
        plus_elems[d] = set(int(e) for e in pe)
        rows = el_cut[pe]
        rows[rows == d] = slave[k]
        el_cut[pe] = rows

    # boundary faces owned by a "+" element read the slave
    out_bf = {}
    for tag, tris in bf.items():
        if tag == wake_tag:
            continue
        t = np.asarray(tris, dtype=np.int64).copy()
        touched = np.flatnonzero(np.isin(t, dup).any(axis=1))
        if len(touched):
            lo2, hi2, _ = _owners(sk, so, t[touched])
            if np.any(hi2 - lo2 != 1):
                raise AssertionError(f"a '{tag}' face is not owned by exactly one tet")
            for r, e in zip(touched, so[lo2]):
                for j in range(3):
                    d = int(t[r, j])
                    if slave_of[d] >= 0 and int(e) in plus_elems[d]:
                        t[r, j] = slave_of[d]
        out_bf[tag] = t.astype(np.int32)
    wplus = wake.copy()
    m = slave_of[wplus] >= 0
    wplus[m] = slave_of[wplus[m]]
    out_bf["wake_minus"] = wake.astype(np.int32)
    out_bf["wake_plus"] = wplus.astype(np.int32)

    wc = WakeCut()
    wc.n_nodes_orig = n
    wc.master_nodes = dup
    wc.slave_nodes = slave
    wc.free_nodes = free
    wc.te_nodes = te
    wc.station_z = station_z
    wc.node_station = node_station
    wc.te_station = te_station
    wc.wake_faces_minus = out_bf["wake_minus"]
    wc.wake_faces_plus = out_bf["wake_plus"]
    wc.kutta_upper, wc.kutta_lower = _kutta_probes(nodes, np.asarray(bf[wall_tag]), te,
                                                   wake_nodes, hint, z_tol)
    mesh_cut = SimpleNamespace(nodes=np.vstack([nodes, nodes[dup]]),
                               elements=el_cut.astype(np.int32), boundary_faces=out_bf)
    return mesh_cut, wc


def _kutta_probes(nodes, wall_faces, te, wake_nodes, hint, z_tol):
    """Per TE node the nearest wall-edge neighbour off the sheet on the "+"
    and on the "-" side, in the TE node's z-plane when one exists
    (pyfp3d/mesh/wake_cut.py:444-506)."""
    wf = np.asarray(wall_faces, dtype=np.int64)
    is_sheet = np.zeros(len(nodes), dtype=bool)
    is_sheet[wake_nodes] = True
    up = np.full(len(te), -1, dtype=np.int64)
    lo = np.full(len(te), -1, dtype=np.int64)
    for k, t in enumerate(te):
        nb = np.unique(wf[(wf == t).any(axis=1)])
        nb = nb[(nb != t) & ~is_sheet[nb]]
        for same_plane in (True, False):
            cand = nb
            if same_plane:
                cand = nb[np.abs(nodes[nb, 2] - nodes[t, 2]) <= 10 * z_tol + 1e-12]
            off = nodes[cand] - nodes[t]
            side = off @ hint
            dist = np.linalg.norm(off, axis=1)
            pu, pl = cand[side > 0], cand[side < 0]
            if len(pu) and len(pl):
                up[k] = pu[np.argmin(dist[side > 0])]
                lo[k] = pl[np.argmin(dist[side < 0])]
                break
    if np.any(up < 0) or np.any(lo < 0):
        raise AssertionError("a TE node has no wall neighbour on both sides of the sheet")
    return up, lo

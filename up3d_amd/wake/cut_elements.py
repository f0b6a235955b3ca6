"""Cut-element census and extended-DOF numbering of a level-set wake
(pyfp3d/wake/cut_elements.py:73-333, CutElementMap).

Rules (same as the reference):
  * node side +1 / -1, never 0: nodes with |s| < eps_rel * h_node (h_node =
    shortest incident edge) and the TE nodes are shifted to "+";
  * TE nodes: |s| and |d| both below te_tol_rel * h_node (restricted to
    wall nodes when given);
  * cut elements: sign-changing elements with an edge crossing s = 0 that
    lies downstream of the TE (d > 0) and within the spanwise extent
    (0 <= q <= span_length [+ outboard_fringe]);
  * the below-TE fan (elements with a TE node whose other nodes are all "-")
    is never cut; it references the TE through its aux DOF;
  * `minus_nodes` (no reference counterpart): nodes forced to the "-"
    side, for the lower-face copies of a zero-thickness wall that lies on
    the sheet plane upstream of the TE (a cracked embedded plate);
  * one aux DOF per distinct cut-element node, numbered n_main + k in
    ascending node order; dofs_upper = main at "+" nodes, aux at "-" nodes,
    dofs_lower the complement.
Host numpy, once per mesh and level-set direction.
"""

from __future__ import annotations

from typing import Dict, Optional

import numpy as np

_EDGES = np.array([(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)],
                  dtype=np.int64)


def node_min_edge_length(nodes, elements) -> np.ndarray:
    """Shortest incident edge per node (+inf for a node in no tet), in the
    library's host code: one pass over the tets' six edges."""
    from .. import _native

    nd = np.ascontiguousarray(nodes, dtype=np.float64)
    el = np.ascontiguousarray(elements, dtype=np.int32)
    h = np.empty(len(nd))
    _native.check(_native.lib().fp3d_host_node_min_edge_length(
        len(el), el.ctypes.data, len(nd), nd.ctypes.data, h.ctypes.data),
        "node_min_edge_length")
    return h


class CutElementMap:
    def __init__(self, nodes, elements, levelset, wall_nodes=None,
                 eps_rel: float = 1e-6, te_tol_rel: float = 1e-3,
                 inboard_clip=None, outboard_fringe: float = 0.0,
                 minus_nodes=None):
        nodes = np.asarray(nodes, dtype=np.float64)
        el = np.asarray(elements, dtype=np.int64)
        n = len(nodes)
        if outboard_fringe < 0.0:
            raise ValueError("outboard_fringe must be >= 0")
        self.n_main = n
        self.wall_nodes = None if wall_nodes is None else \
            np.asarray(wall_nodes, dtype=np.int64)
        self.span_length = levelset.span_length
        self.outboard_fringe = float(outboard_fringe)
        s_raw, d, q = levelset.evaluate(nodes)
        self.s_raw, self.d, self.q = s_raw, d, q
        h = node_min_edge_length(nodes, el)

        te = (np.abs(s_raw) < te_tol_rel * h) & (np.abs(d) < te_tol_rel * h)
        if self.wall_nodes is not None:
            on_wall = np.zeros(n, dtype=bool)
            on_wall[self.wall_nodes] = True
            te &= on_wall
        self.te_nodes = np.flatnonzero(te)

        eps = eps_rel * h
        shift = (np.abs(s_raw) < eps) | te
        s = s_raw.copy()
        s[shift] = eps[shift]
        if minus_nodes is not None:
            # lower-face copies of a cracked surface lying on the sheet
            mn = np.asarray(minus_nodes, dtype=np.int64)
            if np.any(te[mn]):
                raise ValueError("a minus node is a TE node")
            s[mn] = -np.maximum(np.abs(s_raw[mn]), eps[mn])
            shift[mn] = True
        self.s = s
        self.shifted_nodes = np.flatnonzero(shift)
        self.node_side = np.where(s > 0.0, 1, -1).astype(np.int8)

        # candidates: elements whose nodes change side
        side_e = self.node_side[el]
        cand = np.flatnonzero(side_e.min(axis=1) != side_e.max(axis=1))
        ce = el[cand]
        sa, sb = s[ce[:, _EDGES[:, 0]]], s[ce[:, _EDGES[:, 1]]]
        cross = sa * sb < 0.0
        with np.errstate(divide="ignore", invalid="ignore"):
            t = np.where(cross, sa / (sa - sb), 0.0)
        da, db = d[ce[:, _EDGES[:, 0]]], d[ce[:, _EDGES[:, 1]]]
        qa, qb = q[ce[:, _EDGES[:, 0]]], q[ce[:, _EDGES[:, 1]]]
        d_x = da + t * (db - da)
        q_x = qa + t * (qb - qa)
        downstream = cross & (d_x > 0.0)
        if np.isinf(self.outboard_fringe):
            on = downstream
        else:
            on = downstream & (q_x <= self.span_length + self.outboard_fringe)
        if inboard_clip is None:
            on &= q_x >= 0.0
        else:
            xa = nodes[ce[:, _EDGES[:, 0]]]
            xb = nodes[ce[:, _EDGES[:, 1]]]
            pts = xa + t[:, :, None] * (xb - xa)
            keep = np.asarray(inboard_clip(pts), dtype=bool)
            if keep.shape != cross.shape:
                raise ValueError("inboard_clip must return an (n, 6) mask")
            on &= keep
        self.cut_elems = cand[on.any(axis=1)]
        self.beyond_tip_elems = cand[downstream.any(axis=1) & ~on.any(axis=1)]

        # below-TE fan: a TE node and every other node "-" (not cut)
        is_te = np.zeros(n, dtype=bool)
        is_te[self.te_nodes] = True
        te_e = is_te[el]
        fan = te_e.any(axis=1) & np.all(te_e | (side_e == -1), axis=1)
        self.te_lower_elems = np.flatnonzero(fan)
        self.cut_elems = self.cut_elems[~fan[self.cut_elems]]

        cut_nodes = np.unique(el[self.cut_elems])
        self.n_ext_dofs = len(cut_nodes)
        self.n_total_dofs = n + self.n_ext_dofs
        self.ext_dof_of_node = np.full(n, -1, dtype=np.int64)
        self.ext_dof_of_node[cut_nodes] = n + np.arange(len(cut_nodes))
        if len(self.te_nodes) and self.n_ext_dofs:
            if not np.all(self.ext_dof_of_node[self.te_nodes] >= 0):
                raise ValueError("a TE node has no aux DOF: the wake sheet "
                                 "does not reach the trailing edge")
        conn = el[self.cut_elems]
        aux = self.ext_dof_of_node[conn]
        plus = self.node_side[conn] == 1
        self.dofs_upper = np.where(plus, conn, aux)
        self.dofs_lower = np.where(plus, aux, conn)

    def summary(self) -> Dict[str, int]:
        return {
            "n_main": int(self.n_main), "n_ext_dofs": int(self.n_ext_dofs),
            "n_cut_elems": int(len(self.cut_elems)),
            "n_te_nodes": int(len(self.te_nodes)),
            "n_te_lower_elems": int(len(self.te_lower_elems)),
            "n_beyond_tip_elems": int(len(self.beyond_tip_elems)),
            "n_shifted_nodes": int(len(self.shifted_nodes)),
            "n_side_plus": int(np.sum(self.node_side == 1)),
            "n_side_minus": int(np.sum(self.node_side == -1)),
        }

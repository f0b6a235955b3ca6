"""Lifting case on an embedded zero-thickness swept plate (mesh from
`mesh.plate_box.plate_wing_box`): a planar level-set wake z = 0 downstream
of the trailing edge, cut-element aux DOFs carrying the potential jump
Gamma(y), the pressure-equality Kutta condition at every TE node, and a
freestream far field in the x-z plane (alpha rotates x toward +z).

  * `PlateWingCase.initial_state(alpha_deg, gamma0)`: freestream main DOFs
    and an elliptic jump Gamma0 sqrt(1 - (y/b)^2) on the aux DOFs;
  * `lift_coefficient(phi_ext)`: Kutta-Joukowski lift of the full wing from
    the TE jumps, CL = 2 * 2 int_0^b Gamma dy / (U S), S = 2 b c
    (trapezoid rule over the TE nodes);
  * `lift_gradient(phi_ext)`: dCL/dphi_ext (the adjoint's load).
"""

from __future__ import annotations

import numpy as np

from .cut_elements import CutElementMap
from .levelset import WakeLevelSet
from .multivalued import MultivaluedOperator


class PlateWingCase:
    def __init__(self, mesh, u_inf: float = 1.0):
        self.mesh = mesh
        self.u_inf = float(u_inf)
        meta = mesh.meta
        self.b = float(meta["half_span"])
        self.c = float(meta["chord"])
        self.area = float(meta["area_full"])
        self.levelset = WakeLevelSet(mesh.te_line, direction=(1.0, 0.0, 0.0))
        b = self.b
        # the sheet's spanwise end: only crossings strictly inboard of the
        # tip count, so every aux DOF lies on a streamwise line that starts
        # at a TE node (a line beyond the tip would carry an undetermined
        # jump: the wake rows constrain its x and z derivatives only); the
        # single-valued tets just outboard drop the tip jump over one cell
        self.cm = CutElementMap(mesh.nodes, mesh.elements, self.levelset,
                                wall_nodes=mesh.wall_nodes,
                                minus_nodes=mesh.lower_nodes,
                                inboard_clip=lambda p: p[..., 1] < b * (1.0 - 1e-9),
                                outboard_fringe=1e-9 * self.levelset.span_length)
        if len(self.cm.te_nodes) != len(mesh.te_nodes):
            raise ValueError("the level set did not find every trailing-edge node")
        self.mvop = MultivaluedOperator(mesh.nodes, mesh.elements, self.cm,
                                        levelset=self.levelset)
        te = self.cm.te_nodes
        order = np.argsort(mesh.nodes[te, 1])
        self.te_sorted = te[order]
        y = mesh.nodes[self.te_sorted, 1]
        w = np.zeros(len(y))
        dy = np.diff(y)
        w[:-1] += 0.5 * dy
        w[1:] += 0.5 * dy
        self.te_weights = w                       # int_0^b f dy ~ sum w f

    def initial_state(self, alpha_deg: float, gamma0=None) -> np.ndarray:
        """Freestream main DOFs, aux DOFs = main - side Gamma(y) with the
        elliptic Gamma0 sqrt(1 - (y/b)^2) (Gamma0 by default from lifting-line
        theory, CL = 2 pi alpha AR / (AR + 2), AR = 2 b / c)."""
        from ..solve.picard_ls import _freestream

        mv, cm = self.mvop, self.cm
        nodes = self.mesh.nodes
        if gamma0 is None:
            ar = 2.0 * self.b / self.c
            cl = 2.0 * np.pi * np.deg2rad(alpha_deg) * ar / (ar + 2.0)
            gamma0 = cl * self.u_inf * self.area / (np.pi * self.b)
        phi = np.zeros(mv.n_total)
        phi[: mv.n_main] = _freestream(nodes, alpha_deg, self.u_inf, lift_axis=2)
        yy = np.clip(nodes[mv.cut_nodes, 1] / self.b, 0.0, 1.0)
        g = gamma0 * np.sqrt(1.0 - yy * yy)
        phi[mv.cut_aux] = phi[mv.cut_nodes] - cm.node_side[mv.cut_nodes] * g
        return phi

    def solve_kwargs(self, alpha_deg: float) -> dict:
        """Keyword arguments of the level-set solvers for this case: the
        freestream Dirichlet far field in the x-z plane and the Neumann
        outflow load."""
        return {"alpha_deg": float(alpha_deg), "farfield": "freestream",
                "lift_axis": 2, "external_rhs": self.outflow_load(alpha_deg),
                "u_inf": self.u_inf}

    def outflow_load(self, alpha_deg: float, derivative: bool = False) -> np.ndarray:
        """Consistent P1 load of the freestream flux rho_inf (u . n) A / 3 on
        the outflow faces, u = U (cos a, 0, sin a) (its alpha-derivative per
        degree with derivative=True).  A face below the sheet gives nothing
        to a "+" node that carries an aux DOF: that node's lower-side row is
        a wake row, not a mass balance."""
        mv, cm = self.mvop, self.cm
        f = self.mesh.boundary_faces["outflow"].astype(np.int64)
        p = self.mesh.nodes[f]
        cr = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
        area = 0.5 * np.linalg.norm(cr, axis=1)
        nhat = cr / (2.0 * area)[:, None]
        a = np.deg2rad(alpha_deg)
        if derivative:
            u = self.u_inf * np.deg2rad(1.0) * np.array([-np.sin(a), 0.0, np.cos(a)])
        else:
            u = self.u_inf * np.array([np.cos(a), 0.0, np.sin(a)])
        share = np.repeat((nhat @ u) * area / 3.0, 3).reshape(-1, 3)
        below = p[:, :, 2].mean(axis=1) < 0.0
        skip = below[:, None] & (cm.node_side[f] == 1) & (cm.ext_dof_of_node[f] >= 0)
        share[skip] = 0.0
        out = np.zeros(mv.n_total)
        np.add.at(out, f.ravel(), share.ravel())
        return out

    def gamma_distribution(self, phi_ext):
        """(y, Gamma(y)) at the TE nodes, root to tip."""
        return (self.mesh.nodes[self.te_sorted, 1],
                self.mvop.node_jump(phi_ext, self.te_sorted))

    def lift_coefficient(self, phi_ext) -> float:
        _, g = self.gamma_distribution(phi_ext)
        return float(4.0 * (self.te_weights * g).sum() / (self.u_inf * self.area))

    def lift_gradient(self, phi_ext=None) -> np.ndarray:
        """dCL/dphi_ext (CL is linear in phi_ext)."""
        mv, cm = self.mvop, self.cm
        t = self.te_sorted
        aux = cm.ext_dof_of_node[t]
        side = cm.node_side[t].astype(np.float64)
        c = 4.0 * self.te_weights * side / (self.u_inf * self.area)
        g = np.zeros(mv.n_total)
        np.add.at(g, t, c)
        np.add.at(g, aux, -c)
        return g

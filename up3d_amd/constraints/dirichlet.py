"""Far-field potentials (pyfp3d/constraints/dirichlet.py:29-100)."""

import numpy as np


def freestream_phi(points: np.ndarray, alpha_deg: float = 0.0,
                   u_inf: float = 1.0) -> np.ndarray:
    """phi_inf = U (x cos(alpha) + y sin(alpha))."""
    a = np.deg2rad(alpha_deg)
    return u_inf * (points[:, 0] * np.cos(a) + points[:, 1] * np.sin(a))


def vortex_phi_2d(points: np.ndarray, gamma_total, center=(0.25, 0.0),
                  lower_branch_mask=None, beta: float = 1.0) -> np.ndarray:
    """Prandtl-Glauert point-vortex potential with the branch cut along +x
    from `center` (pyfp3d/constraints/dirichlet.py:37-100):
    phi_v = -(Gamma / 2 pi) theta_w, theta_w = atan2(beta dy, dx) taken in
    [0, 2 pi); points flagged in lower_branch_mask that lie on the cut
    (dx > 0) take theta_w = 2 pi."""
    dx = points[:, 0] - center[0]
    dy = points[:, 1] - center[1]
    th = np.arctan2(beta * dy, dx)
    th = np.where(th < 0.0, th + 2.0 * np.pi, th)
    if lower_branch_mask is not None:
        th = np.where((dx > 0.0) & lower_branch_mask, 2.0 * np.pi, th)
    return -(np.asarray(gamma_total) / (2.0 * np.pi)) * th


def farfield_dirichlet(mesh_cut, wc, alpha_deg: float, gamma_stations,
                       u_inf: float = 1.0, vortex_center=(0.25, 0.0),
                       farfield_tag: str = "farfield", beta: float = 1.0,
                       spanwise_gamma: bool = False):
    """(nodes, values) of the far-field Dirichlet data on a cut mesh
    (pyfp3d/constraints/dirichlet.py:98-180): freestream plus the
    Prandtl-Glauert vortex of the mean station circulation, the wake's
    master nodes on the cut taking the lower branch (the slaves, at the
    same positions, the upper one); spanwise_gamma interpolates Gamma(z)
    over the stations, tapered to 0 at and beyond the sheet tip."""
    ff = np.unique(np.asarray(mesh_cut.boundary_faces[farfield_tag], dtype=np.int64))
    pts = mesh_cut.nodes[ff]
    g = np.atleast_1d(np.asarray(gamma_stations, dtype=np.float64))
    master = np.isin(ff, wc.master_nodes)
    values = freestream_phi(pts, alpha_deg, u_inf)
    if spanwise_gamma and wc.n_stations > 1:
        if np.any(g != 0.0):
            order = np.argsort(wc.station_z)
            zp, gp = wc.station_z[order], g[order]
            z_tip = float(mesh_cut.nodes[wc.master_nodes][:, 2].max())
            if z_tip > zp[-1]:
                zp, gp = np.append(zp, z_tip), np.append(gp, 0.0)
            gz = np.interp(pts[:, 2], zp, gp, right=0.0)
            values = values + vortex_phi_2d(pts, gz, vortex_center,
                                            lower_branch_mask=master, beta=beta)
    else:
        total = float(np.mean(g))
        if total != 0.0:
            values = values + vortex_phi_2d(pts, total, vortex_center,
                                            lower_branch_mask=master, beta=beta)
    return ff, values

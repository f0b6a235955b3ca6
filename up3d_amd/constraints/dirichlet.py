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

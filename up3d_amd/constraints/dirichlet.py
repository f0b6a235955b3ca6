"""Far-field potential (pyfp3d/constraints/dirichlet.py:29-34)."""

import numpy as np


def freestream_phi(points: np.ndarray, alpha_deg: float = 0.0,
                   u_inf: float = 1.0) -> np.ndarray:
    """phi_inf = U (x cos(alpha) + y sin(alpha))."""
    a = np.deg2rad(alpha_deg)
    return u_inf * (points[:, 0] * np.cos(a) + points[:, 1] * np.sin(a))

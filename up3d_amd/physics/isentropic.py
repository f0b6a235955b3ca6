"""Isentropic constitutive laws (pyfp3d/physics/isentropic.py).

Scalar laws are host Python (used for set-up constants such as the
limiter cap); the *_field functions run on the device through
libfp3d_hip.so.  They accept a numpy array (copied to HBM and back) or a
float64 device tensor (computed in place into a new device tensor).
"""

from __future__ import annotations

import numpy as np
import torch

from .. import _native
from ..device import empty, ptr, stream_handle, to_dev, to_host

GAMMA = 1.4
R_GAS = 287.05
M_CRIT_DEFAULT = 0.95
UPWIND_CONST_DEFAULT = 1.5


def density_isentropic(q_squared, M_inf, gamma=GAMMA):
    """rho = [1 + (g-1)/2 M^2 (1 - q^2)]^(1/(g-1)), base floored at 1e-10
    (pyfp3d/physics/isentropic.py:29-58)."""
    exponent = 1.0 / (gamma - 1.0)
    base = 1.0 + 0.5 * (gamma - 1.0) * (M_inf * M_inf) * (1.0 - q_squared)
    if base < 1e-10:
        base = 1e-10
    return base ** exponent


def speed_of_sound_squared(q_squared, M_inf, gamma=GAMMA):
    return 1.0 / (M_inf ** 2) + 0.5 * (gamma - 1.0) * (1.0 - q_squared)


def mach_number_squared(q_squared, M_inf, gamma=GAMMA):
    """(pyfp3d/physics/isentropic.py:79-97)."""
    return q_squared * (M_inf * M_inf) / (
        1.0 + 0.5 * (gamma - 1.0) * (M_inf * M_inf) * (1.0 - q_squared))


def critical_speed_squared(M_inf, gamma=GAMMA):
    """q*^2 where M = 1 (pyfp3d/physics/isentropic.py:100-118)."""
    return (2.0 + (gamma - 1.0) * M_inf ** 2) / ((gamma + 1.0) * M_inf ** 2)


def pressure_coefficient(q_squared, M_inf, gamma=GAMMA):
    """Exact isentropic Cp = 2/(g M^2) (rho^g - 1) (:121-143)."""
    rho = density_isentropic(q_squared, M_inf, gamma)
    return 2.0 / (gamma * M_inf ** 2) * (rho ** gamma - 1.0)


def pressure_coefficient_incompressible(q_squared):
    return 1.0 - q_squared


def density_derivative_wrt_q_sq(q_squared, M_inf, gamma=GAMMA):
    """-(M^2/2) rho^(2-g) (:163-180)."""
    rho = density_isentropic(q_squared, M_inf, gamma)
    return -0.5 * (M_inf * M_inf) * (rho ** (2.0 - gamma))


def mach_squared_derivative_wrt_q_sq(q_squared, M_inf, gamma=GAMMA):
    """M^2 (1 + (g-1)/2 M^2) / D^2 (:183-206)."""
    D = 1.0 + 0.5 * (gamma - 1.0) * (M_inf * M_inf) * (1.0 - q_squared)
    return (M_inf * M_inf) * (1.0 + 0.5 * (gamma - 1.0) * (M_inf * M_inf)) / (D * D)


def upwind_factor(q_squared, M_inf, M_crit=M_CRIT_DEFAULT,
                  C=UPWIND_CONST_DEFAULT, gamma=GAMMA):
    """nu = C max(0, 1 - Mc^2 / max(M^2, Mc^2)) (:209-232)."""
    m2 = mach_number_squared(q_squared, M_inf, gamma)
    mc2 = M_crit * M_crit
    return C * max(0.0, 1.0 - mc2 / max(m2, mc2))


def q2_at_mach(mach, M_inf, gamma=GAMMA):
    """q^2 at local Mach `mach`; +inf at M_inf = 0 (:261-289)."""
    if M_inf == 0.0:
        return np.inf
    m2 = mach * mach
    return (m2 * (1.0 + 0.5 * (gamma - 1.0) * M_inf ** 2)
            / (M_inf ** 2 * (1.0 + 0.5 * (gamma - 1.0) * m2)))


def _field(name, q_squared, *args):
    host = not isinstance(q_squared, torch.Tensor)
    q2 = to_dev(q_squared, torch.float64)
    out = empty(q2.shape)
    _native.call(name, q2.numel(), ptr(q2), *args, ptr(out), stream_handle())
    return to_host(out) if host else out


def density_field(q_squared, M_inf, gamma=GAMMA):
    """Elementwise rho(q^2) on the device (:235-258); rho == 1 bitwise at
    M_inf = 0."""
    return _field("fp3d_density_field", q_squared, float(M_inf), float(gamma))


def mach_squared_field(q_squared, M_inf, gamma=GAMMA):
    """Elementwise M^2(q^2) on the device (:317-334)."""
    return _field("fp3d_mach_squared_field", q_squared, float(M_inf),
                  float(gamma))


def limit_q2_field(q_squared, M_inf, m_cap, gamma=GAMMA):
    """q2_eff = min(q^2, q2_at_mach(m_cap)) on the device (:292-314)."""
    return _field("fp3d_limit_q2_field", q_squared,
                  float(q2_at_mach(m_cap, M_inf, gamma)))


def density_derivative_wrt_q_sq_field(q_squared, M_inf, gamma=GAMMA):
    """Elementwise d rho / d q^2 on the device (:337-351)."""
    return _field("fp3d_density_derivative_field", q_squared, float(M_inf),
                  float(gamma))

"""Mach continuation schedule (pyfp3d/solve/continuation.py:72-78)."""

import numpy as np


def mach_schedule(m_target: float, m_start: float = 0.70, dm: float = 0.05):
    """Ascending Mach ramp ending exactly at m_target."""
    if m_target <= m_start:
        return [float(m_target)]
    ms = list(np.arange(m_start, m_target - 1e-12, dm))
    ms.append(m_target)
    return [float(m) for m in ms]

"""up3d_amd -- MI355X-native full-potential nonlinear solve on unstructured
P1 tetrahedra, a drop-in for the numba hot path of pyfp3d (swayli94/UP3D).

Python host code mirrors pyfp3d's kernels / physics / solve call
signatures; every computation runs in hand-written HIP kernels for gfx950
(libfp3d_hip.so, C-ABI in include/fp3d.h) on HBM-resident data.
"""

__version__ = "0.1.0"

"""Embedded swept-plate lifting case on the device: level-set Newton with a
freestream far field and Neumann outflow, Kutta at every TE node, and the
adjoint dCL/dalpha (no reference counterpart: the reference has neither
the embedded plate nor an adjoint).  Pinned by: convergence to
||R||_inf < 1e-10 with 0 limited / floored; a spanwise loading that falls
from root to tip; adjoint = tangent (same linearisation, transposed) and
= central differences of converged solves (step 1e-3 deg); the autograd
Function returns the adjoint value."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def plate(gpu):
    from up3d_amd.mesh.plate_box import plate_wing_box
    from up3d_amd.wake.plate_case import PlateWingCase

    mesh = plate_wing_box(32, 24, 24)
    print(f"[plate] n_tets={len(mesh.elements)}", flush=True)
    return PlateWingCase(mesh)


def _solve(case, m, a, phi=None, n_seed=20):
    from up3d_amd.solve.newton_ls import solve_multivalued_newton

    return solve_multivalued_newton(
        case.mvop, case.mesh, m, phi_init=case.initial_state(a) if phi is None else phi,
        n_seed=n_seed if phi is None else 0, freeze_tol=1e-6, **case.solve_kwargs(a))


@pytest.mark.parametrize("m_inf", [0.5, 0.78])
def test_plate_newton_converges(plate, m_inf):
    r = _solve(plate, m_inf, 2.0)
    y, g = plate.gamma_distribution(r["phi_ext"])
    cl = plate.lift_coefficient(r["phi_ext"])
    print(f"[plate] M{m_inf}: CL={cl:.6f} n={r['n_newton']} Mmax={np.sqrt(r['mach2_max']):.3f} "
          f"Gamma root/tip {g[0]:.5f}/{g[-1]:.5f}", flush=True)
    assert r["converged"] and r["residual_history"][-1] < 1e-10
    assert r["n_limited"] == 0 and r["n_floored"] == 0
    assert np.all(np.diff(g) < 0.0) and g[-1] > 0.0
    # lifting-line order of magnitude: 2 pi alpha AR / (AR + 2), AR 2
    cl_ll = 2.0 * np.pi * np.deg2rad(2.0) * 0.5
    assert 0.5 * cl_ll < cl < 1.5 * cl_ll


def test_plate_adjoint_matches_tangent_and_differences(plate):
    from up3d_amd.solve.adjoint_ls import lift_alpha_sensitivity

    r = _solve(plate, 0.5, 2.0)
    s = lift_alpha_sensitivity(plate, r["phi_ext"], 0.5, 2.0)
    d = 1e-3
    cl = [plate.lift_coefficient(_solve(plate, 0.5, 2.0 + da, phi=r["phi_ext"])["phi_ext"])
          for da in (-d, d)]
    fd = (cl[1] - cl[0]) / (2 * d)
    print(f"[plate] dCL/dalpha adjoint {s['dcl_dalpha']:.10f} tangent "
          f"{s['dcl_dalpha_tangent']:.10f} FD {fd:.10f} (per deg)", flush=True)
    assert s["adjoint_info"] == 0 and s["tangent_info"] == 0
    assert abs(s["dcl_dalpha"] - s["dcl_dalpha_tangent"]) <= 1e-9 * abs(fd)
    assert abs(s["dcl_dalpha"] - fd) <= 1e-6 * abs(fd)


def test_plate_lift_autograd(plate):
    import torch

    from up3d_amd.solve.adjoint_ls import PlateLift, lift_alpha_sensitivity

    a = torch.tensor(2.0, dtype=torch.float64, requires_grad=True)
    cl = PlateLift.apply(a, plate, 0.5, {"n_seed": 20, "freeze_tol": 1e-6})
    cl.backward()
    r = _solve(plate, 0.5, 2.0)
    s = lift_alpha_sensitivity(plate, r["phi_ext"], 0.5, 2.0, tangent=False)
    assert abs(float(cl.detach()) - plate.lift_coefficient(r["phi_ext"])) < 1e-9
    assert abs(float(a.grad) - s["dcl_dalpha"]) <= 1e-8 * abs(s["dcl_dalpha"])

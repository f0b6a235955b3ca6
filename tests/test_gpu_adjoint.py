"""Discrete adjoint of the channel solve: gradients of the wall x-force and
lift w.r.t. M_inf, alpha and the floor-node y coordinates, checked against
the forward Jacobian-vector product and central finite differences of the
converged solve (step 1e-6), to the 1e-6 bar."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prob(gpu):
    from up3d_amd.mesh.channel import boundary_nodes, bump_channel
    from up3d_amd.solve.adjoint import ChannelForceProblem

    m = bump_channel(24, 8, 2, jitter=0.1, seed=3, rcm=True)
    dn = boundary_nodes(m, ["inflow", "outflow"])
    floor = np.unique(m.boundary_faces["floor"])
    p = ChannelForceProblem(m.nodes, m.elements, dn, m.boundary_faces["floor"], floor)
    print(f"[adjoint] mesh 24x8x2 n_tets={m.n_tets} floor nodes={len(floor)}", flush=True)
    return p


def _solve_ramp(prob, m_target, alpha):
    sol = None
    for m in (0.5, 0.6, 0.65, 0.7):
        if m > m_target + 1e-12:
            break
        sol = prob.solve(None, m, alpha, phi_init=None if sol is None else sol["phi"])
        assert sol["converged"], sol["residual_history"]
    if sol["m_inf"] != m_target:
        sol = prob.solve(None, m_target, alpha, phi_init=sol["phi"])
    return sol


@pytest.mark.parametrize("m_inf", [0.5, 0.7])
def test_adjoint_vs_jvp_and_fd(prob, m_inf):
    alpha = 1.0
    sol = _solve_ramp(prob, m_inf, alpha)
    live = sol["residual_unfrozen"] if sol["froze"] else sol["residual_history"][-1]
    assert live < 1e-10
    print(f"[adjoint] M={m_inf} nu_max={sol['nu_max']:.3f} "
          f"t3={sol['n_term3_active']}", flush=True)
    for w in ((1.0, 0.0), (0.0, 1.0)):
        g = prob.gradients(sol, weights=w)
        assert g["adjoint_info"] == 0
        # forward tangent
        jm = prob.jvp_params(sol, 1.0, 0.0, weights=w)
        ja = prob.jvp_params(sol, 0.0, 1.0, weights=w)
        assert g["d_m_inf"] == pytest.approx(jm, rel=1e-8, abs=1e-12)
        assert g["d_alpha_deg"] == pytest.approx(ja, rel=1e-8, abs=1e-12)
        # central finite differences of the converged solve
        h = 1e-6

        def F(y=None, m=m_inf, a=alpha):
            s = prob.solve(y, m, a, phi_init=sol["phi"])
            assert s["converged"]
            return prob.functional(s, weights=w)[0]

        fd_m = (F(m=m_inf + h) - F(m=m_inf - h)) / (2 * h)
        fd_a = (F(a=alpha + h) - F(a=alpha - h)) / (2 * h)
        assert g["d_m_inf"] == pytest.approx(fd_m, rel=1e-6, abs=1e-9)
        assert g["d_alpha_deg"] == pytest.approx(fd_a, rel=1e-6, abs=1e-9)
        # seeded unit-norm perturbations of the floor-node y coordinates
        y0 = prob.base_nodes[prob.floor_nodes, 1]
        rng = np.random.default_rng(16)
        for k in range(3):
            d = rng.standard_normal(len(y0))
            d /= np.linalg.norm(d)
            fd = (F(y=y0 + h * d) - F(y=y0 - h * d)) / (2 * h)
            ad = float(g["d_y_floor"] @ d)
            assert ad == pytest.approx(fd, rel=1e-6, abs=1e-9), (k, ad, fd)


def test_autograd_function(prob):
    from up3d_amd.solve.adjoint import ChannelForce

    y0 = torch.tensor(prob.base_nodes[prob.floor_nodes, 1], dtype=torch.float64,
                      requires_grad=True)
    m = torch.tensor(0.5, dtype=torch.float64, requires_grad=True)
    a = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    out = ChannelForce.apply(y0, m, a, prob)
    (out[0] + 0.5 * out[1]).backward()
    sol = prob.solve(None, 0.5, 1.0)
    g = prob.gradients(sol, weights=(1.0, 0.5))
    assert np.allclose(y0.grad.numpy(), g["d_y_floor"], rtol=1e-9, atol=1e-14)
    assert float(m.grad) == pytest.approx(g["d_m_inf"], rel=1e-9)
    assert float(a.grad) == pytest.approx(g["d_alpha_deg"], rel=1e-9)


def test_adjoint_vs_frozen_selection_fd(prob):
    """Central differences of the frozen-selection solve (the smooth piece
    the adjoint differentiates) agree with the adjoint far below the free
    re-solve's 1e-6 bar."""
    alpha, m_inf = 1.0, 0.7
    sol = _solve_ramp(prob, m_inf, alpha)
    frozen = prob.frozen_selection(sol)
    g = prob.gradients(sol, weights=(1.0, 0.0))
    h = 1e-6

    def F(y=None, m=m_inf, a=alpha):
        s = prob.solve_frozen(y, m, a, frozen, sol["phi"])
        assert s["converged"], s["residual_history"]
        return prob.functional(s)[0]

    base = prob.solve_frozen(None, m_inf, alpha, frozen, sol["phi"])
    assert base["converged"] and len(base["residual_history"]) <= 3
    fd_m = (F(m=m_inf + h) - F(m=m_inf - h)) / (2 * h)
    assert g["d_m_inf"] == pytest.approx(fd_m, rel=1e-7, abs=1e-10)
    y0 = prob.base_nodes[prob.floor_nodes, 1]
    rng = np.random.default_rng(16)
    for k in range(2):
        d = rng.standard_normal(len(y0))
        d /= np.linalg.norm(d)
        fd = (F(y=y0 + h * d) - F(y=y0 - h * d)) / (2 * h)
        assert float(g["d_y_floor"] @ d) == pytest.approx(fd, rel=1e-7, abs=1e-10), k


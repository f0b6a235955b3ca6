"""Solution-level parity of the device drivers with the oracle drivers
(direct sparse solves): same discrete problem, both converged to
||R||_inf < 1e-10, potential / wall Cp / wall force within 1e-8."""

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def chan(gpu):
    from up3d_amd.mesh.channel import bump_channel

    m = bump_channel(24, 8, 3, jitter=0.15, seed=5, rcm=True)
    print(f"[mesh] channel 24x8x3 n_tets={m.n_tets}", flush=True)
    return m


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_picard_subsonic_neumann_matches_oracle(chan):
    """Config-1 form: freestream Neumann flux in/out, one pinned node."""
    from up3d_amd.mesh.channel import neumann_flux_rhs
    from up3d_amd.post.surface import wall_force_coefficients
    from up3d_amd.solve.picard import solve_subsonic

    b = neumann_flux_rhs(chan)
    pin = np.array([int(np.argmin(np.linalg.norm(chan.nodes - [0, 0.5, 0.5], axis=1)))])
    vals = chan.nodes[pin, 0]
    res = solve_subsonic(chan.nodes, chan.elements, pin, vals, 0.5,
                         body_source_rhs=b)
    ref = O.solve_subsonic(chan.nodes, chan.elements, pin, vals, 0.5,
                           body_source_rhs=b)
    assert res["converged"] and ref["converged"]
    assert res["residual_history"][-1] < 1e-10
    assert _rel(res["phi"], ref["phi"]) < 1e-8
    wf = chan.boundary_faces["floor"]
    f1 = wall_force_coefficients(chan.nodes, chan.elements, wf, res["phi"], m_inf=0.5)
    f0 = O.wall_force_coefficients(chan.nodes, chan.elements, wf, ref["phi"], m_inf=0.5)
    assert np.abs(f1["cp_tri"] - f0["cp_tri"]).max() < 1e-8
    assert np.abs(f1["cf"] - f0["cf"]).max() < 1e-8


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_newton_transonic_matches_oracle(chan, precision):
    from up3d_amd.mesh.channel import boundary_nodes
    from up3d_amd.post.surface import wall_force_coefficients
    from up3d_amd.solve.newton import solve_newton

    dn = boundary_nodes(chan, ["inflow", "outflow"])
    dv = chan.nodes[dn, 0].copy()
    ows = O.OracleNewton(chan.nodes, chan.elements, dn, dv)
    phi = chan.nodes[:, 0].copy()
    ramp = (0.6, 0.65, 0.7, 0.75)
    for m in ramp:
        r0 = ows.solve(m, phi, n_newton_max=40, freeze_tol=1e-6)
        assert r0["converged"], (m, r0["residual_history"])
        phi = r0["phi"]
    # the oracle's final state must satisfy the LIVE residual (unique
    # discrete solution, independent of the freeze history)
    R_live, _ = ows.eval_residual(phi[ows.free], 0.75)
    assert np.abs(R_live).max() < 1e-10
    ws = None
    phi_d = chan.nodes[:, 0].copy()
    for m in ramp:
        r = solve_newton(chan.nodes, chan.elements, dn, dv, m, phi_init=phi_d,
                         n_picard_seed=0, freeze_tol=1e-6, n_newton_max=40,
                         workspace=ws, precision=precision, amg_rebuild_every=1)
        ws = r["workspace"]
        print(f"[newton {precision}] M={m} its={r['n_newton']} "
              f"|R|={r['residual_history'][-1]:.2e}", flush=True)
        assert r["converged"], r["residual_history"]
        phi_d = r["phi"]
    live = r["residual_unfrozen"] if r["froze"] else r["residual_history"][-1]
    assert live < 1e-10
    assert r["nu_max"] > 0.0 and r["n_term3_active"] > 0
    assert _rel(r["phi"], phi) < 1e-8
    wf = chan.boundary_faces["floor"]
    f1 = wall_force_coefficients(chan.nodes, chan.elements, wf, r["phi"], m_inf=0.75)
    f0 = O.wall_force_coefficients(chan.nodes, chan.elements, wf, phi, m_inf=0.75)
    assert np.abs(f1["cp_tri"] - f0["cp_tri"]).max() < 1e-8
    assert np.abs(f1["cf"] - f0["cf"]).max() < 1e-8

"""The one-call Arnoldi step (fp3d_arnoldi_step, which gmres_steps uses for a
Newton-Jacobian matvec with an AMG V-cycle) against the same step issued
launch by launch: GMRES on a transonic channel's Newton system (Term 3
active) gives the same solution, iteration count and residual bit for bit,
with the float64 and the float32 Krylov basis, with and without the
preconditioner, and the Newton solve through it equals the one without."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def newton_system(gpu):
    from up3d_amd.device import to_dev
    from up3d_amd.mesh.channel import boundary_nodes, bump_channel
    from up3d_amd.solve.newton import NewtonWorkspace

    m = bump_channel(24, 8, 6, jitter=0.15, seed=3, rcm=True)
    dn = boundary_nodes(m, ["inflow", "outflow"])
    ws = NewtonWorkspace(m.nodes, m.elements, dn, m.nodes[dn, 0].copy())
    ws.set_mach(0.80)
    s = ws.states[0]
    x = m.nodes[:, 0]
    s.phi.copy_(to_dev(1.25 * x + 0.35 * np.sin(1.7 * x)))
    s.phi.index_copy_(0, ws.dn, ws.dv)
    ws.eval_residual(s, 1.5, 0.95, 0.05)
    ws.assemble_coupled(s, 1.5, 0.95, 0.05, None)
    M = ws.build_preconditioner(s, "jacobian_spd", 0.9, se=ws.se_used)
    print(f"[arnoldi] channel 24x8x6 n_tets={m.n_tets} n_active={ws.jac.n_active}",
          flush=True)
    assert ws.jac.n_active > 0
    return ws, s, M


@pytest.mark.parametrize("basis", ["f64", "f32"])
@pytest.mark.parametrize("with_amg", [True, False])
def test_one_call_step_equals_launch_by_launch(newton_system, basis, with_amg):
    import torch

    from up3d_amd.device import to_host
    from up3d_amd.solve.linear import gmres_dev

    ws, s, M = newton_system
    b = torch.empty_like(s.R)
    b.copy_(s.R).mul_(-1.0)
    op = ws.jac
    out = {}
    for label, mv, ps in (("one call", op.matvec, M.apply if with_amg else None),
                          ("by launch", lambda x, y: op.matvec(x, y),
                           (lambda x, y: M.apply(x, y)) if with_amg else None)):
        for rep in range(2):
            print(f"[arnoldi] {basis} amg={with_amg} {label} run {rep}", flush=True)
            x, info, it = gmres_dev(mv, b, ps, rtol=1e-10, restart=30, maxiter=8,
                                    basis=basis)
            torch.cuda.synchronize()
            out[(label, rep)] = (to_host(x), info, it)
    ref = out[("by launch", 0)]
    for key, val in out.items():
        assert np.array_equal(val[0], ref[0]), key
        assert val[1:] == ref[1:], key
    assert ref[2] > 30            # at least one restart


def test_newton_solve_through_one_call_steps(gpu):
    """A Newton Mach ramp 0.6 -> 0.75 (its GMRES takes the one-call steps)
    equals the same ramp with the steps issued launch by launch."""
    from up3d_amd.mesh.channel import boundary_nodes, bump_channel
    from up3d_amd.solve import linear
    from up3d_amd.solve.newton import solve_newton_transonic

    m = bump_channel(24, 8, 4, jitter=0.15, seed=4, rcm=True)
    dn = boundary_nodes(m, ["inflow", "outflow"])

    def ramp():
        return solve_newton_transonic(
            m.nodes, m.elements, dn, m.nodes[dn, 0].copy(), 0.75, m_start=0.6, dm=0.05,
            upwind_c=2.0, newton_kw={"phi_init": m.nodes[:, 0].copy(), "n_picard_seed": 0,
                                     "precond_matrix": "jacobian_spd", "gmres_restart": 40,
                                     "n_newton_max": 30, "amg_rebuild_every": 0})

    r1 = ramp()
    saved = linear._arnoldi_args
    linear._arnoldi_args = lambda *a: None
    try:
        r2 = ramp()
    finally:
        linear._arnoldi_args = saved
    print(f"[arnoldi] ramp levels {r1['level_history']} gmres {r1['n_gmres_total']}",
          flush=True)
    assert r1["converged"] and r2["converged"]
    assert np.array_equal(r1["phi"], r2["phi"])
    assert r1["residual_history"] == r2["residual_history"]
    assert r1["level_history"] == r2["level_history"]
    assert r1["n_gmres_total"] == r2["n_gmres_total"]

"""Lockstep Newton solves (up3d_amd/solve/batch.py) against the same solves
run one at a time:

  * several transonic Mach ramps on one channel, at different M_inf and
    alpha, advanced together by the BatchExecutor -- with the batched
    Arnoldi steps and residual evaluations (one launch per kernel for every
    solve on the mesh) and with per-solve launches -- give every solve's
    potential, residual history and Krylov count bit for bit as the solve
    run alone, for the float64 and the float32-basis precisions;
  * the batched sweep converges every condition of a small two-mesh sweep
    and agrees with the sequential sweep on its subcritical conditions."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def channel(gpu):
    from up3d_amd.mesh.channel import boundary_nodes, bump_channel

    m = bump_channel(24, 8, 4, jitter=0.15, seed=5, rcm=True)
    dn = boundary_nodes(m, ["inflow", "outflow"])
    print(f"[batch] channel 24x8x4 n_tets={m.n_tets}", flush=True)
    return m, dn


CONDS = [(0.80, 1.0), (0.74, 2.5), (0.82, 0.0)]


def _ramp(mesh, dn, m_inf, alpha, precision, stepped):
    from up3d_amd.constraints.dirichlet import freestream_phi
    from up3d_amd.solve.newton import transonic_steps

    dv = freestream_phi(mesh.nodes[dn], alpha)
    phi0 = freestream_phi(mesh.nodes, alpha)
    kw = {"precision": precision, "gmres_restart": 40, "gmres_maxiter": 5,
          "n_newton_max": 30, "amg_rebuild_every": 0, "precond_matrix": "jacobian_spd",
          "amg_max_coarse": 200, "phi_init": phi0, "n_picard_seed": 0}
    return transonic_steps(mesh.nodes, mesh.elements, dn, dv, m_inf, m_start=0.7,
                           dm=0.05, upwind_c=2.0, newton_kw=kw, stepped=stepped)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("batched", [True, False])
def test_lockstep_solves_equal_solves_alone(channel, precision, batched):
    from up3d_amd.solve.batch import BatchExecutor
    from up3d_amd.solve.linear import drive

    mesh, dn = channel
    alone = []
    for m, a in CONDS:
        print(f"[batch] alone {precision} M{m} a{a}", flush=True)
        alone.append(drive(_ramp(mesh, dn, m, a, precision, False)))
    ex = BatchExecutor(batched=batched)
    print(f"[batch] lockstep {precision} batched={batched}", flush=True)
    res = ex.run({i: _ramp(mesh, dn, m, a, precision, True)
                  for i, (m, a) in enumerate(CONDS)})
    st = ex.stats()
    print(f"[batch] executor {st}", flush=True)
    assert st["max_steps_per_round"] == len(CONDS)
    if batched:
        assert st["batched_groups"] > 0 and st["batched_eval_groups"] > 0
        assert st["grouped_vcycles"] > 0
    for i, ref in enumerate(alone):
        r = res[i]
        assert ref["converged"] and r["converged"]
        assert np.array_equal(r["phi"], ref["phi"])
        assert r["residual_history"] == ref["residual_history"]
        assert r["n_gmres_total"] == ref["n_gmres_total"]
        assert r["level_history"] == ref["level_history"]
        assert r["n_term3_active"] > 0


@pytest.mark.parametrize("by_mesh", [False, True])
def test_batched_sweep_converges_and_matches_sequential(gpu, by_mesh):
    from up3d_amd.mesh.channel import bump_channel
    from up3d_amd.solve.batch import solve_sweep_batched
    from up3d_amd.solve.sweep import solve_sweep

    meshes = [bump_channel(24, 8, 4, jitter=0.15, seed=21, rcm=True),
              bump_channel(30, 10, 4, jitter=0.15, seed=22, rcm=True)]
    conds = [(0, 0.62, 0.5), (0, 0.70, 2.0), (0, 0.78, 1.0), (0, 0.66, 2.8),
             (1, 0.64, 1.5), (1, 0.76, 0.3), (1, 0.72, 2.2)]
    kw = {"gmres_restart": 40, "amg_max_coarse": 200}
    print("[batch] sweep: batched", flush=True)
    rb = solve_sweep_batched(meshes, conds, chains_per_mesh=2, newton_kw=kw,
                             mesh_by_mesh=by_mesh)
    print(f"[batch] sweep executor {rb['executor']} wall {rb['wall_s']:.2f}s", flush=True)
    assert rb["n_converged"] == len(conds)
    assert rb["executor"]["batched_groups"] > 0
    print("[batch] sweep: sequential", flush=True)
    rs = solve_sweep(meshes, conds, newton_kw=kw)
    assert rs["n_converged"] == len(conds)
    for a, b in zip(rb["instances"], rs["instances"]):
        assert a["residual_inf"] < 1e-10
        if a["m_inf"] < 0.7:          # subcritical: one discrete solution
            assert np.abs(np.array(a["cf"]) - np.array(b["cf"])).max() < 1e-8


def test_coarse_route_sweep_sequential_equals_lockstep(gpu):
    """Supercritical conditions started from a companion-channel solve (the
    coarse route): every condition converges by that route, and the
    lockstep sweep with one chain per mesh reproduces the sequential sweep's
    wall forces bit for bit."""
    from up3d_amd.mesh.channel import bump_channel
    from up3d_amd.solve.batch import solve_sweep_batched
    from up3d_amd.solve.sweep import companion_mesh, solve_sweep

    meshes = [bump_channel(36, 12, 4, jitter=0.15, seed=31, rcm=True)]
    comp = companion_mesh((18, 6, 4), seed=32)
    conds = [(0, 0.78, 1.0), (0, 0.74, 2.0), (0, 0.64, 0.5)]
    kw = {"gmres_restart": 40, "amg_max_coarse": 200}
    print(f"[batch] coarse route: {len(meshes[0].elements)} tets, companion "
          f"{len(comp[0].elements)}", flush=True)
    rs = solve_sweep(meshes, conds, newton_kw=kw, coarse=comp, coarse_min_tets=0,
                     coarse_min_mach=0.72)
    rb = solve_sweep_batched(meshes, conds, chains_per_mesh=1, newton_kw=kw, coarse=comp,
                             coarse_min_tets=0, coarse_min_mach=0.72)
    # coarse-route conditions do not depend on their chain: one chain per
    # condition, mesh by mesh, gives them the same bits
    rw = solve_sweep_batched(meshes, conds, chains_per_mesh=len(conds), newton_kw=kw,
                             coarse=comp, coarse_min_tets=0, coarse_min_mach=0.72,
                             mesh_by_mesh=True)
    assert rw["executor"]["max_steps_per_round"] > 1
    for a, b in zip(rs["instances"][:2], rw["instances"][:2]):
        assert a["route"] == b["route"] == "coarse" and a["cf"] == b["cf"]
    assert rs["n_converged"] == rb["n_converged"] == len(conds)
    routes = [i["route"] for i in rs["instances"]]
    print(f"[batch] routes {routes}", flush=True)
    assert routes[:2] == ["coarse", "coarse"] and routes[2] != "coarse"
    for a, b in zip(rs["instances"], rb["instances"]):
        assert a["route"] == b["route"] and a["n_newton"] == b["n_newton"]
        assert a["cf"] == b["cf"] and a["residual_inf"] < 1e-10

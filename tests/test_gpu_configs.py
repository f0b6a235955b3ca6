"""The workloads at their full sizes on the device:

  * 48x16x16 channel (73,728 tets), M 0.5, Neumann flux in/out with one
    pinned node: the Picard solve equals the oracle's Picard (direct sparse
    solves) in phi, wall Cp and wall force to 1e-8;
  * 96x32x32 jittered channel (589,824 tets), M 0.80: the adjoint of the
    wall x-force equals the forward tangent, and central differences (step
    1e-6) of the solve with the upstream selection held, along seeded
    unit-norm floor perturbations and in M_inf / alpha, to 1e-6;
  * the four jittered sweep channels (144,000 / 589,824 / 1,152,000 /
    2,387,718 tets), two seeded conditions each: every instance converges to
    ||R_free||inf < 1e-10 with 0 limited / floored elements;
  * the embedded swept plate in a 128x96x96 box (7,077,888 tets), M 0.78,
    alpha 2 deg: level-set Newton to ||R||inf < 1e-10 with 0 limited /
    floored elements, and the adjoint dCL/dalpha equal to the tangent."""

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_config1_picard_full_size_matches_oracle(gpu):
    from up3d_amd.mesh.channel import config_mesh, neumann_flux_rhs
    from up3d_amd.post.surface import wall_force_coefficients
    from up3d_amd.solve.picard import solve_subsonic

    m = config_mesh("c1")
    print(f"[configs] c1 n_tets={m.n_tets}", flush=True)
    b = neumann_flux_rhs(m)
    pin = np.array([int(np.argmin(np.linalg.norm(m.nodes - [0, 0.5, 0.5], axis=1)))])
    vals = m.nodes[pin, 0]
    res = solve_subsonic(m.nodes, m.elements, pin, vals, 0.5, body_source_rhs=b)
    ref = O.solve_subsonic(m.nodes, m.elements, pin, vals, 0.5, body_source_rhs=b)
    assert res["converged"] and ref["converged"]
    assert res["residual_history"][-1] < 1e-10
    rel = np.abs(res["phi"] - ref["phi"]).max() / np.abs(ref["phi"]).max()
    wf = m.boundary_faces["floor"]
    f1 = wall_force_coefficients(m.nodes, m.elements, wf, res["phi"], m_inf=0.5)
    f0 = O.wall_force_coefficients(m.nodes, m.elements, wf, ref["phi"], m_inf=0.5)
    print(f"[configs] c1: {res['n_picard']} Picard its, phi rel {rel:.1e}, "
          f"cf_x {f1['cf'][0]:.10f} vs {f0['cf'][0]:.10f}", flush=True)
    assert rel < 1e-8
    assert np.abs(f1["cp_tri"] - f0["cp_tri"]).max() < 1e-8
    assert np.abs(f1["cf"] - f0["cf"]).max() < 1e-8


def test_config3_adjoint_full_size(gpu):
    from up3d_amd.constraints.dirichlet import freestream_phi
    from up3d_amd.mesh.channel import boundary_nodes, bump_channel, config_mesh
    from up3d_amd.solve.adjoint import ChannelForceProblem
    from up3d_amd.solve.sequencing import solve_newton_sequenced

    m = config_mesh("c3")
    dn = boundary_nodes(m, ["inflow", "outflow"])
    floor = np.unique(m.boundary_faces["floor"])
    prob = ChannelForceProblem(m.nodes, m.elements, dn, m.boundary_faces["floor"], floor,
                               newton_kw={"amg_max_coarse": 1000})
    print(f"[configs] c3 n_tets={m.n_tets} floor nodes={len(floor)}", flush=True)
    m_inf, alpha = 0.80, 0.0
    # grid-sequenced state: Mach ramp at upwind_c 3 on 48x16x16, then the
    # 96x32x32 channel from its interpolant at upwind_c 3 -> 2 -> 1.5
    coarse = bump_channel(48, 16, 16, jitter=0.15, seed=20260103, rcm=True)
    dc = boundary_nodes(coarse, ["inflow", "outflow"])
    seq = solve_newton_sequenced(
        [{"nodes": coarse.nodes, "elements": coarse.elements, "dirichlet_nodes": dc,
          "dirichlet_values": freestream_phi(coarse.nodes[dc], alpha)},
         {"nodes": m.nodes, "elements": m.elements, "dirichlet_nodes": dn,
          "dirichlet_values": freestream_phi(m.nodes[dn], alpha)}],
        m_inf, m_start=0.6, dm=0.05, upwind_c_ramp=3.0, upwind_c_post=(2.0, 1.5),
        post_stages_on="finest",
        newton_kw={"gmres_restart": 200, "gmres_maxiter": 5, "n_newton_max": 30,
                   "amg_rebuild_every": 0, "precond_matrix": "jacobian_spd",
                   "amg_max_coarse": 1000})
    assert seq["converged"]
    sol = prob.solve(None, m_inf, alpha, phi_init=seq["phi"])
    print(f"[configs] c3 M0.80: converged={sol['converged']} frozen residual at "
          f"acceptance {sol['frozen_residual_at_accept']:.1e}, "
          f"{sol['frozen_polish_steps']} settling steps", flush=True)
    assert sol["converged"] and sol["n_limited"] == 0
    frozen = prob.frozen_selection(sol)
    g = prob.gradients(sol, weights=(1.0, 0.0))
    assert g["adjoint_info"] == 0
    tan_m = prob.jvp_params(sol, 1.0, 0.0)
    tan_a = prob.jvp_params(sol, 0.0, 1.0)
    assert abs(g["d_m_inf"] - tan_m) <= 1e-8 * abs(tan_m)
    assert abs(g["d_alpha_deg"] - tan_a) <= 1e-8 * max(abs(tan_a), 1e-12)
    h = 1e-6

    def F(y=None, mm=m_inf, a=alpha):
        s = prob.solve_frozen(y, mm, a, frozen, sol["phi"])
        assert s["converged"], s["residual_history"]
        return prob.functional(s)[0]

    fd_m = (F(mm=m_inf + h) - F(mm=m_inf - h)) / (2 * h)
    fd_a = (F(a=alpha + h) - F(a=alpha - h)) / (2 * h)
    errs = [abs(g["d_m_inf"] - fd_m) / abs(fd_m)]
    if abs(fd_a) > 1e-9:
        errs.append(abs(g["d_alpha_deg"] - fd_a) / abs(fd_a))
    y0 = prob.base_nodes[prob.floor_nodes, 1]
    rng = np.random.default_rng(16)
    for k in range(4):
        d = rng.standard_normal(len(y0))
        d /= np.linalg.norm(d)
        fd = (F(y0 + h * d) - F(y0 - h * d)) / (2 * h)
        errs.append(abs(float(g["d_y_floor"] @ d) - fd) / abs(fd))
    print(f"[configs] c3 adjoint vs central differences (h 1e-6): "
          f"{['%.1e' % e for e in errs]}", flush=True)
    assert max(errs) <= 1e-6


def test_config4_sweep_every_mesh(gpu):
    from up3d_amd.solve.sweep import solve_sweep, sweep_conditions, sweep_meshes

    meshes = sweep_meshes()
    conds = sweep_conditions(len(meshes), 2)
    print(f"[configs] c4 meshes {[len(m.elements) for m in meshes]} conditions "
          f"{[(k, round(mm, 3), round(a, 2)) for k, mm, a in conds]}", flush=True)
    r = solve_sweep(meshes, conds)
    for i in r["instances"]:
        print(f"[configs] c4 mesh {i['mesh']} M{i['m_inf']:.3f} a{i['alpha_deg']:.2f}: "
              f"converged={i['converged']} |R|={i['residual_inf']:.1e} "
              f"n={i['n_newton']} {i['seconds']:.1f}s", flush=True)
    assert r["n_converged"] == len(conds)
    for i in r["instances"]:
        assert i["residual_inf"] < 1e-10
    assert sorted({i["mesh"] for i in r["instances"]}) == [0, 1, 2, 3]


def test_config5_plate_full_size(gpu):
    from up3d_amd.mesh.plate_box import plate_wing_box
    from up3d_amd.solve.adjoint_ls import lift_alpha_sensitivity
    from up3d_amd.solve.newton_ls import solve_multivalued_newton
    from up3d_amd.wake.plate_case import PlateWingCase

    mesh = plate_wing_box(128, 96, 96)
    print(f"[configs] c5 n_tets={len(mesh.elements)}", flush=True)
    case = PlateWingCase(mesh)
    r = solve_multivalued_newton(case.mvop, mesh, 0.78, phi_init=case.initial_state(2.0),
                                 n_seed=20, freeze_tol=1e-6, **case.solve_kwargs(2.0))
    y, gam = case.gamma_distribution(r["phi_ext"])
    print(f"[configs] c5: converged={r['converged']} |R|={r['residual_history'][-1]:.1e} "
          f"lim/flr={r['n_limited']}/{r['n_floored']} Gamma root/tip {gam[0]:.5f}/"
          f"{gam[-1]:.5f}", flush=True)
    assert r["converged"] and r["residual_history"][-1] < 1e-10
    assert r["n_limited"] == 0 and r["n_floored"] == 0
    assert np.all(np.diff(gam) < 0.0) and gam[-1] > 0.0
    s = lift_alpha_sensitivity(case, r["phi_ext"], 0.78, 2.0)
    print(f"[configs] c5 dCL/dalpha adjoint {s['dcl_dalpha']:.12f} tangent "
          f"{s['dcl_dalpha_tangent']:.12f}", flush=True)
    assert s["adjoint_info"] == 0 and s["tangent_info"] == 0
    assert abs(s["dcl_dalpha"] - s["dcl_dalpha_tangent"]) <= 1e-9 * abs(s["dcl_dalpha"])

"""Level-set Newton on the device (pyfp3d/solve/newton_ls.py), on the
reference's NACA0012 2.5D meshes:

  * the assembled Jacobian (device plain part + host special / Kutta
    correction) equals central finite differences of the level-set residual
    at a frozen per-side upwind selection, with a supersonic pocket present
    (nu > 0 elements), for directions on every DOF, on the aux DOFs alone
    and on the sheet;
  * subsonic M0.5 alpha 1.25: the circulation of the reference's level-set
    Picard (0.087173, docs/demo_report/track_a.md:146-151);
  * transonic alpha 1.25, per-side upwinding C = 1.5: the converged level-set
    circulations of the reference's table (docs/design_track_b.md:799,
    873-878; 4 significant digits), the same state from a Picard seed at the
    target Mach and from the Mach ramp (the discrete solution is isolated),
    0 limited / floored, shock pocket M_max in (1, 1.3);
  * medium M0.70: the reference's figure there (0.1193,
    docs/design_track_b.md:879,882) is a level-set PICARD iterate stopped at
    ||R||inf ~ 1e-7; the same procedure here gives it to 0.1 %, and Newton
    started from that very Picard state converges (||R|| < 1e-10) to the
    same root as the Newton Mach ramp, 0.118312: the Picard's circulation
    mode converges slowly, and continued past that stop the Picard iterate
    walks onto the Newton root (0.118307 at ||R|| 2e-9 after 3000 outers)."""

import json
from pathlib import Path

import numpy as np
import pytest

from tests.test_gpu_wake_ls import _mesh, _setup

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).parent / "golden"


def _gold():
    return json.loads((GOLD / "golden.json").read_text())["ls_gamma_naca_a125"]["values"]


def test_ls_jacobian_matches_central_differences(gpu):
    from up3d_amd.device import empty, to_dev, to_host
    from up3d_amd.solve.newton_ls import LSNewtonSystem
    from up3d_amd.solve.picard_ls import _farfield_main, solve_multivalued_lifting

    mesh = _mesh("coarse")
    print(f"[wake-newton] jacobian FD: coarse M0.70 n_tets={len(mesh.elements)}", flush=True)
    _, mv = _setup(mesh)
    seed = solve_multivalued_lifting(mv, mesh, 0.70, alpha_deg=1.25, upwind_c=1.5,
                                     damping_theta=0.2, n_outer_max=30)
    phi = seed["phi_ext"].copy()
    ff, vals = _farfield_main(mesh.nodes, mesh.boundary_faces["farfield"], 1.25,
                              seed["gamma"], 1.0, (0.25, 0.0), np.sqrt(1 - 0.49))
    phi[ff] = vals
    sysm = LSNewtonSystem(mv, 0.70)
    fz = sysm.freeze(phi)
    R0, sides, C, w = sysm.residual(phi, fz, sens=True)
    assert mv.n_nu_active > 0 and mv.n_limited == 0 and mv.n_floored == 0
    sysm.linearize(phi, sides, C, w, ff)
    free = np.ones(mv.n_total, dtype=bool)
    free[ff] = False
    rng = np.random.default_rng(1)
    sheet = np.unique(mv.el[mv.special])
    for kind in ("all", "aux", "sheet"):
        v = np.zeros(mv.n_total)
        if kind == "all":
            v[:] = rng.standard_normal(mv.n_total)
        else:
            if kind == "sheet":
                v[sheet] = rng.standard_normal(len(sheet))
            v[mv.n_main:] = rng.standard_normal(mv.n_total - mv.n_main)
        v[~free] = 0.0
        Jv = to_host(sysm.matvec(to_dev(v), empty(mv.n_total)))
        h = 1e-6
        fd = (sysm.residual(phi + h * v, fz)[0] - sysm.residual(phi - h * v, fz)[0]) / (2 * h)
        err = float(np.abs(Jv - fd)[free].max())
        scale = float(np.abs(fd[free]).max())
        print(f"[wake-newton] {kind}: |Jv - FD| = {err:.2e} of {scale:.2e}", flush=True)
        assert err <= 1e-7 * scale


def test_ls_newton_subsonic_gamma(gpu):
    from up3d_amd.solve.newton_ls import solve_multivalued_newton

    mesh = _mesh("coarse")
    _, mv = _setup(mesh)
    r = solve_multivalued_newton(mv, mesh, 0.5, n_seed=10)
    print(f"[wake-newton] M0.5: gamma={r['gamma']:.6f} n={r['n_newton']} "
          f"|R|={r['residual_history'][-1]:.2e}", flush=True)
    assert r["converged"] and r["accept_reason"] == "tol"
    assert r["residual_history"][-1] < 1e-10
    assert abs(r["gamma"] - 0.087173) <= 2e-6


def test_ls_newton_transonic_matches_reference_table(gpu):
    from up3d_amd.solve.newton_ls import (solve_multivalued_newton,
                                          solve_multivalued_newton_transonic)

    gold = _gold()["coarse"]
    mesh = _mesh("coarse")
    _, mv = _setup(mesh)
    ramp = solve_multivalued_newton_transonic(mv, mesh, 0.75, dm=0.05)
    print("[wake-newton] ramp " + str([(lv["m_inf"], round(lv["gamma"], 6), lv["n_newton"],
                                        lv["converged"], round(lv["mach_max"], 4))
                                       for lv in ramp["levels"]]), flush=True)
    assert ramp["target_reached"]
    for lv in ramp["levels"]:
        assert lv["converged"] and lv["n_limited"] == 0 and lv["n_floored"] == 0
        key = f"{lv['m_inf']:.2f}"
        if key in gold:
            assert abs(lv["gamma"] - gold[key]) <= 0.5e-4, (key, lv["gamma"])
    m70 = [lv for lv in ramp["levels"] if abs(lv["m_inf"] - 0.70) < 1e-12][0]
    assert 1.0 < m70["mach_max"] < 1.3
    direct = solve_multivalued_newton(mv, mesh, 0.70, n_seed=40, freeze_tol=1e-6)
    print(f"[wake-newton] M0.70 from a Picard seed: gamma={direct['gamma']:.8f} "
          f"(ramp {m70['gamma']:.8f})", flush=True)
    assert direct["converged"]
    assert abs(direct["gamma"] - m70["gamma"]) < 1e-8


def test_ls_newton_medium_m070(gpu):
    from up3d_amd.solve.newton_ls import (LSNewtonSystem, solve_multivalued_newton,
                                          solve_multivalued_newton_transonic)
    from up3d_amd.solve.picard_ls import _farfield_main, solve_multivalued_transonic

    mesh = _mesh("medium")
    print(f"[wake-newton] medium M0.70 n_tets={len(mesh.elements)}", flush=True)
    _, mv = _setup(mesh)
    # the reference's procedure: level-set Picard Mach ramp, stopped at
    # ||R||inf < 1e-7 (tol_residual); it reproduces the quoted 0.1193 to 0.2 %
    # (the stop falls at 0.1192-0.1194 depending on the far-field variant)
    pic = solve_multivalued_transonic(mv, mesh, 0.70, n_outer_level=600)
    ref = _gold()["medium"]["0.70"]
    print(f"[wake-newton] medium M0.70 Picard (stop at 1e-7): gamma={pic['gamma']:.6f}",
          flush=True)
    assert pic["converged"] and pic["n_limited"] == 0 and pic["n_floored"] == 0
    assert abs(pic["gamma"] - ref) <= 0.002 * ref
    # ...but that state is not a discrete solution at the 1e-10 bar
    phi = pic["phi_ext"].copy()
    ff, vals = _farfield_main(mesh.nodes, mesh.boundary_faces["farfield"], 1.25,
                              float(pic["gamma"]), 1.0, (0.25, 0.0), np.sqrt(1 - 0.49))
    phi[ff] = vals
    free = np.ones(mv.n_total, dtype=bool)
    free[ff] = False
    R, _, _, _ = LSNewtonSystem(mv, 0.70).residual(phi, None)
    r_pic = float(np.abs(R[free]).max())
    assert r_pic > 1e-9, r_pic
    # Newton from the Picard state and the Newton Mach ramp reach one root
    from_pic = solve_multivalued_newton(mv, mesh, 0.70, phi_init=pic["phi_ext"],
                                        gamma_init=float(pic["gamma"]), n_seed=0,
                                        freeze_tol=1e-6)
    r = solve_multivalued_newton_transonic(mv, mesh, 0.70)
    print(f"[wake-newton] medium M0.70: Newton from Picard gamma={from_pic['gamma']:.8f} "
          f"(|R| {from_pic['residual_history'][-1]:.1e}), ramp gamma={r['gamma']:.8f}, "
          f"Picard state |R|={r_pic:.1e}", flush=True)
    for x in (from_pic, r):
        assert x["converged"] and x["residual_history"][-1] < 1e-10
        assert x["n_limited"] == 0 and x["n_floored"] == 0
    assert r["target_reached"]
    assert abs(from_pic["gamma"] - r["gamma"]) < 1e-8
    # the root sits below the Picard stop, as the continued Picard shows
    assert r["gamma"] < pic["gamma"]


def test_schur_reduced_system_solves_fused_matrix(gpu):
    """SchurReducedSystem on the free-reduced fused level-set matrix equals a
    direct solve; the aux rows hold exactly by back-substitution
    (tests/test_b14_schur_ls.py's honesty identity)."""
    import scipy.sparse.linalg as spla

    from up3d_amd.constraints.dirichlet import freestream_phi
    from up3d_amd.solve.schur_ls import SchurReducedSystem, main_block_preconditioner

    mesh = _mesh("coarse")
    _, mv = _setup(mesh)
    phi = np.zeros(mv.n_total)
    phi[: mv.n_main] = freestream_phi(mesh.nodes, 2.0)
    phi[mv.cut_aux] = phi[mv.cut_nodes] - 0.1 * mv.cm.node_side[mv.cut_nodes]
    E = mv.assemble_matrix(closure="wake_ls", te_kutta="pressure", phi_ext=phi)
    ff = np.unique(mesh.boundary_faces["farfield"])
    free = np.setdiff1d(np.arange(mv.n_total), ff)
    A = E[free][:, free].tocsr()
    b = np.random.default_rng(3).standard_normal(len(free))
    S = SchurReducedSystem(A, free, mv.n_main, n_aux_expected=mv.n_ext)
    M = main_block_preconditioner(mv, None, S.main_free)
    x, n_it, info = S.solve(b, M=M, rtol=1e-12, restart=60, maxiter=40)
    xd = spla.spsolve(A.tocsc(), b)
    print(f"[wake-newton] schur: {n_it} GMRES iterations, |x - direct| = "
          f"{np.abs(x - xd).max():.2e}", flush=True)
    assert info == 0
    assert np.abs(x - xd).max() <= 1e-8 * np.abs(xd).max()
    # aux rows exact to LU roundoff: componentwise backward error
    r = A @ x - b
    scale = abs(A) @ np.abs(x) + np.abs(b)
    assert (np.abs(r[S.n_mf:]) / scale[S.n_mf:]).max() <= 1e-10
    with pytest.raises(ValueError):
        SchurReducedSystem(A, free, mv.n_main, n_aux_expected=mv.n_ext + 1)


def test_ls_newton_schur_mode(gpu):
    from up3d_amd.solve.newton_ls import solve_multivalued_newton

    mesh = _mesh("coarse")
    _, mv = _setup(mesh)
    r = solve_multivalued_newton(mv, mesh, 0.70, n_seed=40, freeze_tol=1e-6,
                                 linear="schur")
    print(f"[wake-newton] schur M0.70: gamma={r['gamma']:.8f} n={r['n_newton']} "
          f"gmres={r['n_gmres_total']}", flush=True)
    assert r["converged"] and r["n_gmres_stalled"] == 0
    assert abs(r["gamma"] - _gold()["coarse"]["0.70"]) <= 0.5e-4

"""The numerical failure-recovery paths of the drivers, forced from the
host so every kernel still exits normally:

  * GMRES retry at twice the restart, then RuntimeError
    (pyfp3d/solve/linear.py:186-201), and on_fail="return";
  * a Newton level that runs out of its n_newton_max budget;
  * the level fail-fast (pyfp3d/solve/newton.py:860-863): a residual 100x
    above the level's best ends the level unconverged;
  * the Mach ramp's step halving around a level that fails
    (pyfp3d/solve/newton.py:1259-1266), end to end on the device."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _channel():
    from up3d_amd.mesh.channel import boundary_nodes, bump_channel

    m = bump_channel(24, 8, 4, jitter=0.15, seed=5, rcm=True)
    dn = boundary_nodes(m, ["inflow", "outflow"])
    return m, dn


def test_gmres_retry_at_double_restart(gpu, monkeypatch):
    import scipy.sparse as sp

    from up3d_amd.solve import linear as L

    restarts = []
    real = L.gmres_dev

    def spy(mv, b, ps, x0, rtol, atol, restart, maxiter, *a, **k):
        restarts.append(restart)
        return real(mv, b, ps, x0, rtol, atol, restart, maxiter, *a, **k)

    monkeypatch.setattr(L, "gmres_dev", spy)
    rng = np.random.default_rng(7)
    # 4 unknowns: GMRES(2) x 1 cycle cannot finish, GMRES(4) solves exactly
    A4 = sp.csr_matrix(np.eye(4) * 4.0 + rng.standard_normal((4, 4)))
    b4 = rng.standard_normal(4)
    print("[recovery] gmres retry 4x4", flush=True)
    x, n_it, info = L.solve_gmres(A4, b4, rtol=1e-10, restart=2, maxiter=1)
    assert restarts == [2, 4] and info == 0 and n_it > 2
    assert np.allclose(x, np.linalg.solve(A4.toarray(), b4), rtol=1e-8, atol=1e-12)
    # 8 unknowns: both attempts fail -> RuntimeError naming both restarts
    restarts.clear()
    A8 = sp.csr_matrix(np.eye(8) * 4.0 + rng.standard_normal((8, 8)))
    b8 = rng.standard_normal(8)
    print("[recovery] gmres retry failure 8x8", flush=True)
    with pytest.raises(RuntimeError, match="restart 2 then 4"):
        L.solve_gmres(A8, b8, rtol=1e-12, restart=2, maxiter=1)
    assert restarts == [2, 4]
    # on_fail="return": one attempt, nonzero info, no exception
    restarts.clear()
    x, n_it, info = L.solve_gmres(A8, b8, rtol=1e-12, restart=2, maxiter=1,
                                  on_fail="return")
    assert restarts == [2] and info != 0 and n_it == 2 and np.all(np.isfinite(x))


def test_newton_budget_exhausted(gpu):
    from up3d_amd.solve.newton import solve_newton

    m, dn = _channel()
    print(f"[recovery] n_newton_max=2 at M0.80, n_tets={m.n_tets}", flush=True)
    r = solve_newton(m.nodes, m.elements, dn, m.nodes[dn, 0], 0.80,
                     phi_init=m.nodes[:, 0].copy(), n_picard_seed=0, n_newton_max=2)
    assert not r["converged"] and r["accept_reason"] is None
    assert len(r["residual_history"]) == 2 and r["residual_history"][-1] > 1e-10


def test_level_fail_fast(gpu, monkeypatch):
    from up3d_amd.solve import newton as N

    m, dn = _channel()
    real = N.NewtonWorkspace.eval_residual
    calls = [0]

    def inflated(self, state, *a, **k):
        out = real(self, state, *a, **k)
        calls[0] += 1
        if calls[0] >= 3:          # from the second Newton step's state on
            state.r_norm = max(state.r_norm, 1e-3) * 1e4
        return out

    monkeypatch.setattr(N.NewtonWorkspace, "eval_residual", inflated)
    print("[recovery] fail-fast with an inflated residual norm", flush=True)
    r = N.solve_newton(m.nodes, m.elements, dn, m.nodes[dn, 0], 0.5,
                       phi_init=m.nodes[:, 0].copy(), n_picard_seed=0, n_newton_max=20,
                       line_search=False)
    h = r["residual_history"]
    assert not r["converged"]
    assert len(h) < 20 and h[-1] > 100.0 * min(h[:-1]) and h[-1] > 1e-3


def test_mach_ramp_halves_after_a_failed_level(gpu, monkeypatch):
    from up3d_amd.solve import newton as N

    m, dn = _channel()
    real = N.solve_newton
    seen = []

    def fail_first_top(*a, m_inf, **k):
        seen.append(round(m_inf, 10))
        if round(m_inf, 10) == 0.7 and seen.count(0.7) == 1:
            k = dict(k, n_newton_max=1)    # a level that cannot converge
        return real(*a, m_inf=m_inf, **k)

    monkeypatch.setattr(N, "solve_newton", fail_first_top)
    print("[recovery] mach ramp 0.6 -> 0.7 with the first 0.7 attempt starved",
          flush=True)
    r = N.solve_newton_transonic(m.nodes, m.elements, dn, m.nodes[dn, 0], 0.70,
                                 m_start=0.60, dm=0.05)
    assert seen == [0.6, 0.65, 0.7, 0.675, 0.7]
    conv = [lv["converged"] for lv in r["level_results"]]
    assert conv == [True, True, False, True, True]
    assert r["converged"] and r["residual_history"][-1] < 1e-10

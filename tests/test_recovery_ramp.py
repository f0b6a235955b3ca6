"""Mach-step halving of the Newton continuation (pyfp3d/solve/newton.py:
1259-1266) on the CPU: the per-level solve is replaced by a stand-in that
fails where the test says, so the ramp's control flow -- retry at the
midpoint of the failed step, warm start from the last converged level only,
give up once the step would drop below dm_min or the first level fails --
is checked without a device."""

import pytest

from up3d_amd.solve import newton as N
from up3d_amd.solve.timing import new_timings


class _FakeLevels:
    def __init__(self, fail):
        self.fail = fail          # fail(m, attempt_index_at_m) -> bool
        self.calls = []           # (m, phi_init tag)
        self.seen = {}

    def __call__(self, nodes, elements, dn, dv, m_inf, workspace=None, verbose=False,
                 **kw):
        k = self.seen.get(round(m_inf, 10), 0)
        self.seen[round(m_inf, 10)] = k + 1
        self.calls.append((round(m_inf, 10), kw.get("phi_init")))
        ok = not self.fail(round(m_inf, 10), k)
        return {"converged": ok, "accept_reason": "tol" if ok else None,
                "n_newton": 3, "residual_history": [1.0, 1e-11 if ok else 1e-3],
                "step_records": [], "n_gmres_total": 7, "timings": new_timings(),
                "workspace": "ws", "phi_dev": f"phi@{m_inf:.4f}"}


def _run(monkeypatch, fail, **kw):
    fake = _FakeLevels(fail)
    monkeypatch.setattr(N, "solve_newton", fake)
    r = N.solve_newton_transonic(None, None, None, None, kw.pop("m_inf", 0.80),
                                 m_start=0.70, dm=0.05, **kw)
    return r, fake


def test_failed_level_is_retried_at_half_step(monkeypatch):
    # 0.80 fails once: the ramp inserts 0.775, then reaches 0.80 from it
    r, fake = _run(monkeypatch, lambda m, k: m == 0.8 and k == 0)
    ms = [m for m, _ in fake.calls]
    assert ms == [0.7, 0.75, 0.8, 0.775, 0.8]
    assert r["converged"]
    # warm starts come from the last converged level, never the failed one
    seeds = [p for _, p in fake.calls]
    assert seeds[0] is None
    assert seeds[3] == "phi@0.7500" and seeds[4] == "phi@0.7750"
    assert [lv["m"] for lv in r["level_results"]] == ms
    assert [lv["converged"] for lv in r["level_results"]] == [True, True, False, True, True]
    assert len(r["level_history"]) == 5
    assert r["timings_total"]["wall"] == 0.0 and "workspace" not in r


def test_halving_stops_below_dm_min(monkeypatch):
    # 0.80 never converges: 0.775, 0.7625 fail too; the next half step
    # (0.00625) is below dm_min = 0.01, so the ramp stops unconverged
    r, fake = _run(monkeypatch, lambda m, k: m > 0.75)
    ms = [m for m, _ in fake.calls]
    assert ms == [0.7, 0.75, 0.8, 0.775, 0.7625]
    assert not r["converged"]


def test_first_level_failure_ends_the_ramp(monkeypatch):
    r, fake = _run(monkeypatch, lambda m, k: m == 0.7)
    assert [m for m, _ in fake.calls] == [0.7]
    assert not r["converged"]


def test_upward_only(monkeypatch):
    with pytest.raises(ValueError):
        _run(monkeypatch, lambda m, k: False, m_inf=0.6)

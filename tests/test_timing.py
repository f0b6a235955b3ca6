"""The `timings` schema helpers (up3d_amd/solve/timing.py) on the CPU: the
key set of pyfp3d/solve/timing.py:33-71, exclusive phases, `other` as the
unexplained remainder, level totals, per-step deltas and the column view of
step records (the reference's contract test, tests/test_a1_instrumentation.py,
checks the same keys on every driver)."""

import time

import pytest

from up3d_amd.solve import timing as T


def test_schema_keys_and_zero_start():
    t = T.new_timings()
    assert tuple(t) == ("seed", "assembly", "precond", "linsolve", "residual",
                        "kutta", "wall", "other")
    assert all(v == 0.0 for v in t.values())


def test_phase_charges_only_its_key_and_rejects_unknown_keys():
    t = T.new_timings()
    with T.phase(t, "assembly", sync=False):
        time.sleep(0.02)
    with T.phase(t, "assembly", sync=False):
        time.sleep(0.01)
    assert 0.025 <= t["assembly"] < 0.5
    assert all(t[k] == 0.0 for k in t if k != "assembly")
    with pytest.raises(KeyError):
        with T.phase(t, "wall", sync=False):
            pass
    # an exception inside the block still charges the time and propagates
    with pytest.raises(ZeroDivisionError):
        with T.phase(t, "kutta", sync=False):
            time.sleep(0.01)
            1 / 0
    assert t["kutta"] >= 0.009


def test_finalize_other_is_wall_minus_phases():
    t = T.new_timings()
    t.update(seed=0.5, assembly=1.25, precond=0.25, linsolve=2.0, residual=0.75,
             kutta=0.125)
    out = T.finalize(t, 5.0)
    assert out is t
    assert t["wall"] == 5.0
    assert t["other"] == pytest.approx(5.0 - 4.875, abs=1e-15)
    assert abs(t["wall"] - (sum(t[p] for p in T.PHASES) + t["other"])) < 1e-12


def test_sum_timings_adds_levels_keywise():
    a = T.finalize(dict(T.new_timings(), assembly=1.0, linsolve=2.0), 4.0)
    b = T.finalize(dict(T.new_timings(), assembly=0.5, residual=1.0), 2.0)
    tot = T.sum_timings([a, b, {"kutta": 0.25}])
    assert tot["assembly"] == 1.5 and tot["linsolve"] == 2.0
    assert tot["residual"] == 1.0 and tot["kutta"] == 0.25
    assert tot["wall"] == 6.0 and tot["other"] == pytest.approx(1.5)
    assert set(tot) == set(T.SCHEMA)
    assert T.sum_timings([]) == T.new_timings()


def test_step_delta_and_snapshot():
    t = T.new_timings()
    prev = T.snapshot(t)
    t["linsolve"] += 0.75
    t["residual"] += 0.25
    d = T.step_delta(t, prev)
    assert set(d) == {"t_" + p for p in T.PHASES}
    assert d["t_linsolve"] == 0.75 and d["t_residual"] == 0.25
    assert d["t_seed"] == 0.0
    snap = T.snapshot(t)
    t["linsolve"] += 1.0        # a snapshot is a copy, not a view
    assert snap["linsolve"] == 0.75


def test_records_to_columns_keeps_first_seen_order():
    recs = [{"i": 0, "residual": 1.0}, {"i": 1, "eta": 0.1, "residual": 0.5}]
    cols = T.records_to_columns(recs)
    assert list(cols) == ["i", "residual", "eta"]
    assert cols["eta"] == [None, 0.1]
    assert cols["residual"] == [1.0, 0.5]
    assert T.records_to_columns([]) == {}

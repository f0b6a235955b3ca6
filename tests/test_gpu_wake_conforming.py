"""Conforming explicit-circulation lifting solve on the device
(pyfp3d/solve/picard.py:88-268 solve_laplace_lifting on
pyfp3d/mesh/wake_cut.py's cut): the circulation the reference measured on
its NACA0012 meshes with this path, alpha 2 deg, incompressible
(tests/test_b4_te_control_volume.py:186-189: conforming 0.1175 coarse,
0.1200 medium; 4 significant digits)."""

import json
from pathlib import Path

import numpy as np
import pytest

from tests.test_wake_cut import _mesh

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).parent / "golden"


@pytest.mark.parametrize("level,idx", [("coarse", 0), ("medium", 1)])
def test_conforming_laplace_lifting_gamma(gpu, level, idx):
    from up3d_amd.constraints.wake import kutta_targets
    from up3d_amd.mesh.wake_cut import cut_wake
    from up3d_amd.solve.picard import solve_laplace_lifting

    gold = json.loads((GOLD / "golden.json").read_text())["ls_gamma_naca_a2"]["values"]
    m = _mesh(level)
    mc, wc = cut_wake(m)
    print(f"[conforming] {level}: {len(mc.elements)} tets, {len(wc.slave_nodes)} slaves",
          flush=True)
    r = solve_laplace_lifting(mc, wc, alpha_deg=2.0)
    g = float(np.mean(r["gamma"]))
    print(f"[conforming] {level}: gamma={g:.6f} (reference {gold['conforming'][idx]}) "
          f"updates={r['n_kutta_updates']} cg={r['n_cg_total']}", flush=True)
    assert r["kutta_converged"] and r["n_kutta_updates"] < 20
    assert abs(g - gold["conforming"][idx]) <= 0.6e-4
    assert np.allclose(kutta_targets(r["phi"], wc), r["gamma"], atol=1e-8)
    # the prescribed-circulation path returns the same field at that Gamma
    r2 = solve_laplace_lifting(mc, wc, alpha_deg=2.0, gamma_fixed=r["gamma"])
    assert np.abs(r2["phi"] - r["phi"]).max() < 1e-8


def _gold_newton():
    return json.loads((GOLD / "golden.json").read_text())["conforming_newton_naca_a125"]["values"]


def test_conforming_newton_ramp_matches_reference(gpu):
    """Coupled (phi, Gamma) Newton, Mach ramp 0.60 -> 0.75 at alpha 1.25
    deg: the reference's conforming-Newton circulations at M 0.65 / 0.70 /
    0.75 on the coarse mesh (4 digits), each level converged with 0
    limited / floored; at M 0 it keeps the incompressible solution."""
    from up3d_amd.mesh.wake_cut import cut_wake
    from up3d_amd.solve.newton_lifting import (solve_newton_lifting,
                                               solve_newton_lifting_transonic)

    mc, wc = cut_wake(_mesh("coarse"))
    r0 = solve_newton_lifting(mc, wc, 0.0, alpha_deg=2.0)
    assert r0["converged"] and abs(float(r0["gamma"][0]) - 0.117545) < 5e-6
    r = solve_newton_lifting_transonic(mc, wc, 0.75, alpha_deg=1.25)
    print("[conforming] ramp " + str([(lv["m_inf"], round(lv["gamma"], 6), lv["n_newton"])
                                      for lv in r["levels"]]), flush=True)
    assert r["target_reached"]
    gold = _gold_newton()["coarse"]
    for lv in r["levels"]:
        assert lv["converged"] and lv["n_limited"] == 0 and lv["n_floored"] == 0
        key = f"{lv['m_inf']:.2f}"
        if key in gold:
            assert abs(lv["gamma"] - gold[key]) <= 0.5e-4, (key, lv["gamma"])


def test_conforming_newton_medium_m070(gpu):
    from up3d_amd.mesh.wake_cut import cut_wake
    from up3d_amd.solve.newton_lifting import solve_newton_lifting_transonic

    mc, wc = cut_wake(_mesh("medium"))
    r = solve_newton_lifting_transonic(mc, wc, 0.70, alpha_deg=1.25)
    g = r["levels"][-1]["gamma"]
    print(f"[conforming] medium M0.70: gamma={g:.6f} (reference 0.1190)", flush=True)
    assert r["target_reached"] and abs(g - _gold_newton()["medium"]["0.70"]) <= 0.5e-4

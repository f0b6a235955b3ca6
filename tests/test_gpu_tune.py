"""Autotuner (no reference counterpart): the element block size changes no
result bit; tuning picks candidates it measured, caches them per mesh
shape, and the tuned Newton converges like the default one."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _channel():
    from up3d_amd.mesh.channel import boundary_nodes, bump_channel

    m = bump_channel(24, 8, 8, jitter=0.1, seed=3)
    dn = boundary_nodes(m, ["inflow", "outflow"])
    return m, dn, m.nodes[dn, 0].copy()


def test_element_block_is_bitwise_neutral(gpu):
    from up3d_amd import _native
    from up3d_amd.device import to_dev, to_host
    from up3d_amd.solve.newton import NewtonWorkspace
    from up3d_amd.tune import apply_element_block

    m, dn, dv = _channel()
    ws = NewtonWorkspace(m.nodes, m.elements, dn, dv)
    ws.set_mach(0.8)
    s = ws.states[0]
    phi = m.nodes[:, 0] + 0.05 * np.sin(3 * m.nodes[:, 0]) * m.nodes[:, 1]
    out = {}
    prev = int(_native.lib().fp3d_get_element_block())
    try:
        for t in (64, 128, 192, 256):
            apply_element_block(t)
            s.phi.copy_(to_dev(phi))
            s.phi.index_copy_(0, ws.dn, ws.dv)
            ws.eval_residual(s, 1.5, 0.95, 0.05)
            ws.assemble_coupled(s, 1.5, 0.95, 0.05)
            out[t] = (to_host(s.R).copy(), to_host(ws.jac.data).copy())
    finally:
        apply_element_block(prev)
    for t in (64, 128, 192):
        assert np.array_equal(out[t][0], out[256][0])
        assert np.array_equal(out[t][1], out[256][1])
    with pytest.raises(ValueError):
        apply_element_block(100)


def test_gather_chunk_is_bitwise_neutral(gpu):
    from up3d_amd import _native
    from up3d_amd.device import to_dev, to_host
    from up3d_amd.solve.newton import NewtonWorkspace
    from up3d_amd.tune import apply_gather_chunk

    m, dn, dv = _channel()
    ws = NewtonWorkspace(m.nodes, m.elements, dn, dv)
    ws.set_mach(0.8)
    s = ws.states[0]
    phi = m.nodes[:, 0] + 0.05 * np.sin(3 * m.nodes[:, 0]) * m.nodes[:, 1]
    s.phi.copy_(to_dev(phi))
    s.phi.index_copy_(0, ws.dn, ws.dv)
    ws.eval_residual(s, 1.5, 0.95, 0.05)
    out = {}
    prev = int(_native.lib().fp3d_get_gather_chunk())
    try:
        for c in (256, 512, 1024, 2048):
            apply_gather_chunk(c)
            ws.jac.data.fill_(float("nan"))
            ws.assemble_coupled(s, 1.5, 0.95, 0.05)
            out[c] = to_host(ws.jac.data).copy()
    finally:
        apply_gather_chunk(prev)
    for c in (256, 512, 2048):
        assert np.array_equal(out[c], out[1024])
    with pytest.raises(ValueError):
        apply_gather_chunk(300)


def test_tune_newton_workspace_and_cache(gpu, tmp_path):
    from up3d_amd import _native
    from up3d_amd.solve.newton import NewtonWorkspace, solve_newton
    from up3d_amd.tune import TuningCache, tune_newton_workspace, tuned_newton_kwargs

    m, dn, dv = _channel()
    ws = NewtonWorkspace(m.nodes, m.elements, dn, dv)
    cache = TuningCache(str(tmp_path / "tune.json"))
    prev = int(_native.lib().fp3d_get_element_block())
    prev_w = int(_native.lib().fp3d_get_walk_width())
    try:
        r = tune_newton_workspace(ws, m.nodes[:, 0].copy(), 0.7, precond_matrix="picard",
                                  restarts=(30, 60), degrees=(1, 2), max_coarse=(10, 500),
                                  cache=cache)
        print(f"[tune] {r['element_block']['best']} {r['solver']['best']}", flush=True)
        assert r["cached"] is False
        assert r["element_block"]["best"] in (64, 128, 192, 256)
        assert int(_native.lib().fp3d_get_element_block()) == r["element_block"]["best"]
        assert r["walk_width"]["best"] in (1, 2, 3) and len(r["walk_width"]["ms"]) == 3
        assert int(_native.lib().fp3d_get_walk_width()) == r["walk_width"]["best"]
        assert len(r["solver"]["trials"]) == 8 and any(t["converged"] for t in r["solver"]["trials"])
        again = tune_newton_workspace(ws, m.nodes[:, 0].copy(), 0.7, cache=cache)
        assert again["cached"] is True and again["solver"]["best"] == r["solver"]["best"]
        kw = tuned_newton_kwargs(r)
        res = solve_newton(m.nodes, m.elements, dn, dv, 0.6, n_picard_seed=0,
                           phi_init=m.nodes[:, 0].copy(), **kw)
        base = solve_newton(m.nodes, m.elements, dn, dv, 0.6, n_picard_seed=0,
                            phi_init=m.nodes[:, 0].copy())
        assert res["converged"] and base["converged"]
        assert np.abs(res["phi"] - base["phi"]).max() < 1e-8
    finally:
        from up3d_amd.tune import apply_element_block, apply_walk_width

        apply_element_block(prev)
        apply_walk_width(prev_w)

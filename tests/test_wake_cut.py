"""Conforming wake cut and its constraint on the CPU
(pyfp3d/mesh/wake_cut.py, pyfp3d/constraints/wake.py), on the reference's
NACA0012 2.5D meshes (tests/golden fixtures with their wake and symmetry
face groups): topology of the cut (every wake face becomes two boundary
faces, stars split in two, one station per TE line, probes one wall edge
off the TE on both sides) and the exact algebra of the elimination."""

from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest
import scipy.sparse as sp

from up3d_amd.constraints.dirichlet import farfield_dirichlet
from up3d_amd.constraints.wake import WakeConstraint, kutta_targets
from up3d_amd.mesh.metrics import build_face_adjacency
from up3d_amd.mesh.wake_cut import cut_wake

GOLD = Path(__file__).parent / "golden"


def _mesh(level):
    z = np.load(GOLD / f"naca0012_{level}_bc.npz")
    return SimpleNamespace(nodes=z["nodes"], elements=z["elements"],
                           boundary_faces={k: z[k] for k in ("wall", "farfield", "wake", "symmetry")})


def _bfaces(el):
    fn = build_face_adjacency(el)
    fn = fn[0] if isinstance(fn, tuple) else fn
    return int((fn < 0).sum())


@pytest.mark.parametrize("level", ["coarse", "medium"])
def test_cut_topology(level):
    m = _mesh(level)
    mc, wc = cut_wake(m)
    nw = len(m.boundary_faces["wake"])
    assert _bfaces(mc.elements) == _bfaces(m.elements) + 2 * nw
    assert len(mc.nodes) == len(m.nodes) + len(wc.slave_nodes)
    assert np.array_equal(mc.nodes[wc.slave_nodes], m.nodes[wc.master_nodes])
    assert len(wc.free_nodes) == 0 and wc.n_stations == 1
    # one TE node per z-plane, at (1, 0)
    assert len(wc.te_nodes) == len(np.unique(m.nodes[:, 2]))
    assert np.allclose(m.nodes[wc.te_nodes, :2], [1.0, 0.0])
    # probes: wall nodes off the sheet, one above and one below each TE node
    wall = np.unique(m.boundary_faces["wall"])
    assert np.isin(wc.kutta_upper, wall).all() and np.isin(wc.kutta_lower, wall).all()
    assert (m.nodes[wc.kutta_upper, 1] > 0).all() and (m.nodes[wc.kutta_lower, 1] < 0).all()
    # "+" side elements read slaves, "-" side elements the masters
    el = mc.elements.astype(np.int64)
    cen_y = mc.nodes[el, 1].mean(axis=1)
    uses_slave = np.isin(el, wc.slave_nodes).any(axis=1)
    uses_master = np.isin(el, wc.master_nodes).any(axis=1)
    assert (cen_y[uses_slave] > 0).all() and (cen_y[uses_master & ~uses_slave] < 0).all()


def test_constraint_algebra():
    m = _mesh("coarse")
    mc, wc = cut_wake(m)
    n = len(mc.nodes)
    rng = np.random.default_rng(0)
    R = sp.random(n, n, density=5.0 / n, random_state=2, format="csr")
    A = (R + R.T + sp.eye(n) * 10.0).tocsr()
    con = WakeConstraint(A, wc)
    phi_red = rng.standard_normal(wc.n_nodes_orig)
    g = np.array([0.37])
    full = con.expand(phi_red, g)
    assert np.allclose(full[wc.slave_nodes] - full[wc.master_nodes], 0.37)
    lhs = con.T.T @ (A @ full)
    assert np.allclose(lhs, con.A_reduced @ phi_red + g @ con._h, atol=1e-11)
    b = rng.standard_normal(n)
    assert np.allclose(con.reduced_rhs(b, g), con.T.T @ b - g @ con._h)
    # Kutta target of a field whose only jump is the sheet's: the probe jump
    assert np.allclose(kutta_targets(full, wc),
                       np.mean(full[wc.kutta_upper] - full[wc.kutta_lower]))
    nd, vd = con.to_reduced_dirichlet(np.array([0, 1, wc.slave_nodes[0]]), np.array([1., 2., 3.]))
    assert list(nd) == [0, 1] and list(vd) == [1.0, 2.0]


def test_farfield_jump_across_the_cut():
    m = _mesh("coarse")
    mc, wc = cut_wake(m)
    nd, vd = farfield_dirichlet(mc, wc, 2.0, np.array([0.2]))
    val = dict(zip(nd.tolist(), vd.tolist()))
    pairs = [(s, ms) for s, ms in zip(wc.slave_nodes, wc.master_nodes)
             if s in val and ms in val]
    assert pairs
    for s, ms in pairs:
        assert abs((val[s] - val[ms]) - 0.2) < 1e-12


def test_node_min_edge_length_equals_numpy_formula():
    """The library's shortest incident edge per node equals the numpy
    restatement (edge norms, minimum.at) bit for bit; a node in no tet
    keeps +inf."""
    import numpy as np

    from up3d_amd.mesh.channel import bump_channel
    from up3d_amd.wake.cut_elements import _EDGES, node_min_edge_length

    m = bump_channel(10, 4, 4, jitter=0.2, seed=7, rcm=True)
    nodes = np.vstack([m.nodes, [[9.0, 9.0, 9.0]]])          # an unused node
    el = np.asarray(m.elements, dtype=np.int64)
    i, j = el[:, _EDGES[:, 0]].ravel(), el[:, _EDGES[:, 1]].ravel()
    L = np.linalg.norm(nodes[i] - nodes[j], axis=1)
    ref = np.full(len(nodes), np.inf)
    np.minimum.at(ref, i, L)
    np.minimum.at(ref, j, L)
    got = node_min_edge_length(nodes, m.elements)
    assert np.array_equal(got, ref)
    assert np.isinf(got[-1])

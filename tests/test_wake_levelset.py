"""Level-set wake set-up on the CPU: the WakeLevelSet frame and the
CutElementMap census (pyfp3d/wake/levelset.py, pyfp3d/wake/cut_elements.py),
pinned by the reference's synthetic known answers
(tests/test_b1_cut_elements.py:67-131) and by invariants on the reference's
NACA0012 2.5D meshes (committed as tests/golden fixtures)."""

from pathlib import Path

import numpy as np
import pytest

from up3d_amd.wake.cut_elements import CutElementMap
from up3d_amd.wake.levelset import WakeLevelSet

GOLD = Path(__file__).parent / "golden"


def _strip(x0):
    nodes = np.array([[1.0 + x0, -0.5, 0.0], [2.0 + x0, -0.5, 0.0],
                      [1.0 + x0, 0.5, 0.0], [1.0 + x0, 0.0, 1.0],
                      [2.0 + x0, 0.5, 1.0]])
    return nodes, np.array([[0, 1, 2, 3], [1, 2, 3, 4]])


def test_downstream_strip_is_cut():
    nodes, el = _strip(0.5)
    cm = CutElementMap(nodes, el, WakeLevelSet([1.0, 0.0, 0.0], (1.0, 0.0, 0.0)))
    assert set(cm.cut_elems) == {0, 1}
    assert cm.n_ext_dofs == 5
    assert cm.node_side[0] == -1 and cm.node_side[2] == 1 and cm.node_side[3] == 1


def test_upstream_strip_not_cut():
    nodes, el = _strip(-5.0)
    cm = CutElementMap(nodes, el, WakeLevelSet([1.0, 0.0, 0.0], (1.0, 0.0, 0.0)))
    assert len(cm.cut_elems) == 0 and cm.n_ext_dofs == 0


def test_te_fan_split():
    nodes = np.array([[1.0, 0.0, 0.0], [0.5, -0.3, 0.0], [0.9, -0.4, 1.0],
                      [0.4, -0.2, 0.5]])
    cm = CutElementMap(nodes, np.array([[0, 1, 2, 3]]),
                       WakeLevelSet([1.0, 0.0, 0.0], (1.0, 0.0, 0.0)))
    assert len(cm.cut_elems) == 0
    assert list(cm.te_lower_elems) == [0] and list(cm.te_nodes) == [0]


def test_dof_assignment_convention():
    nodes, el = _strip(0.5)
    cm = CutElementMap(nodes, el, WakeLevelSet([1.0, 0.0, 0.0], (1.0, 0.0, 0.0)))
    for k, e in enumerate(cm.cut_elems):
        for loc, n in enumerate(el[e]):
            aux = cm.ext_dof_of_node[n]
            if cm.node_side[n] == 1:
                assert cm.dofs_upper[k, loc] == n and cm.dofs_lower[k, loc] == aux
            else:
                assert cm.dofs_upper[k, loc] == aux and cm.dofs_lower[k, loc] == n


def test_levelset_frame_closed_forms():
    """Oblique frame on a swept TE: x = a + u v + d g + s n exactly; q is
    the unclamped arclength, s the signed offset along the panel normal."""
    te = np.array([[1.0, 0.0, 0.0], [1.5, 0.0, 2.0]])       # swept segment
    g = np.array([np.cos(0.1), np.sin(0.1), 0.0])
    wls = WakeLevelSet(te, direction=g)
    v = te[1] - te[0]
    n = np.cross(v, g)
    n /= np.linalg.norm(n)
    rng = np.random.default_rng(4)
    u, d, s = rng.uniform(0.05, 0.95, 20), rng.uniform(0.1, 3.0, 20), rng.uniform(-0.3, 0.3, 20)
    x = te[0] + u[:, None] * v + d[:, None] * g + s[:, None] * n
    s2, d2, q2 = wls.evaluate(x)
    assert np.allclose(s2, s, atol=1e-13) and np.allclose(d2, d, atol=1e-13)
    assert np.allclose(q2, u * np.linalg.norm(v), atol=1e-13)
    assert np.allclose(wls.surface_normals(x), n)
    # off the tip: q reports the excess, not the clamp
    _, _, q3 = wls.evaluate(te[1][None, :] + 0.5 * v[None, :] / np.linalg.norm(v) + g)
    assert q3[0] > wls.span_length


@pytest.mark.parametrize("level", ["coarse", "medium"])
def test_naca_census(level):
    z = np.load(GOLD / f"naca0012_{level}_bc.npz")
    nodes, el = z["nodes"], z["elements"]
    wall = np.unique(z["wall"])
    zz = nodes[:, 2]
    wls = WakeLevelSet(np.array([[1.0, 0.0, zz.min()], [1.0, 0.0, zz.max()]]),
                       direction=(1.0, 0.0, 0.0))
    cm = CutElementMap(nodes, el, wls, wall_nodes=wall)
    # one TE node per z-layer of the quasi-2D mesh
    assert len(cm.te_nodes) == len(np.unique(np.round(zz, 12)))
    assert np.allclose(nodes[cm.te_nodes, :2], [1.0, 0.0], atol=1e-9)
    assert np.all(cm.node_side != 0)
    y = nodes[:, 1]
    off = np.abs(y) > 1e-9
    assert np.all(cm.node_side[off] == np.sign(y[off]).astype(np.int8))
    assert np.all(cm.node_side[~off] == 1)           # on-sheet nodes go "+"
    # aux numbering is contiguous after the main DOFs
    ext = cm.ext_dof_of_node[cm.ext_dof_of_node >= 0]
    assert ext.min() == cm.n_main and ext.max() == cm.n_total_dofs - 1
    assert len(np.unique(ext)) == cm.n_ext_dofs
    # the below-TE fan is exactly the TE elements whose other nodes are "-"
    is_te = np.zeros(len(nodes), dtype=bool)
    is_te[cm.te_nodes] = True
    te_e = is_te[el]
    fan = np.flatnonzero(te_e.any(1) & np.all(te_e | (cm.node_side[el] == -1), 1))
    assert set(fan) == set(cm.te_lower_elems)
    assert not np.isin(fan, cm.cut_elems).any()
    # every cut element straddles the sheet downstream of the TE
    ce = el[cm.cut_elems]
    assert np.all(cm.node_side[ce].min(1) == -1) and np.all(cm.node_side[ce].max(1) == 1)
    assert np.all(nodes[ce, 0].max(1) > 1.0)

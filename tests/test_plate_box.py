"""Embedded swept-plate box (the lifting level-set case): mesh invariants
and the wake census on the CPU.  No reference counterpart (the reference's
lifting meshes are body-fitted); pinned by closed forms: the box volume,
the plate's face count, outward normals, the crack (no face adjacency
through the plate), and the level-set census rules of
pyfp3d/wake/cut_elements.py applied to a planar sheet on grid faces."""

import numpy as np
import pytest

from up3d_amd.mesh.metrics import build_face_adjacency, precompute_element_geometry
from up3d_amd.mesh.plate_box import plate_wing_box
from up3d_amd.wake.cut_elements import CutElementMap
from up3d_amd.wake.levelset import WakeLevelSet


@pytest.fixture(scope="module")
def box():
    return plate_wing_box(32, 24, 24)


def _adjacency(el):
    fn = build_face_adjacency(el)
    return fn[0] if isinstance(fn, tuple) else fn


def test_volume_and_boundary_tags(box):
    geo = precompute_element_geometry(box.nodes, box.elements)
    V = geo[1] if isinstance(geo, tuple) else geo["V"]
    assert V.min() > 0.0
    assert abs(V.sum() - 10.0 * 4.0 * 8.0) < 1e-9       # xi 10 x y 4 x z 8 (shear keeps volume)
    fn = _adjacency(box.elements)
    tagged = sum(len(box.boundary_faces[t]) for t in
                 ("inflow", "outflow", "side", "bottom", "top", "symmetry", "wall"))
    assert tagged == int((fn < 0).sum())
    assert len(box.boundary_faces["farfield"]) == sum(
        len(box.boundary_faces[t]) for t in ("inflow", "side", "bottom", "top"))


def test_plate_is_a_two_sided_wall(box):
    m = box.meta
    nc, ns = m["i_te"] - m["i_le"], m["j_tip"]
    w = box.boundary_faces["wall"]
    # two triangles per plate cell and side, less the one leading-edge /
    # tip corner triangle whose nodes all lie on the planform's edges
    assert len(w) == 2 * 2 * nc * ns - 2
    p = box.nodes[w]
    nrm = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    lower = np.isin(w, box.lower_nodes).any(axis=1)
    assert (nrm[lower, 2] > 0).all() and (nrm[~lower, 2] < 0).all()
    assert lower.sum() == (~lower).sum()
    # the duplicated nodes sit where their upper twins are
    assert np.array_equal(box.nodes[box.lower_nodes], box.nodes[box.upper_nodes])
    # tets below the plate never read an upper twin, tets above never a lower copy
    el = box.elements.astype(np.int64)
    zc = box.nodes[el, 2].mean(axis=1)
    assert not np.isin(el[zc < 0], box.upper_nodes).any()
    assert not np.isin(el[zc > 0], box.lower_nodes).any()


def test_trailing_edge_line(box):
    te = box.nodes[box.te_nodes]
    t = np.tan(np.deg2rad(30.0))
    assert np.allclose(te[:, 0], 1.0 + t * te[:, 1]) and np.allclose(te[:, 2], 0.0)
    assert np.isclose(te[:, 1].min(), 0.0) and np.isclose(te[:, 1].max(), 1.0)


def test_wake_census_on_the_plate(box):
    b = box.meta["half_span"]
    wls = WakeLevelSet(box.te_line, direction=(1.0, 0.0, 0.0))
    assert np.allclose(wls.surface_normals(box.nodes[:1]), [0.0, 0.0, 1.0])
    cm = CutElementMap(box.nodes, box.elements, wls, wall_nodes=box.wall_nodes,
                       minus_nodes=box.lower_nodes,
                       inboard_clip=lambda p: p[..., 1] < b * (1.0 - 1e-9),
                       outboard_fringe=1e-9 * wls.span_length)
    assert set(cm.te_nodes) == set(box.te_nodes)
    assert (cm.node_side[box.lower_nodes] == -1).all()
    el = box.elements.astype(np.int64)
    ce = el[cm.cut_elems]
    # cut tets straddle the sheet downstream of the TE, inboard of the tip
    assert (cm.node_side[ce].min(1) == -1).all() and (cm.node_side[ce].max(1) == 1).all()
    xte = 1.0 + np.tan(np.deg2rad(30.0)) * box.nodes[ce, 1]
    assert (box.nodes[ce, 0].max(1) > xte.min(1) - 1e-12).all()
    aux_nodes = np.flatnonzero(cm.ext_dof_of_node >= 0)
    assert box.nodes[aux_nodes, 1].max() <= b + 1e-12
    # every aux node lies on a streamwise grid line that starts at a TE node
    assert np.isin(np.round(box.nodes[aux_nodes, 1], 9),
                   np.round(box.nodes[box.te_nodes, 1], 9)).all()
    # the fan below the TE reads only lower-face copies and lower nodes
    fan = el[cm.te_lower_elems]
    is_te = np.isin(fan, cm.te_nodes)
    assert (is_te | (cm.node_side[fan] == -1)).all()

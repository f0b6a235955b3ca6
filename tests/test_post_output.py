"""VTU output and spanwise sections (pyfp3d/post/vtk_out.py,
pyfp3d/post/section_cut.py), pinned by exact round trips and closed forms:
binary float64 fields survive write/read bit for bit; a plane section of
an axis-aligned box has the box's cross-section area and reproduces linear
nodal fields exactly (P1 interpolation along cut edges); the symmetry-plane
path returns that plane's triangulation; on a flat plate in uniform flow
every sectional Cp equals 1 - cos^2(alpha) and the section spans one chord."""

import numpy as np
import pytest

from up3d_amd.mesh.channel import bump_channel
from up3d_amd.mesh.plate_box import plate_wing_box
from up3d_amd.post.section_cut import section_cp_curve, section_cut
from up3d_amd.post.vtk_out import read_vtu, write_vtu


def _area(sec):
    p = sec.points2d[sec.triangles]
    u, v = p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]
    return 0.5 * np.abs(u[:, 0] * v[:, 1] - u[:, 1] * v[:, 0]).sum()


def test_vtu_round_trip(tmp_path):
    m = bump_channel(6, 3, 3, jitter=0.1, seed=1)
    rng = np.random.default_rng(0)
    pd = {"phi": rng.standard_normal(m.n_nodes), "u": rng.standard_normal((m.n_nodes, 3))}
    cd = {"rho": rng.standard_normal(m.n_tets)}
    f = tmp_path / "out" / "sol.vtu"
    write_vtu(f, m.nodes, m.elements, pd, cd)
    nodes, el, fields = read_vtu(f)
    assert np.array_equal(nodes, m.nodes) and np.array_equal(el, m.elements)
    assert np.array_equal(fields["phi"], pd["phi"]) and np.array_equal(fields["u"], pd["u"])
    assert np.array_equal(fields["cell_rho"], cd["rho"])
    with pytest.raises(ValueError):
        write_vtu(f, m.nodes, m.elements, {"bad": np.zeros(3)})


@pytest.mark.parametrize("z", [0.37, 0.5, 0.0625])
def test_marching_section_of_a_box(z):
    m = bump_channel(12, 4, 4, bump_height=0.0)            # [0,3]x[0,1]x[0,1]
    lin = 2.0 * m.nodes[:, 0] - m.nodes[:, 1] + 3.0 * m.nodes[:, 2] + 1.0
    sec = section_cut(m, {"f": lin}, z)
    assert abs(_area(sec) - 3.0) < 1e-12
    x, y = sec.points2d[:, 0], sec.points2d[:, 1]
    assert np.allclose(sec.fields["f"], 2.0 * x - y + 3.0 * z + 1.0, atol=1e-12)


def test_symmetry_plane_section():
    box = plate_wing_box(16, 12, 12)
    sec = section_cut(box, {"y": box.nodes[:, 1]}, 0.0, span_axis=1)
    assert abs(_area(sec) - 10.0 * 8.0) < 1e-9              # x in [-4, 6], z in [-4, 4]
    assert np.all(sec.fields["y"] == 0.0)


def test_plate_section_cp_uniform_flow():
    box = plate_wing_box(32, 24, 24)
    a = np.deg2rad(3.0)
    phi = box.nodes[:, 0] * np.cos(a) + box.nodes[:, 2] * np.sin(a)
    c = section_cp_curve(box, phi, eta=0.5, b_semi=1.0, upper_hint=(0.0, 0.0, 1.0),
                         span_axis=1, side_by="normal")
    for k in ("cp_upper", "cp_lower"):
        assert np.allclose(c[k], 1.0 - np.cos(a) ** 2, atol=1e-12)
    assert len(c["x_upper"]) == len(c["x_lower"]) >= 5
    h = 1.0 / (box.meta["i_te"] - box.meta["i_le"])        # chordwise cell
    assert 1.0 - 2 * h < c["chord"] <= 1.0
    assert abs(c["x_le"] - 0.5 * np.tan(np.deg2rad(30.0))) < h
    with pytest.raises(ValueError):
        section_cp_curve(box, phi, eta=0.5, upper_hint=(0.0, 0.0, 1.0), span_axis=1)


def _normals_full_search(nodes, elements, wall_faces):
    """Wall normals with the owner found among all 4 N_t faces."""
    el = np.asarray(elements, dtype=np.int64)
    wf = np.asarray(wall_faces, dtype=np.int64)
    n = int(max(el.max(), wf.max())) + 1

    def keys(tri):
        s = np.sort(tri, axis=1)
        return (s[:, 0] * n + s[:, 1]) * n + s[:, 2]

    faces = np.concatenate([el[:, [1, 2, 3]], el[:, [0, 2, 3]], el[:, [0, 1, 3]],
                            el[:, [0, 1, 2]]])
    owner = np.tile(np.arange(len(el)), 4)
    order = np.argsort(keys(faces), kind="stable")
    ks, os_ = keys(faces)[order], owner[order]
    own = os_[np.searchsorted(ks, keys(wf))]
    p0 = nodes[wf[:, 0]]
    nv = np.cross(nodes[wf[:, 1]] - p0, nodes[wf[:, 2]] - p0)
    nrm = nv / np.linalg.norm(nv, axis=1)[:, None]
    flip = np.einsum("ij,ij->i", nrm, nodes[el[own]].mean(axis=1) - nodes[wf].mean(axis=1)) < 0
    nrm[flip] *= -1.0
    return nrm


def test_wall_normals_candidate_search_equals_full_search():
    """The owner search over tets with three wall nodes gives the normals of
    the search over every face, bit for bit, on each channel wall; a face
    that is not a boundary face is refused."""
    from up3d_amd.post.surface import wall_outward_normals

    m = bump_channel(12, 4, 4, jitter=0.2, seed=3, rcm=True)
    for name, wf in m.boundary_faces.items():
        got = wall_outward_normals(m.nodes, m.elements, wf)
        assert np.array_equal(got, _normals_full_search(m.nodes, m.elements, wf)), name
    interior = m.elements[:1, :3]            # a face of the first tet
    faces = {tuple(sorted(f)) for f in np.concatenate(list(m.boundary_faces.values()))}
    if tuple(sorted(interior[0])) not in faces:
        with pytest.raises(ValueError, match="not owned"):
            wall_outward_normals(m.nodes, m.elements, interior)


def test_array_digest_and_cached_locations():
    """The mesh content key separates different contents and matches equal
    ones; repeated P1 transfers reuse the located points and give the same
    values."""
    from up3d_amd.kernels.store import array_digest
    from up3d_amd.mesh.transfer import interpolate_p1, locate_points_cached

    a = np.arange(12.0)
    assert array_digest(a) == array_digest(a.copy())
    b = a.copy()
    b[3] += 1e-12
    assert array_digest(a) != array_digest(b)
    assert array_digest(a.reshape(3, 4)) == array_digest(a)          # bytes only
    src = bump_channel(6, 2, 2, jitter=0.1, seed=1, rcm=True)
    dst = bump_channel(12, 4, 4, jitter=0.1, seed=2, rcm=True)
    f = 1.0 + 2.0 * src.nodes[:, 0] - src.nodes[:, 1] + 0.5 * src.nodes[:, 2]
    v1 = interpolate_p1(src.nodes, src.elements, f, dst.nodes)
    h1 = locate_points_cached(src.nodes, src.elements, dst.nodes)
    v2 = interpolate_p1(src.nodes, src.elements, 2.0 * f, dst.nodes)
    h2 = locate_points_cached(src.nodes, src.elements, dst.nodes)
    assert h1[0] is h2[0] and not h1[0].flags.writeable
    assert np.array_equal(v2, 2.0 * v1)

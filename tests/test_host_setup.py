"""Host-side set-up of the product (native C++ in libfp3d_hip.so, runs on
the CPU) against the oracle, the reference's fixtures and closed forms,
plus the C-ABI export check.  No GPU needed."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from up3d_amd import _native
from up3d_amd.kernels.store import csr_pattern, elem_to_csr
from up3d_amd.mesh import channel
from up3d_amd.mesh.coloring import color_partition_csr, greedy_coloring, validate_coloring
from up3d_amd.mesh.metrics import build_face_adjacency, precompute_element_geometry

from .helpers import jittered_cube, structured_cube_mesh

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_every_declared_symbol():
    hdr = (ROOT / "include" / "fp3d.h").read_text()
    decl = set(re.findall(r"\b(fp3d_[a-z0-9_]+)\s*\(", hdr))
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    missing = [s for s in sorted(decl) if not hasattr(lib, s)]
    assert not missing, missing
    assert set(_native.EXPORTED) <= decl


def test_library_loads_without_gpu():
    assert _native.lib().fp3d_version() == 1
    assert _native.device_count() >= 0


@pytest.mark.parametrize("mesh", ["kuhn8", "jcube", "channel"])
def test_coloring_bit_identical_to_oracle(mesh):
    if mesh == "kuhn8":
        _, el = structured_cube_mesh(8)
    elif mesh == "jcube":
        _, el = jittered_cube(6, seed=2)
    else:
        el = channel.bump_channel(12, 5, 4, jitter=0.15, seed=4, rcm=True).elements
    c1, n1 = greedy_coloring(el)
    c2, n2 = O.greedy_coloring(el)
    assert n1 == n2 and np.array_equal(c1, c2)
    o1 = color_partition_csr(el)
    o2 = O.color_partition_csr(el)
    assert np.array_equal(o1[0], o2[0]) and np.array_equal(o1[1], o2[1])
    assert validate_coloring(el, c1)


def test_coloring_golden_kuhn8_product():
    _, el = structured_cube_mesh(8)
    c, n = greedy_coloring(el)
    cnt = np.bincount(c)
    assert (n, cnt.min(), cnt.max()) == (31, 3, 153)


def test_coloring_limit_raises():
    # a node shared by > 255 elements forces > 256 colours
    k = 300
    el = np.stack([np.zeros(k), 1 + 3 * np.arange(k), 2 + 3 * np.arange(k),
                   3 + 3 * np.arange(k)], axis=1).astype(np.int32)
    with pytest.raises(ValueError, match="256"):
        greedy_coloring(el)


def test_csr_pattern_and_elem_to_csr_bit_identical():
    nodes, el = jittered_cube(5, seed=1)
    ip, ix = csr_pattern(el, len(nodes))
    pat = O.build_csr_pattern(el, len(nodes))
    assert np.array_equal(ip, pat.indptr) and np.array_equal(ix, pat.indices)
    assert np.array_equal(elem_to_csr(el, ip, ix),
                          O.build_elem_to_csr(el, pat.indptr, pat.indices))
    # a channel large enough for every host thread (row- and element-
    # parallel passes), and a pattern missing one element's couplings
    m = channel.bump_channel(24, 8, 8, jitter=0.1, seed=2, rcm=True)
    ip2, ix2 = csr_pattern(m.elements, m.n_nodes)
    pat2 = O.build_csr_pattern(m.elements, m.n_nodes)
    assert np.array_equal(ip2, pat2.indptr) and np.array_equal(ix2, pat2.indices)
    assert np.array_equal(elem_to_csr(m.elements, ip2, ix2),
                          O.build_elem_to_csr(m.elements, pat2.indptr, pat2.indices))
    ip3, ix3 = csr_pattern(m.elements[1:], m.n_nodes)
    with pytest.raises(ValueError, match="missing"):
        elem_to_csr(m.elements, ip3, ix3)


def test_face_adjacency_identical_to_oracle():
    _, el = jittered_cube(5, seed=1)
    fn, _ = build_face_adjacency(el)
    assert np.array_equal(fn, O.build_face_adjacency(el))


def test_geometry_matches_oracle():
    nodes, el = jittered_cube(5, seed=1)
    B, V = precompute_element_geometry(nodes, el)
    Bo, Vo = O.precompute_element_geometry(nodes, el)
    assert np.abs(B - Bo).max() < 1e-13 * np.abs(Bo).max()
    assert np.abs(V - Vo).max() < 1e-15 * np.abs(Vo).max()
    bad = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], dtype=float)
    with pytest.raises(ValueError, match="degenerate"):
        precompute_element_geometry(bad, np.array([[0, 1, 2, 3]]))


@pytest.mark.parametrize("shape", [(48, 16, 16), (6, 4, 3)])
def test_channel_sizes_closed_form(shape):
    nx, ny, nz = shape
    if nx == 48:
        ijk, el = channel.kuhn_box(nx, ny, nz)
        assert len(el) == 73728 and len(ijk) == 14161
        return
    m = channel.bump_channel(nx, ny, nz, jitter=0.15, seed=3, rcm=True)
    assert m.n_tets == 6 * nx * ny * nz
    assert m.n_nodes == (nx + 1) * (ny + 1) * (nz + 1)
    ip, ix = csr_pattern(m.elements, m.n_nodes)
    N = m.n_nodes
    s = sum((nx + 1 - a) * (ny + 1 - b) * (nz + 1 - c)
            for a, b, c in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (0, 1, 1),
                            (1, 0, 1), (1, 1, 1)])
    assert len(ix) == N + 2 * s
    _, V = precompute_element_geometry(m.nodes, m.elements)
    assert V.min() > 0.0
    # flat boundaries: inflow/outflow areas are 1, side walls area 3 - bump
    for tag, area in (("inflow", 1.0), ("outflow", 1.0), ("ceiling", 3.0)):
        f = m.boundary_faces[tag]
        p = m.nodes[f]
        a = 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1)
        assert a.sum() == pytest.approx(area, rel=1e-12)
    b = channel.neumann_flux_rhs(m)
    assert abs(b.sum()) < 1e-14 and b.min() < 0 < b.max()


@pytest.mark.parametrize("alpha", [0.0, 25.0])
def test_neumann_flux_rhs_is_the_boundary_term_of_a_linear_field(alpha):
    """For phi = u . x (rho = 1, M 0) the Galerkin residual is
    R_i = int grad phi . grad N_i = oint N_i (u . n) dS (divergence theorem),
    the consistent P1 load of the flux through every boundary face: with the
    loads of all six faces of a flat channel it equals R to rounding, at
    every node (interior nodes jittered).  Per face the load sums to
    rho u . n times the area: -cos(alpha) at the inflow, +cos(alpha) at the
    outflow (area 1), -3 sin(alpha) at the floor (area 3)."""
    m = channel.bump_channel(6, 4, 3, bump_height=0.0, jitter=0.15, seed=9)
    a = np.deg2rad(alpha)
    phi = m.nodes[:, 0] * np.cos(a) + m.nodes[:, 1] * np.sin(a)
    tags = ("inflow", "outflow", "floor", "ceiling", "side_lo", "side_hi")
    b = channel.neumann_flux_rhs(m, tags=tags, alpha_deg=alpha)
    R = O.OraclePicardOperator(m.nodes, m.elements).assemble_residual(
        phi, np.ones(m.n_tets))
    assert np.abs(R - b).max() < 1e-13 * np.abs(b).max()      # rounding only
    for tag, total in (("inflow", -np.cos(a)), ("outflow", np.cos(a)),
                       ("floor", -3.0 * np.sin(a)), ("ceiling", 3.0 * np.sin(a)),
                       ("side_lo", 0.0)):
        bt = channel.neumann_flux_rhs(m, tags=(tag,), alpha_deg=alpha, rho_inf=1.0)
        assert bt.sum() == pytest.approx(total, abs=1e-14)
    # the default (inflow + outflow) is what the subsonic channel carries;
    # rho_inf scales it
    b2 = channel.neumann_flux_rhs(m, rho_inf=2.0)
    assert np.allclose(b2, 2.0 * channel.neumann_flux_rhs(m), rtol=0, atol=1e-16)


def test_bump_profile():
    x = np.array([0.5, 1.0, 1.5, 2.0, 2.5])
    y = channel.bump_profile(x, 0.1)
    assert y[0] == 0 and y[-1] == 0 and abs(y[1]) < 1e-15 and abs(y[3]) < 1e-15
    assert y[2] == pytest.approx(0.1, rel=1e-14)


def test_mach_schedule_and_ew():
    from up3d_amd.solve.continuation import mach_schedule
    from up3d_amd.solve.newton import _ew_forcing

    assert mach_schedule(0.84, 0.7, 0.05) == pytest.approx([0.7, 0.75, 0.8, 0.84])
    assert mach_schedule(0.6) == [0.6]
    assert _ew_forcing(1.0, None, None) == 1e-2
    assert _ew_forcing(1e-3, 1e-1, 1e-2) == pytest.approx(0.9e-4)
    assert _ew_forcing(1e-12, 1.0, 1e-2) == 1e-10


def test_p1_transfer_exact_for_linear_fields():
    """Mesh-to-mesh P1 transfer (grid sequencing): linear fields are
    reproduced to rounding on a finer jittered mesh of the same domain,
    coarse-node values are returned exactly at coinciding points, and the
    located barycentric coordinates sum to one."""
    from up3d_amd.mesh.channel import bump_channel
    from up3d_amd.mesh.transfer import interpolate_p1, locate_points

    coarse = bump_channel(12, 4, 4, jitter=0.15, seed=1, rcm=True)
    fine = bump_channel(24, 8, 8, jitter=0.15, seed=2, rcm=True)
    f = lambda p: 1.3 * p[:, 0] - 0.7 * p[:, 1] + 0.2 * p[:, 2] + 0.05  # noqa: E731
    v = interpolate_p1(coarse.nodes, coarse.elements, f(coarse.nodes), fine.nodes)
    assert np.abs(v - f(fine.nodes)).max() < 1e-13
    g = np.sin(coarse.nodes[:, 0]) * np.cos(2 * coarse.nodes[:, 1])
    assert np.abs(interpolate_p1(coarse.nodes, coarse.elements, g, coarse.nodes)
                  - g).max() < 1e-13
    elem, bary = locate_points(coarse.nodes, coarse.elements, fine.nodes)
    assert np.abs(bary.sum(axis=1) - 1.0).max() < 1e-13
    # interior fine nodes lie inside their located coarse tet
    inside = bary.min(axis=1) >= -1e-10
    assert inside.mean() > 0.95


def test_apply_givens_matches_rotation_product():
    """fp3d_host_apply_givens: the chain of rotations equals the product of
    the 2x2 rotation matrices (an orthogonal map: the column norm is kept),
    and gives the bits of the host loop it replaced in gmres_steps."""
    rng = np.random.default_rng(3)
    for col in (0, 1, 7, 60):
        hrow = rng.standard_normal(col + 2)
        ang = rng.uniform(0, 2 * np.pi, col)
        givens = np.ascontiguousarray(np.stack([np.cos(ang), np.sin(ang)], axis=1))
        out = hrow.copy()
        _native.call("fp3d_host_apply_givens", col, out.ctypes.data, givens.ctypes.data)
        Q = np.eye(col + 2)
        for j in range(col):
            c, s = givens[j]
            G = np.eye(col + 2)
            G[j:j + 2, j:j + 2] = [[c, s], [-s, c]]
            Q = G @ Q
        assert np.allclose(out, Q @ hrow, rtol=0, atol=1e-13)
        assert abs(np.linalg.norm(out) - np.linalg.norm(hrow)) < 1e-13
        ref = hrow.copy()
        for j in range(col):
            c, s = givens[j]
            n0, n1 = ref[j], ref[j + 1]
            ref[j], ref[j + 1] = c * n0 + s * n1, -s * n0 + c * n1
        assert np.array_equal(out, ref)


def test_masked_sum_csr_is_the_masked_dense_sum():
    """The adjoint's merged operator: rows of J12 + T with the Dirichlet rows
    emptied, duplicate (row, column) pairs of the two kept as two entries,
    equals diag(fm) (J12 + T) densely, and its row order is ascending."""
    from types import SimpleNamespace

    import scipy.sparse as sp
    import torch

    from up3d_amd.solve.adjoint import _csr_from_coo, _masked_sum_csr

    rng = np.random.default_rng(5)
    n = 23
    A = sp.random(n, n, density=0.3, random_state=1, format="csr") + sp.eye(n)
    A = A.tocsr()
    A.sort_indices()
    st = SimpleNamespace(n_nodes=n, indptr=torch.from_numpy(A.indptr.astype(np.int32)),
                         indices=torch.from_numpy(A.indices.astype(np.int32)))
    J12 = torch.from_numpy(A.data.astype(np.float64))
    m = 60                       # COO with repeated pairs, some shared with A
    r = rng.integers(0, n, m)
    c = rng.integers(0, n, m)
    r[:5], c[:5] = r[5:10], c[5:10]
    v = rng.standard_normal(m)
    T = _csr_from_coo(torch.from_numpy(r.astype(np.int32)), torch.from_numpy(c.astype(np.int32)),
                      torch.from_numpy(v), n)
    fm = torch.from_numpy((rng.random(n) > 0.3).astype(np.uint8))
    M = _masked_sum_csr(st, J12, T, fm)
    Tt = sp.coo_matrix((v, (r, c)), shape=(n, n)).toarray()
    want = np.diag(fm.numpy().astype(float)) @ (A.toarray() + Tt)
    ip, ix, dv = M.indptr.numpy(), M.indices.numpy(), M.data.numpy()
    got = sp.csr_matrix((dv, ix, ip), shape=(n, n)).toarray()
    assert np.allclose(got, want, rtol=0, atol=1e-14)
    assert M.nnz == int(np.diff(A.indptr)[fm.numpy() > 0].sum() + (fm.numpy()[r] > 0).sum())
    assert all(np.all(np.diff(ix[ip[i]:ip[i + 1]]) >= 0) for i in range(n))


def test_gmres_rejects_restart_beyond_the_basis_limit():
    """The device multi-dot orthogonalises against at most 256 vectors, so a
    longer restart is refused before any device work."""
    import torch

    from up3d_amd.solve.linear import MAX_RESTART, gmres_dev

    b = torch.ones(10, dtype=torch.float64)
    with pytest.raises(ValueError, match="restart"):
        gmres_dev(lambda x, y: y.copy_(x), b, restart=MAX_RESTART + 1)
    with pytest.raises(ValueError, match="restart"):
        gmres_dev(lambda x, y: y.copy_(x), b, restart=0)


def test_executor_host_tasks_on_workers_and_inline():
    """HostTasks of stepped solves: on worker threads (host_workers > 0) or
    inline (0, and drive_inline), the same results; a task's exception
    reaches the caller."""
    import threading

    from up3d_amd.solve.batch import BatchExecutor
    from up3d_amd.solve.newton import HostTask, drive_inline

    main = threading.get_ident()
    seen = []

    def co(i):
        a = yield HostTask(lambda: (seen.append(threading.get_ident()), i * i)[1])
        b = yield HostTask(lambda: a + 1)
        return b

    for workers in (0, 4):
        seen.clear()
        ex = BatchExecutor(host_workers=workers)
        assert ex.run({i: co(i) for i in range(8)}) == {i: i * i + 1 for i in range(8)}
        assert ex.stats()["host_tasks"] == 16
        assert all((t == main) == (workers == 0) for t in seen)
    assert drive_inline(co(3)) == 10

    def bad():
        yield HostTask(lambda: 1 / 0)

    with pytest.raises(ZeroDivisionError):
        BatchExecutor(host_workers=2).run({0: bad()})

"""Pins the CPU oracle (oracle/) against the reference's golden fixtures,
closed forms, brute force and finite differences.  CPU only."""

import re
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

from .helpers import (cube_boundary_mask, jittered_cube, structured_cube_mesh,
                      transonic_field)

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def kuhn8():
    return structured_cube_mesh(8)


@pytest.fixture(scope="module")
def jcube():
    return jittered_cube(6, seed=3)


# ------------------------------------------------------------ physics
def test_physics_closed_forms():
    m = 0.5
    assert O.density_isentropic(1.0, m) == 1.0
    assert O.density_isentropic(0.0, m) == pytest.approx(1.05 ** 2.5, rel=1e-15)
    assert O.mach_number_squared(1.0, m) == pytest.approx(m * m, rel=1e-15)
    qs = O.critical_speed_squared(m)
    assert O.mach_number_squared(qs, m) == pytest.approx(1.0, rel=1e-14)
    q2 = np.array([0.0, 0.3, 1.0, 1.7, 2.4])
    assert np.all(O.density_field(q2, 0.0) == 1.0)          # bitwise Laplace limit
    assert O.q2_at_mach(3.0, 0.0) == np.inf
    lim = O.limit_q2_field(np.linspace(0, 2, 11), 0.5, 3.0)
    assert np.array_equal(lim, np.linspace(0, 2, 11))
    assert O.limit_q2_field(np.array([50.0]), 0.8, 3.0)[0] == O.q2_at_mach(3.0, 0.8)
    assert O.mach_number_squared(O.q2_at_mach(3.0, 0.8), 0.8) == pytest.approx(9.0)


def test_physics_derivatives_fd():
    q2 = np.linspace(0.2, 2.0, 19)
    h = 1e-6
    for m in (0.5, 0.84):
        fd = (O.density_field(q2 + h, m) - O.density_field(q2 - h, m)) / (2 * h)
        assert np.allclose(O.density_derivative_field(q2, m), fd, rtol=1e-7)
        fd2 = (O.mach_squared_field(q2 + h, m) - O.mach_squared_field(q2 - h, m)) / (2 * h)
        assert np.allclose(O.mach_squared_derivative_field(q2, m), fd2, rtol=1e-7)


# ------------------------------------------------------------ geometry
def test_geometry_invariants(jcube):
    nodes, el = jcube
    B, V = O.precompute_element_geometry(nodes, el)
    assert np.abs(B.sum(axis=1)).max() < 1e-12 * np.abs(B).max()
    assert abs(V.sum() - 1.0) < 1e-13
    a = np.array([0.3, -1.2, 2.5])
    phi = nodes @ a + 0.7
    g = np.einsum("ni,nid->nd", phi[el], B)
    assert np.abs(g - a).max() < 1e-12


@pytest.mark.parametrize("n", [4, 16])
def test_kuhn_volumes_exact(n):
    # cases/demo/p0_infrastructure/results/g01_volumes.csv:4-5 (sum V = 1 exactly)
    nodes, el = structured_cube_mesh(n) if n == 4 else _fast_cube(n)
    _, V = O.precompute_element_geometry(nodes, el)
    assert abs(V.sum() - 1.0) < 1e-14


def _fast_cube(n):
    from up3d_amd.mesh.channel import kuhn_box

    ijk, el = kuhn_box(n, n, n)
    return ijk / float(n), el.astype(np.int32)


def test_degenerate_tet_raises():
    nodes = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], dtype=float)
    with pytest.raises(ValueError):
        O.precompute_element_geometry(nodes, np.array([[0, 1, 2, 3]]))


# ------------------------------------------------------------ colouring
def test_coloring_golden_kuhn8(kuhn8):
    # cases/demo/p0_infrastructure/results/g03_coloring.csv:2
    _, el = kuhn8
    colors, nc = O.greedy_coloring(el)
    counts = np.bincount(colors)
    assert (len(el), nc, counts.min(), counts.max()) == (3072, 31, 3, 153)
    assert O.validate_coloring(el, colors)


def test_coloring_brute_force_validity(jcube):
    _, el = jcube
    colors, _ = O.greedy_coloring(el)
    # brute force: any two elements sharing a node differ in colour
    n = len(el)
    sets = [set(map(int, r)) for r in el]
    for e in range(0, n, 7):
        for f in range(n):
            if f != e and sets[e] & sets[f]:
                assert colors[e] != colors[f]


def test_color_partition_order(kuhn8):
    _, el = kuhn8
    off, ce, nc = O.color_partition_csr(el)
    colors, _ = O.greedy_coloring(el)
    for c in range(nc):
        seg = ce[off[c]:off[c + 1]]
        assert np.all(np.diff(seg) > 0) and np.all(colors[seg] == c)


# ------------------------------------------------------------ CSR
def _nnz_closed_form(nx, ny, nz):
    N = (nx + 1) * (ny + 1) * (nz + 1)
    s = 0
    for d in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (0, 1, 1), (1, 0, 1), (1, 1, 1)]:
        s += (nx + 1 - d[0]) * (ny + 1 - d[1]) * (nz + 1 - d[2])
    return N + 2 * s


def test_csr_nnz_closed_form(kuhn8):
    nodes, el = kuhn8
    pat = O.build_csr_pattern(el, len(nodes))
    assert pat.nnz == _nnz_closed_form(8, 8, 8)
    e2c = O.build_elem_to_csr(el, pat.indptr, pat.indices)
    r = np.repeat(el, 4, axis=1).reshape(-1, 4, 4)
    c = np.tile(el, (1, 4)).reshape(-1, 4, 4)
    rows = np.repeat(np.arange(len(nodes)), np.diff(pat.indptr))
    assert np.array_equal(rows[e2c], r) and np.array_equal(pat.indices[e2c], c)


# ------------------------------------------------------------ residual / A
def test_freestream_residual_machine_zero(jcube):
    nodes, el = jcube
    op = O.OraclePicardOperator(nodes, el)
    phi = nodes[:, 0].copy()
    for m in (0.0, 0.5, 0.8):
        g, q2 = op.velocities(phi)
        rho = O.density_field(q2, m)
        upw = O.OracleUpwind(nodes, el)
        rt = upw.rho_tilde(g, q2, rho, m, 1.5, 0.95)
        R = op.assemble_residual(phi, rt)
        assert np.abs(R[~cube_boundary_mask(nodes)]).max() < 1e-12


@pytest.mark.parametrize("order", ["serial", "colored"])
def test_colored_vs_serial_reference(jcube, order):
    # tests/test_p3_assembly.py:42-69: fast assembly vs the serial
    # reference (geometry recomputed, scipy COO->CSR) to 1e-13
    nodes, el = jcube
    op = O.OraclePicardOperator(nodes, el, order=order)
    A = op.assemble_matrix()
    Aref = O.stiffness_matrix_reference(nodes, el)
    assert np.array_equal(A.indptr, Aref.indptr) and np.array_equal(A.indices, Aref.indices)
    assert np.abs(A.data - Aref.data).max() < 1e-13 * np.abs(Aref.data).max()
    phi = np.sin(nodes[:, 0]) + 0.3 * nodes[:, 1] ** 2 - 0.1 * nodes[:, 2]
    R = op.assemble_residual(phi)
    Rr = O.residual_reference(nodes, el, phi)
    assert np.abs(R - Rr).max() < 1e-13 * np.abs(Rr).max()
    # A phi == R for the Laplace residual (linearity)
    assert np.abs(A @ phi - R).max() < 1e-12 * np.abs(R).max()


def test_serial_order_is_element_order(jcube):
    """The serial oracle sums each CSR slot in ascending element order:
    recompute slot sums with an explicit Python loop over elements for a
    few rows (brute force) and the clamped variant's PSD property."""
    nodes, el = jcube
    op = O.OraclePicardOperator(nodes, el)
    n = len(el)
    rng = np.random.default_rng(2)
    rt = rng.uniform(0.5, 1.5, n)
    g = rng.standard_normal((n, 3))
    se = -np.abs(rng.standard_normal(n))
    J = op.jacobian_terms12(g, rt, se, 1.0)
    ref = np.zeros_like(J)
    for e in range(n):
        B = op.B[e]
        w1 = rt[e] * op.V[e]
        w2 = 2.0 * se[e] * op.V[e]
        for i in range(4):
            gi = g[e] @ B[i]
            for j in range(4):
                val = w1 * (B[i, 0] * B[j, 0] + B[i, 1] * B[j, 1] + B[i, 2] * B[j, 2])
                val += w2 * gi * (g[e, 0] * B[j, 0] + g[e, 1] * B[j, 1] + g[e, 2] * B[j, 2])
                ref[op.elem_to_csr[e, i, j]] += val
    assert np.abs(J - ref).max() <= 1e-15 * np.abs(ref).max()
    Js = op.jacobian_spd_terms12(g, rt, se * 50.0, 1.0, 0.9)
    A = op._csr(Js).toarray()
    assert np.allclose(A, A.T)
    assert np.linalg.eigvalsh(A).min() > -1e-12 * np.abs(A).max()


def test_laplace_harmonic_quadratic_exact(kuhn8):
    # tests/test_laplace_mms.py:8-17: harmonic quadratics are reproduced
    # exactly by the Kuhn-split Laplace stencil at interior nodes
    nodes, el = kuhn8
    op = O.OraclePicardOperator(nodes, el)
    x, y, z = nodes.T
    for phi in (x * x - y * y, x * y, y * z - 0.5 * x):
        R = op.assemble_residual(phi)
        assert np.abs(R[~cube_boundary_mask(nodes)]).max() < 1e-13


def test_matrix_bit_deterministic(jcube):
    nodes, el = jcube
    op = O.OraclePicardOperator(nodes, el)
    rho = np.linspace(0.9, 1.1, len(el))
    assert np.array_equal(op.assemble_matrix(rho).data, op.assemble_matrix(rho).data)


# ------------------------------------------------------------ upwind
def test_walk_strictly_upstream_uniform_flow(jcube):
    nodes, el = structured_cube_mesh(6)
    upw = O.OracleUpwind(nodes, el)
    op = O.OraclePicardOperator(nodes, el)
    g, _ = op.velocities(nodes[:, 0].copy())
    u = upw.upstream(g)
    disp = (upw.centroids[u] - upw.centroids)[:, 0]
    self_up = u == np.arange(len(el))
    assert np.all(disp[~self_up] < 0.0)
    assert np.all(upw.centroids[self_up, 0] < 0.25)


def _walk_py(fn, cen, nodes, el, g, e, max_hops=4):
    p = nodes[el[e]] @ g
    target = 0.8 * (p.max() - p.min())
    cur, reach, hops = e, 0.0, 0
    while hops < max_hops and reach < target:
        best, bd = -1, 0.0
        for f in range(4):
            nb = fn[cur, f]
            if nb >= 0:
                d = ((cen[nb, 0] - cen[cur, 0]) * g[0] + (cen[nb, 1] - cen[cur, 1]) * g[1]
                     + (cen[nb, 2] - cen[cur, 2]) * g[2])
                if d < bd:
                    bd, best = d, nb
        if best < 0:
            break
        cur, reach, hops = best, reach + (-bd), hops + 1
    return cur


def test_walk_matches_pure_python(jcube):
    nodes, el = jcube
    upw = O.OracleUpwind(nodes, el)
    op = O.OraclePicardOperator(nodes, el)
    g, _ = op.velocities(transonic_field(nodes))
    u = upw.upstream(g)
    for e in range(0, len(el), 5):
        assert u[e] == _walk_py(upw.face_neighbors, upw.centroids, nodes, el, g[e], e)


def test_face_adjacency_symmetric(jcube):
    _, el = jcube
    fn = O.build_face_adjacency(el)
    for e in range(len(el)):
        for f in range(4):
            nb = fn[e, f]
            if nb >= 0:
                assert e in fn[nb]
                shared = set(el[e]) & set(el[nb])
                assert len(shared) == 3 and el[e, f] not in shared


def _branch_counts(nodes, el, m=0.8):
    nodes = nodes * np.array([3.0, 1.0, 1.0])
    op = O.OraclePicardOperator(nodes, el)
    upw = O.OracleUpwind(nodes, el)
    g, q2 = op.velocities(transonic_field(nodes))
    rho = O.density_field(q2, m)
    up, br = upw.freeze_upwind_state(g, q2, rho, m, 1.5, 0.95)
    return op, upw, g, q2, rho, up, br


def test_sensitivities_vs_fd(jcube):
    nodes, el = jcube
    m, C, mc, fl = 0.8, 1.5, 0.95, 0.05
    op, upw, g, q2, rho, up, br = _branch_counts(nodes, el, m)
    assert set(np.unique(br)) >= {0, 1, 2}
    se, su = O.rho_tilde_sensitivities(q2, rho, up, m, mc, C, 1.4, fl)
    h = 1e-7
    n = len(q2)
    # d/dq2_e at fixed upstream (perturb only element e's own q2 & rho)
    for e in range(0, n, 13):
        if br[e] == 3:
            continue
        u = up[e]
        qp, qm = q2.copy(), q2.copy()
        qp[e] += h
        qm[e] -= h
        rp = O.density_field(qp, m)
        rm = O.density_field(qm, m)
        _, tp = O.rho_tilde_frozen_sweep(qp, rp, up, br, m, mc, C, 1.4, fl)
        _, tm = O.rho_tilde_frozen_sweep(qm, rm, up, br, m, mc, C, 1.4, fl)
        fd = (tp[e] - tm[e]) / (2 * h)
        if u == e:
            continue
        assert fd == pytest.approx(se[e], rel=1e-5, abs=1e-8)


def test_frozen_equals_live_at_freeze_state(jcube):
    nodes, el = jcube
    m, C, mc, fl = 0.8, 1.5, 0.95, 0.05
    op, upw, g, q2, rho, up, br = _branch_counts(nodes, el, m)
    nu_l, rt_l = O.rho_tilde_sweep(q2, rho, up, m, mc, C, 1.4, fl)
    nu_f, rt_f = O.rho_tilde_frozen_sweep(q2, rho, up, br, m, mc, C, 1.4, fl)
    assert np.array_equal(rt_l, rt_f)
    se_l, su_l = O.rho_tilde_sensitivities(q2, rho, up, m, mc, C, 1.4, fl)
    se_f, su_f = O.rho_tilde_frozen_sensitivities(q2, rho, up, br, m, mc, C, 1.4)
    ok = br != 3
    assert np.array_equal(se_l[ok], se_f[ok]) and np.array_equal(su_l[ok], su_f[ok])


def test_jacobian_equals_picard_at_zero_sensitivity(jcube):
    nodes, el = jcube
    op = O.OraclePicardOperator(nodes, el)
    phi = transonic_field(nodes)
    rho = np.linspace(0.8, 1.2, len(el))
    z = np.zeros(len(el))
    J = op.assemble_newton_jacobian(phi, rho, z, z, np.arange(len(el)))
    assert np.array_equal(J.data, op.assemble_matrix(rho).data)


def test_jacobian_jvp_vs_fd(jcube):
    """Newton Jacobian (Terms 1-3, frozen selection) against central FD of
    the residual with frozen (upstream, branch)."""
    nodes, el = jcube
    nodes = nodes * np.array([3.0, 1.0, 1.0])
    m, C, mc, fl = 0.8, 1.5, 0.95, 0.05
    op = O.OraclePicardOperator(nodes, el)
    upw = O.OracleUpwind(nodes, el)
    phi = transonic_field(nodes)
    g, q2 = op.velocities(phi)
    rho = O.density_field(q2, m)
    up, br = upw.freeze_upwind_state(g, q2, rho, m, C, mc)
    se, su = O.rho_tilde_frozen_sensitivities(q2, rho, up, br, m, mc, C, 1.4)
    _, rt = O.rho_tilde_frozen_sweep(q2, rho, up, br, m, mc, C, 1.4, fl)
    J = op.assemble_newton_jacobian(phi, rt, se, su, up)
    assert op.n_term3_active > 0

    def R(p):
        gg, qq = op.velocities(p)
        rr = O.density_field(qq, m)
        _, t = O.rho_tilde_frozen_sweep(qq, rr, up, br, m, mc, C, 1.4, fl)
        return op.assemble_residual(p, t)

    rng = np.random.default_rng(7)
    v = rng.standard_normal(len(nodes))
    eps = 1e-6
    fd = (R(phi + eps * v) - R(phi - eps * v)) / (2 * eps)
    jv = J @ v
    assert np.abs(jv - fd).max() < 1e-6 * np.abs(jv).max()


# ------------------------------------------------------------ wall force
def test_wall_force_uniform_flow_zero():
    from up3d_amd.mesh.channel import bump_channel

    mesh = bump_channel(6, 3, 2, bump_height=0.0)
    phi = mesh.nodes[:, 0].copy()
    out = O.wall_force_coefficients(mesh.nodes, mesh.elements,
                                    mesh.boundary_faces["floor"], phi, m_inf=0.5)
    assert np.abs(out["cp_tri"]).max() < 1e-12
    assert np.abs(out["cf"]).max() < 1e-12


# ------------------------------------------- pressure coefficient, wall force
def _cp_local_mach(q2, m, g=1.4):
    """Isentropic Cp written through the local Mach number (the textbook
    route, independent of the density form the code uses):
    M_l^2 = q2 M^2 / (1 + (g-1)/2 M^2 (1 - q2)) from the energy equation,
    p_l / p_inf = ((1 + (g-1)/2 M^2) / (1 + (g-1)/2 M_l^2))^(g/(g-1))."""
    k = 0.5 * (g - 1.0)
    ml2 = q2 * m * m / (1.0 + k * m * m * (1.0 - q2))
    return 2.0 / (g * m * m) * (((1.0 + k * m * m) / (1.0 + k * ml2)) ** (g / (g - 1.0))
                                - 1.0)


def _cp_impls():
    from up3d_amd.post.surface import cp_from_q2

    return {"oracle": O.pressure_coefficient,
            "product": lambda q2, m: float(cp_from_q2(np.asarray(q2), m))}


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_pressure_coefficient_closed_forms(impl):
    cp = _cp_impls()[impl]
    # stagnation point at M 0.5: isentropic p0/p = 1.05^3.5 = 1.186212..., so
    # Cp = (p0/p - 1) / (0.7 * 0.25) = 1.0640722 (tables: p/p0 = 0.8430)
    assert abs(cp(0.0, 0.5) - 1.0640722) < 1e-7
    # small-M series of the stagnation value: 1 + M^2/4 + M^4/40 + ...
    m = 0.05
    assert abs(cp(0.0, m) - (1.0 + m * m / 4.0 + m ** 4 / 40.0)) < 1e-10
    # incompressible limit 1 - q2 and freestream Cp = 0
    assert abs(cp(0.64, 1e-4) - 0.36) < 1e-7
    assert abs(cp(1.0, 0.8)) < 1e-15
    # the local-Mach route at q2 != 1, sub- and supersonic
    for q2, mm in [(1.21, 0.7), (0.5, 0.84), (1.6, 0.8), (0.2, 0.3)]:
        assert abs(cp(q2, mm) - _cp_local_mach(q2, mm)) < 1e-12, (q2, mm)
    # sonic pressure coefficient Cp* = 2/(g M^2) [((2 + (g-1) M^2)/(g+1))^(g/(g-1)) - 1]
    # at the sonic speed q*^2 (M 0.8: -0.4346 in the tables)
    for mm in (0.6, 0.8):
        cps = 2.0 / (1.4 * mm * mm) * (((2.0 + 0.4 * mm * mm) / 2.4) ** 3.5 - 1.0)
        assert abs(cp(O.q2_at_mach(1.0, mm), mm) - cps) < 1e-12
    assert abs(cp(O.q2_at_mach(1.0, 0.8), 0.8) + 0.4346) < 1e-4


def _wall_force_impls():
    from up3d_amd.post.surface import wall_force_coefficients

    return {"oracle": O.wall_force_coefficients, "product": wall_force_coefficients}


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_wall_force_closed_forms(impl):
    from up3d_amd.mesh.channel import bump_channel

    wfc = _wall_force_impls()[impl]
    mesh = bump_channel(12, 4, 3)                  # 10 % arc bump on the floor
    floor = mesh.boundary_faces["floor"]
    nodes, el = mesh.nodes, mesh.elements
    n = len(nodes)
    # phi = 0 at M 0: Cp = 1 on every triangle and cf = -(vector area of the
    # floor, normals into the fluid).  A surface's vector area depends only
    # on its boundary loop, here the 3 x 1 rectangle in y = 0, so the bumped
    # floor gives exactly (0, -3, 0): pins the normal's side and the 1/2
    out = wfc(nodes, el, floor, np.zeros(n), m_inf=0.0)
    assert np.allclose(out["cp_tri"], 1.0, atol=1e-15)
    assert np.allclose(out["cf"], [0.0, -3.0, 0.0], atol=1e-12)
    # phi = c z: its gradient is tangent to the spanwise-uniform bump, so
    # q2 = c^2 on every triangle and cf = (0, -3 Cp(c^2), 0)
    c, m = 1.1, 0.5
    out = wfc(nodes, el, floor, c * nodes[:, 2], m_inf=m)
    cp = _cp_local_mach(c * c, m)
    assert np.allclose(out["cp_tri"], cp, rtol=0, atol=1e-12)
    assert np.allclose(out["cf"], [0.0, -3.0 * cp, 0.0], atol=1e-11)
    # the same field at twice the reference speed with phi scaled alike, and S_ref
    out2 = wfc(nodes, el, floor, 2.0 * c * nodes[:, 2], m_inf=m, u_inf=2.0, s_ref=1.5)
    assert np.allclose(out2["cf"], out["cf"] / 1.5, atol=1e-11)
    # lift / pressure-drag axes rotate with alpha
    a = 3.0
    ar = np.deg2rad(a)
    outa = wfc(nodes, el, floor, c * nodes[:, 2], m_inf=m, alpha_deg=a)
    assert abs(outa["cl"] - (-np.sin(ar) * outa["cf"][0] + np.cos(ar) * outa["cf"][1])) < 1e-14
    assert abs(outa["cd_pressure"] - (np.cos(ar) * outa["cf"][0]
                                      + np.sin(ar) * outa["cf"][1])) < 1e-14
    # a nonuniform field on a rotated copy of the mesh: Cp is invariant and
    # the force rotates with the mesh
    phi = nodes[:, 0] + 0.3 * np.sin(2.0 * nodes[:, 0]) + 0.2 * nodes[:, 2]
    base = wfc(nodes, el, floor, phi, m_inf=0.6)
    th = np.deg2rad(37.0)
    Rz = np.array([[np.cos(th), -np.sin(th), 0.0], [np.sin(th), np.cos(th), 0.0],
                   [0.0, 0.0, 1.0]])
    rot = wfc(nodes @ Rz.T, el, floor, phi, m_inf=0.6)
    assert np.allclose(rot["cp_tri"], base["cp_tri"], rtol=0, atol=1e-12)
    assert np.allclose(rot["cf"], Rz @ base["cf"], rtol=0, atol=1e-12)
    assert abs(base["cf"][0]) > 1e-3               # a real x-force, not zero


def test_wall_force_oracle_equals_product_on_a_transonic_state():
    from up3d_amd.mesh.channel import bump_channel
    from up3d_amd.post.surface import wall_force_coefficients

    mesh = bump_channel(12, 4, 3, jitter=0.15, seed=4)
    x = mesh.nodes[:, 0]
    phi = 1.25 * x + 0.35 * np.sin(1.7 * x) + 0.05 * mesh.nodes[:, 1] ** 2
    a = O.wall_force_coefficients(mesh.nodes, mesh.elements, mesh.boundary_faces["floor"],
                                  phi, m_inf=0.8, alpha_deg=1.0)
    b = wall_force_coefficients(mesh.nodes, mesh.elements, mesh.boundary_faces["floor"],
                                phi, m_inf=0.8, alpha_deg=1.0)
    assert np.allclose(a["cp_tri"], b["cp_tri"], rtol=1e-13, atol=1e-14)
    assert np.allclose(a["cf"], b["cf"], rtol=1e-12, atol=1e-14)

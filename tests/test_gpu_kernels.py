"""GPU parity of the HIP element kernels against the oracle (bit-exact for
integer work and for the uncontracted FP64 kernels fed identical inputs;
rho-dependent values within a few ulp because device pow() and glibc pow()
may round differently)."""

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O

from .helpers import jittered_cube, transonic_field

pytestmark = pytest.mark.gpu

M, C, MC, FL, G = 0.8, 1.5, 0.95, 0.05, 1.4


@pytest.fixture(scope="module")
def mesh(gpu):
    from up3d_amd.mesh.channel import bump_channel

    m = bump_channel(18, 6, 4, jitter=0.15, seed=11, rcm=True)
    print(f"[mesh] channel 18x6x4 n_tets={m.n_tets}", flush=True)
    return m


@pytest.fixture(scope="module")
def ora(mesh):
    op = O.OraclePicardOperator(mesh.nodes, mesh.elements)
    upw = O.OracleUpwind(mesh.nodes, mesh.elements)
    phi = transonic_field(mesh.nodes)
    g, q2 = op.velocities(phi)
    rho = O.density_field(q2, M)
    up = upw.upstream(g)
    opc = O.OraclePicardOperator(mesh.nodes, mesh.elements, order="colored")
    return dict(op=op, opc=opc, upw=upw, phi=phi, g=g, q2=q2, rho=rho, up=up)


def _ulp_close(a, b, n_ulp=8):
    a = np.asarray(a)
    b = np.asarray(b)
    tol = n_ulp * np.spacing(np.maximum(np.abs(a), np.abs(b)))
    return np.all(np.abs(a - b) <= tol)


def test_velocity_bitwise(mesh, ora):
    from up3d_amd.kernels.gradient import element_velocity_q2

    n = mesh.n_tets
    g = np.empty((n, 3))
    q2 = np.empty(n)
    element_velocity_q2(mesh.elements, ora["op"].B, ora["phi"], g, q2)
    assert np.array_equal(g, ora["g"]) and np.array_equal(q2, ora["q2"])


def test_physics_fields(gpu):
    from up3d_amd.physics import isentropic as P

    q2 = np.linspace(0.0, 2.5, 1001)
    for m in (0.0, 0.5, 0.84):
        assert _ulp_close(P.density_field(q2, m), O.density_field(q2, m), 4)
        assert np.array_equal(P.mach_squared_field(q2, m), O.mach_squared_field(q2, m))
        assert np.array_equal(P.limit_q2_field(q2, m, 3.0), O.limit_q2_field(q2, m, 3.0))
    assert np.all(P.density_field(q2, 0.0) == 1.0)


def test_upstream_bitwise(mesh, ora):
    from up3d_amd.kernels.upwind import upstream_elements

    out = np.empty(mesh.n_tets, dtype=np.int64)
    upstream_elements(ora["upw"].face_neighbors, ora["upw"].centroids, mesh.nodes,
                      mesh.elements, ora["g"], out)
    assert np.array_equal(out, ora["up"])
    assert np.count_nonzero(out != np.arange(mesh.n_tets)) > mesh.n_tets // 2


def test_sweeps(mesh, ora):
    from up3d_amd.kernels import upwind as U

    n = mesh.n_tets
    q2, rho, up = ora["q2"], ora["rho"], ora["up"]
    nu = np.empty(n)
    rt = np.empty(n)
    U.rho_tilde_sweep(q2, rho, up, M, MC, C, G, FL, nu, rt)
    nu_o, rt_o = O.rho_tilde_sweep(q2, rho, up, M, MC, C, G, FL)
    assert np.array_equal(nu, nu_o) and np.array_equal(rt, rt_o)
    br = np.empty(n, dtype=np.int8)
    U.classify_upwind_branches(q2, rho, up, M, MC, C, G, FL, br)
    br_o = O.classify_upwind_branches(q2, rho, up, M, MC, C, G, FL)
    assert np.array_equal(br, br_o)
    assert set(np.unique(br)) >= {0, 1, 2}
    U.rho_tilde_frozen_sweep(q2, rho, up, br, M, MC, C, G, FL, nu, rt)
    assert np.array_equal(rt, O.rho_tilde_frozen_sweep(q2, rho, up, br, M, MC, C, G, FL)[1])
    se = np.empty(n)
    su = np.empty(n)
    U.rho_tilde_sensitivities_sweep(q2, rho, up, M, MC, C, G, FL, se, su)
    se_o, su_o = O.rho_tilde_sensitivities(q2, rho, up, M, MC, C, G, FL)
    assert _ulp_close(se, se_o) and _ulp_close(su, su_o)
    U.rho_tilde_frozen_sensitivities_sweep(q2, rho, up, br, M, MC, C, G, FL, se, su)
    se_o, su_o = O.rho_tilde_frozen_sensitivities(q2, rho, up, br, M, MC, C, G)
    assert _ulp_close(se, se_o) and _ulp_close(su, su_o)


def test_colored_assembly_bitwise(mesh, ora):
    from up3d_amd.kernels.jacobian import (assemble_jacobian_data_colored,
                                           assemble_matrix_data_colored,
                                           fill_term3_coo)
    from up3d_amd.kernels.residual import assemble_residual_colored

    op = ora["opc"]
    n = mesh.n_tets
    rt = np.linspace(0.7, 1.3, n)
    R = np.empty(mesh.n_nodes)
    assemble_residual_colored(op.color_offsets, op.color_elems, mesh.elements,
                              op.B, op.V, rt, ora["g"], R)
    assert np.array_equal(R, op.assemble_residual(ora["phi"], rt, ora["g"]))
    data = np.zeros(len(op.indices))
    assemble_matrix_data_colored(op.color_offsets, op.color_elems, op.elem_to_csr,
                                 op.B, op.V, rt, data)
    assert np.array_equal(data, op.assemble_matrix(rt).data)
    se = np.sin(np.arange(n)) * (np.arange(n) % 3 != 0)
    data[:] = 0.0
    assemble_jacobian_data_colored(op.color_offsets, op.color_elems,
                                   op.elem_to_csr, op.B, op.V, rt, ora["g"], se,
                                   1.0, data)
    assert np.array_equal(data, op.jacobian_terms12(ora["g"], rt, se, 1.0))
    su = np.cos(np.arange(n)) * (np.arange(n) % 4 == 0)
    active, r_o, c_o, v_o = op.term3_coo(ora["g"], ora["up"], su, 1.0)
    rows = np.empty(16 * len(active), dtype=np.int64)
    cols = np.empty_like(rows)
    vals = np.empty(16 * len(active))
    fill_term3_coo(active, mesh.elements, op.B, op.V, ora["g"], ora["up"], su,
                   1.0, rows, cols, vals)
    assert np.array_equal(rows, r_o) and np.array_equal(cols, c_o)
    assert np.array_equal(vals, v_o)


def test_gather_assembly_bitwise(mesh, ora):
    """Operator hot path (node gather, one launch) == the oracle's serial
    element-order assembly, bit for bit: residual, Picard matrix, Terms 1+2
    with a limiter mask, and the clamped SPD variant."""
    import torch

    from up3d_amd.device import to_dev, to_host
    from up3d_amd.kernels.jacobian import PicardOperator

    op = PicardOperator(mesh.nodes, mesh.elements)
    # the oracle's assembly fed the device's own geometry (the host set-up's
    # cofactor B_e agrees with the oracle's to ~1e-15, checked elsewhere)
    oo = O.OraclePicardOperator(mesh.nodes, mesh.elements)
    oo.B = np.ascontiguousarray(op.store.B_host.reshape(-1, 4, 3))
    oo.V = np.ascontiguousarray(op.store.V_host)
    n = mesh.n_tets
    rt = np.linspace(0.7, 1.3, n)
    g = ora["g"]
    se = np.sin(np.arange(n)) * (np.arange(n) % 3 != 0)
    lim = (np.arange(n) % 5 != 0).astype(np.uint8)
    rtd, gd, sed = to_dev(rt), to_dev(g.reshape(-1, 3)), to_dev(se)
    limd = to_dev(lim, torch.uint8)
    R = to_host(op.residual_dev(rtd, gd, torch.empty(mesh.n_nodes, dtype=torch.float64,
                                                     device=rtd.device)))
    assert np.array_equal(R, oo.assemble_residual(ora["phi"], rt, g))
    # residual terms from the state sweep's grad . B (gdotb): the same bits
    ph = to_dev(ora["phi"])
    gst = torch.empty((n, 3), dtype=torch.float64, device=rtd.device)
    q2n, q2l, rh = (torch.empty(n, dtype=torch.float64, device=rtd.device)
                    for _ in range(3))
    gb = torch.empty(4 * n, dtype=torch.float64, device=rtd.device)
    op.state_dev(ph, 1.0, M, 1.4, 1e30, gst, q2n, q2l, rh, None, gdotb=gb)
    Ra = to_host(op.residual_dev(rtd, gst, torch.empty(mesh.n_nodes, dtype=torch.float64,
                                                       device=rtd.device)))
    Rb = to_host(op.residual_dev(rtd, gst, torch.empty(mesh.n_nodes, dtype=torch.float64,
                                                       device=rtd.device), gdotb=gb))
    assert np.array_equal(Ra, Rb) and np.abs(Ra).max() > 0.0
    out = torch.full((op.nnz,), np.nan, dtype=torch.float64, device=rtd.device)
    assert np.array_equal(to_host(op.matrix_dev(rtd, out)), oo.assemble_matrix(rt).data)
    out.fill_(np.nan)
    J = to_host(op.jacobian_dev(rtd, gd, sed, limd, 1.0, out))
    assert np.array_equal(J, oo._serial(rt, g, se, lim, 1.0, 0.0))
    assert np.array_equal(to_host(op.jacobian_dev(rtd, gd, sed, None, 0.5, out)),
                          oo.jacobian_terms12(g, rt, se, 0.5))
    Js = to_host(op.jacobian_spd_dev(rtd, gd, -(sed.abs() * 40.0), limd, 1.0, 0.9, out))
    assert np.array_equal(Js, oo.jacobian_spd_terms12(g, rt, -np.abs(se) * 40.0, 1.0,
                                                      0.9, lim))
    # element-order terms (inc_pos NULL) + indexed gathers (inc_ent): the
    # same sums as the incidence-ordered path the operator runs
    from up3d_amd import _native
    from up3d_amd.device import ptr, stream_handle

    st = op.store
    t = torch.empty(4 * n, dtype=torch.float64, device=rtd.device)
    R2 = torch.empty(mesh.n_nodes, dtype=torch.float64, device=rtd.device)
    _native.call("fp3d_element_residual_terms", n, ptr(st.B), ptr(st.V), ptr(rtd),
                 ptr(gd), None, None, ptr(t), stream_handle())
    _native.call("fp3d_gather_node_terms", mesh.n_nodes, ptr(st.inc_ptr),
                 ptr(st.inc_ent), ptr(t), ptr(R2), stream_handle())
    assert np.array_equal(to_host(R2), R)
    K = torch.empty(16 * n, dtype=torch.float64, device=rtd.device)
    out.fill_(np.nan)
    _native.call("fp3d_element_matrix_terms", n, ptr(st.B), ptr(st.V), ptr(rtd), ptr(gd),
                 ptr(sed), ptr(limd), 1.0, 0.0, None, ptr(K), stream_handle())
    _native.call("fp3d_gather_csr_terms", mesh.n_nodes, st.gather_max_seg, ptr(st.indptr),
                 ptr(st.inc_ptr), ptr(st.inc_ent), ptr(st.rel8), ptr(K), ptr(out),
                 stream_handle())
    assert np.array_equal(to_host(out), J)
    # serial order vs the reference's colour order: same sums to 1e-13
    Jc = ora["opc"].jacobian_terms12(g, rt, se, 1.0)
    assert np.abs(to_host(op.jacobian_dev(rtd, gd, sed, None, 1.0, out)) - Jc).max() \
        <= 1e-13 * np.abs(Jc).max()


def test_coloured_scatter_fallback(mesh, ora):
    """A mesh the node gather cannot take (a CSR row block beyond the LDS
    bound) assembles its matrices by the coloured scatter: forced here on
    the test channel, the Picard matrix and Terms 1+2 equal the oracle's
    colour-order assembly bit for bit, the SPD variant the serial one to
    1e-13."""
    import torch

    from up3d_amd.device import to_dev, to_host
    from up3d_amd.kernels.jacobian import PicardOperator

    op = PicardOperator(mesh.nodes, mesh.elements)
    st = op.store
    oc = O.OraclePicardOperator(mesh.nodes, mesh.elements, order="colored")
    oc.B = np.ascontiguousarray(st.B_host.reshape(-1, 4, 3))
    oc.V = np.ascontiguousarray(st.V_host)
    assert np.array_equal(oc.color_offsets, st.color_offsets)
    os_ = O.OraclePicardOperator(mesh.nodes, mesh.elements)
    os_.B, os_.V = oc.B, oc.V
    n = mesh.n_tets
    rt = np.linspace(0.7, 1.3, n)
    g = ora["g"]
    se = np.sin(np.arange(n)) * (np.arange(n) % 3 != 0)
    lim = (np.arange(n) % 5 != 0).astype(np.uint8)
    rtd, gd, sed = to_dev(rt), to_dev(g.reshape(-1, 3)), to_dev(se)
    limd = to_dev(lim, torch.uint8)
    out = torch.full((op.nnz,), np.nan, dtype=torch.float64, device=rtd.device)
    assert st.gather_ok
    st.gather_ok = False
    try:
        A = to_host(op.matrix_dev(rtd, out))
        J = to_host(op.jacobian_dev(rtd, gd, sed, limd, 0.5, out))
        Js = to_host(op.jacobian_spd_dev(rtd, gd, -(sed.abs() * 40.0), limd, 1.0, 0.9,
                                         out))
    finally:
        st.gather_ok = True
    assert np.array_equal(A, oc.assemble_matrix(rt).data)
    assert np.array_equal(J, oc.jacobian_terms12(g, rt, se * lim.astype(np.float64), 0.5))
    ref = os_.jacobian_spd_terms12(g, rt, -np.abs(se) * 40.0, 1.0, 0.9, lim)
    assert np.abs(Js - ref).max() <= 1e-13 * np.abs(ref).max()


def test_picard_operator_api(mesh, ora):
    from up3d_amd.kernels.jacobian import PicardOperator

    op = PicardOperator(mesh.nodes, mesh.elements)
    oo = ora["op"]
    assert op.n_colors == oo.n_colors
    g, q2 = op.velocities(ora["phi"])
    assert np.abs(g - ora["g"]).max() < 1e-12 * np.abs(ora["g"]).max()
    rt = np.linspace(0.7, 1.3, mesh.n_tets)
    A = op.assemble_matrix(rt)
    Ao = oo.assemble_matrix(rt)
    assert np.array_equal(A.indptr, Ao.indptr) and np.array_equal(A.indices, Ao.indices)
    assert np.abs(A.data - Ao.data).max() < 1e-12 * np.abs(Ao.data).max()
    R = op.assemble_residual(ora["phi"], rt)
    Ro = oo.assemble_residual(ora["phi"], rt)
    assert np.abs(R - Ro).max() < 1e-12 * np.abs(Ro).max()
    # bit-determinism across calls
    assert np.array_equal(op.assemble_matrix(rt).data, A.data)
    assert np.array_equal(op.assemble_residual(ora["phi"], rt), R)
    # Laplace limit: rho_tilde=None
    assert np.abs(op.assemble_matrix().data - oo.assemble_matrix().data).max() < 1e-12


def test_upwind_operator_and_jacobian(mesh, ora):
    from up3d_amd.kernels.jacobian import PicardOperator
    from up3d_amd.kernels.upwind import UpwindOperator

    op = PicardOperator(mesh.nodes, mesh.elements)
    upw = UpwindOperator(mesh.nodes, mesh.elements)
    q2, rho, g = ora["q2"], ora["rho"], ora["g"]
    rt = upw.rho_tilde(g, q2, rho, M, C, MC)
    rt_o = ora["upw"].rho_tilde(g, q2, rho, M, C, MC)
    assert _ulp_close(rt, rt_o)
    assert upw.nu_max == ora["upw"].nu_max
    assert upw.n_supersonic == ora["upw"].n_supersonic
    se, su, up = upw.rho_tilde_sensitivities(g, q2, rho, M, C, MC)
    se_o, su_o, up_o = ora["upw"].rho_tilde_sensitivities(g, q2, rho, M, C, MC)
    assert np.array_equal(up, up_o)
    assert _ulp_close(se, se_o) and _ulp_close(su, su_o)
    upf, brf = upw.freeze_upwind_state(g, q2, rho, M, C, MC)
    upf_o, brf_o = ora["upw"].freeze_upwind_state(g, q2, rho, M, C, MC)
    assert np.array_equal(upf, upf_o) and np.array_equal(brf, brf_o)
    J = op.assemble_newton_jacobian(ora["phi"], rt, se, su, up)
    Jo = ora["op"].assemble_newton_jacobian(ora["phi"], rt_o, se_o, su_o, up_o)
    assert op.n_term3_active == ora["op"].n_term3_active > 0
    D = (J - Jo).tocsr()
    assert abs(D).max() < 1e-12 * abs(Jo).max()


def test_term3_device_operator(mesh, ora):
    """Matrix-free Term 3 inside the SpMV == scipy (J + COO) on free rows."""
    import torch

    from up3d_amd.device import to_dev, to_host
    from up3d_amd.kernels.jacobian import PicardOperator
    from up3d_amd.kernels.upwind import UpwindOperator
    from up3d_amd.solve.linear import NewtonOperator

    op = PicardOperator(mesh.nodes, mesh.elements)
    upw = UpwindOperator(mesh.nodes, mesh.elements)
    st = op.store
    q2, rho, g = ora["q2"], ora["rho"], ora["g"]
    gs, q2s, rhos = st.to_storage(g.reshape(-1, 3)), st.to_storage(q2), st.to_storage(rho)
    rts = upw.rho_tilde_dev(gs, q2s, rhos, M, C, MC).clone()
    se, su, up = upw.sensitivities_dev(gs, q2s, rhos, M, C, MC)
    lim = torch.ones(mesh.n_tets, dtype=torch.uint8, device=st.device)
    dn = np.unique(mesh.boundary_faces["inflow"])
    fm = np.ones(mesh.n_nodes, dtype=np.uint8)
    fm[dn] = 0
    jop = NewtonOperator(st, to_dev(fm, torch.uint8))
    op.jacobian_dev(rts, gs, se, lim, 1.0, jop.data)
    jop.finalize()
    jop.build_term3(gs, up, su, lim, 1.0)
    na = jop.n_active
    assert na > 0
    # oracle Jacobian (original order) restricted to free dofs
    rt_o = ora["upw"].rho_tilde(g, q2, rho, M, C, MC)
    se_o, su_o, up_o = ora["upw"].rho_tilde_sensitivities(g, q2, rho, M, C, MC)
    Jo = ora["op"].assemble_newton_jacobian(ora["phi"], rt_o, se_o, su_o, up_o)
    free = np.nonzero(fm)[0]
    x = np.random.default_rng(5).standard_normal(mesh.n_nodes)
    x[dn] = 0.0
    y_ref = np.zeros(mesh.n_nodes)
    y_ref[free] = (Jo[free][:, free] @ x[free])
    xd = to_dev(x)
    yd = torch.empty_like(xd)
    for f32 in (False, True):
        jop.is_f32 = 0
        if f32:
            jop.finalize(f32=True)
        jop.matvec(xd, yd)
        y = to_host(yd)
        tol = 1e-12 if not f32 else 1e-6
        assert np.abs(y - y_ref).max() < tol * np.abs(y_ref).max()
    # repeatability
    y1 = to_host(jop.matvec(xd, yd)).copy()
    assert np.array_equal(y1, to_host(jop.matvec(xd, yd)))
    # the build fed the state sweep's grad . B (gdotb) instead of grad and B:
    # the same table, so the same bits
    ph = to_dev(ora["phi"])
    gst = torch.empty((mesh.n_tets, 3), dtype=torch.float64, device=st.device)
    q2n, q2l, rh = (torch.empty(mesh.n_tets, dtype=torch.float64, device=st.device)
                    for _ in range(3))
    gb = torch.empty(4 * mesh.n_tets, dtype=torch.float64, device=st.device)
    op.state_dev(ph, 1.0, M, 1.4, 1e30, gst, q2n, q2l, rh, None, gdotb=gb)
    ys = []
    for g_arg in (None, gb):
        jop.build_term3(gst, up, su, lim, 1.0, gdotb=g_arg)
        ys.append(to_host(jop.matvec(xd, yd)).copy())
    assert jop.n_active == na
    assert np.array_equal(ys[0], ys[1])


def test_fused_walk_sensitivities_bitwise(mesh, ora):
    """Walk + rho_tilde + (s_e, s_u) in one sweep == the walk sweep followed
    by the sensitivities at its upstream (same bits), and both match the
    oracle's upstream / rho_tilde / sensitivities."""
    import torch

    from up3d_amd.device import to_host
    from up3d_amd.kernels.upwind import UpwindOperator

    upw = UpwindOperator(mesh.nodes, mesh.elements)
    st = upw.store
    q2, rho, g = ora["q2"], ora["rho"], ora["g"]
    gs, q2s, rhos = st.to_storage(g.reshape(-1, 3)), st.to_storage(q2), st.to_storage(rho)
    n = mesh.n_tets
    for floor in (0.05, 0.9):
        up_a = torch.empty(n, dtype=torch.int32, device=st.device)
        rt_a = upw.rho_tilde_dev(gs, q2s, rhos, M, C, MC, rho_floor=floor,
                                 up_out=up_a, rt_out=torch.empty_like(q2s)).clone()
        se_a, su_a = (t.clone() for t in upw.sensitivities_at_dev(q2s, rhos, up_a, M, C, MC,
                                                                   rho_floor=floor))
        up_b = torch.empty_like(up_a)
        se_b, su_b = torch.empty_like(q2s), torch.empty_like(q2s)
        rt_b = upw.rho_tilde_dev(gs, q2s, rhos, M, C, MC, rho_floor=floor, up_out=up_b,
                                 rt_out=torch.empty_like(q2s), se_out=se_b, su_out=su_b)
        assert torch.equal(up_a, up_b) and torch.equal(rt_a, rt_b)
        assert torch.equal(se_a, se_b) and torch.equal(su_a, su_b)
        assert np.array_equal(to_host(up_b).astype(np.int64), ora["up"])
        se_o, su_o = O.rho_tilde_sensitivities(q2, rho, ora["up"], M, MC, C, G, floor)
        assert _ulp_close(to_host(se_b), se_o) and _ulp_close(to_host(su_b), su_o)
        if floor > 0.5:
            assert np.count_nonzero(to_host(rt_b) == floor) > 0


def test_walk_width_is_bitwise_neutral(mesh, ora):
    """1 to 4 upstream walks per lane give the same selection, rho_tilde and
    sensitivities, bit for bit, over full and partial blocks (the mesh size
    is not a multiple of W x block), and match the oracle's upstream."""
    import torch

    from up3d_amd import _native
    from up3d_amd.device import to_host
    from up3d_amd.kernels.upwind import UpwindOperator

    upw = UpwindOperator(mesh.nodes, mesh.elements)
    st = upw.store
    q2, rho, g = ora["q2"], ora["rho"], ora["g"]
    gs, q2s, rhos = st.to_storage(g.reshape(-1, 3)), st.to_storage(q2), st.to_storage(rho)
    L = _native.lib()
    prev = int(L.fp3d_get_walk_width())
    out = {}
    try:
        for w in (1, 2, 3, 4):
            _native.check(L.fp3d_set_walk_width(w), "set_walk_width")
            up = torch.full((mesh.n_tets,), -7, dtype=torch.int32, device=st.device)
            se, su = torch.full_like(q2s, float("nan")), torch.full_like(q2s, float("nan"))
            rt = upw.rho_tilde_dev(gs, q2s, rhos, M, C, MC, rho_floor=0.05, up_out=up,
                                   rt_out=torch.full_like(q2s, float("nan")), se_out=se,
                                   su_out=su)
            out[w] = [to_host(t).copy() for t in (up, rt, se, su)]
            # the walk without sensitivities, same width
            up2 = torch.full_like(up, -7)
            rt2 = upw.rho_tilde_dev(gs, q2s, rhos, M, C, MC, rho_floor=0.05, up_out=up2,
                                    rt_out=torch.full_like(q2s, float("nan")))
            assert torch.equal(up2, up) and torch.equal(rt2, rt)
    finally:
        _native.check(L.fp3d_set_walk_width(prev), "set_walk_width")
    assert mesh.n_tets % (2 * int(L.fp3d_get_element_block())) != 0
    for w in (1, 3, 4):
        for a, b in zip(out[w], out[2]):
            assert np.array_equal(a, b)
    assert np.array_equal(out[2][0].astype(np.int64), ora["up"])
    assert L.fp3d_set_walk_width(5) != 0 and "walk_width" in L.fp3d_last_error().decode()


def test_gmres_matches_scipy(mesh, ora, gpu):
    import scipy.sparse.linalg as spla

    from up3d_amd.solve.linear import build_amg_preconditioner, solve_gmres

    rt = ora["upw"].rho_tilde(ora["g"], ora["q2"], ora["rho"], M, C, MC)
    se, su, up = ora["upw"].rho_tilde_sensitivities(ora["g"], ora["q2"], ora["rho"], M, C, MC)
    J = ora["op"].assemble_newton_jacobian(ora["phi"], rt, se, su, up)
    A = ora["op"].assemble_matrix(rt)
    dn = np.unique(mesh.boundary_faces["inflow"])
    fm = np.ones(mesh.n_nodes, bool)
    fm[dn] = False
    free = np.nonzero(fm)[0]
    Jf = J[free][:, free].tocsr()
    Af = A[free][:, free].tocsr()
    b = np.random.default_rng(3).standard_normal(len(free))
    ml, Mp = build_amg_preconditioner(Af)
    x, n_it, info = solve_gmres(Jf, b, M=Mp, rtol=1e-11, restart=30, maxiter=20)
    assert info == 0
    x_ref = spla.spsolve(Jf.tocsc(), b)
    assert np.abs(x - x_ref).max() < 1e-8 * np.abs(x_ref).max()
    assert np.linalg.norm(Jf @ x - b) <= 1.01e-11 * np.linalg.norm(b)


def test_cg_amg_laplace(gpu):
    import scipy.sparse.linalg as spla

    from up3d_amd.solve.linear import solve_cg_amg

    nodes, el = jittered_cube(10, seed=9)
    op = O.OraclePicardOperator(nodes, el)
    from .helpers import cube_boundary_mask

    free = np.nonzero(~cube_boundary_mask(nodes))[0]
    A = op.assemble_matrix()[free][:, free].tocsr()
    b = np.ones(len(free))
    x, n_it = solve_cg_amg(A, b, rtol=1e-12, maxiter=200)
    assert n_it < 60
    x_ref = spla.spsolve(A.tocsc(), b)
    assert np.abs(x - x_ref).max() < 1e-9 * np.abs(x_ref).max()


def test_amg_graph_replay_matches_eager(gpu):
    """The V-cycle replayed from its HIP graphs equals the eager V-cycle on
    every replay, with eager device work in between: graphs captured on the
    caller's own buffers (up to four pairs) and the copy-in/out graph used
    beyond them, for Chebyshev degrees 1, 2 and 3 and both precisions."""
    import torch

    from up3d_amd.device import to_dev, to_host
    from up3d_amd.solve.amg import AMGHierarchy

    from .helpers import cube_boundary_mask

    nodes, el = jittered_cube(12, seed=4)
    op = O.OraclePicardOperator(nodes, el)
    free = np.nonzero(~cube_boundary_mask(nodes))[0]
    A = op.assemble_matrix(np.linspace(0.8, 1.2, len(el)))[free][:, free].tocsr()
    n = A.shape[0]
    for prec, degree in (("f64", 2), ("f32", 2), ("f64", 1), ("f64", 3)):
        ml = AMGHierarchy(A, precision=prec, degree=degree)
        assert len(ml.levels) >= 3          # restriction into a smoothed level
        rng = np.random.default_rng(1)
        xe = torch.empty(n, dtype=torch.float64, device="cuda")
        srcs = [torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(3)]
        dsts = [torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(3)]
        for rep in range(12):
            b, xg = srcs[rep % 3], dsts[(rep // 3) % 3]
            b.copy_(to_dev(rng.standard_normal(n)))
            ml.apply_eager(b, xe)
            ref = to_host(xe).copy()
            ml.apply(b, xg)                    # first call eager, then graphs
            torch.cuda.synchronize()
            assert np.array_equal(to_host(xg), ref), (prec, degree, rep)
            _ = (b * 2.0).sum().item()         # eager work between replays
        # nine buffer pairs were used: four direct graphs, then the copy graph
        assert len(ml._graphs) == ml._MAX_DIRECT_GRAPHS and ml._graph is not None


def test_gather_long_rows_bitwise(gpu):
    """A node of high valence (a star of 236 tets around the centre of a
    120-point sphere: a 121-entry CSR row, 236 incidence entries, just under
    the 256-colour limit the greedy colouring keeps): the node gathers still
    equal the oracle's serial assembly bit for bit, with staged chunks of 256
    and 1024 entries."""
    import ctypes
    import torch
    from scipy.spatial import ConvexHull

    from up3d_amd import _native
    from up3d_amd.device import to_dev, to_host
    from up3d_amd.kernels.jacobian import PicardOperator

    k = np.arange(120) + 0.5
    th = np.arccos(1 - 2 * k / 120)
    ph = np.pi * (1 + 5 ** 0.5) * k
    sph = np.stack([np.cos(ph) * np.sin(th), np.sin(ph) * np.sin(th), np.cos(th)], 1)
    nodes = np.vstack([np.zeros((1, 3)), sph])
    tri = ConvexHull(sph).simplices + 1
    el = np.hstack([np.zeros((len(tri), 1), dtype=np.int64), tri])
    # positive orientation
    d = nodes[el[:, 1:]] - nodes[el[:, :1]]
    neg = np.einsum("ij,ij->i", d[:, 0], np.cross(d[:, 1], d[:, 2])) < 0
    el[neg, 1], el[neg, 2] = el[neg, 2].copy(), el[neg, 1].copy()
    el = el.astype(np.int32)
    op = PicardOperator(nodes, el)
    st = op.store
    row0 = int(st.indptr_host[1] - st.indptr_host[0])
    print(f"[long rows] {len(el)} tets, centre row {row0} entries", flush=True)
    assert row0 == 121 and st.gather_ok
    oo = O.OraclePicardOperator(nodes, el)
    oo.B = np.ascontiguousarray(st.B_host.reshape(-1, 4, 3))
    oo.V = np.ascontiguousarray(st.V_host)
    n = len(el)
    rng = np.random.default_rng(5)
    rt = 0.6 + rng.random(n)
    g = rng.standard_normal((n, 3))
    se = rng.standard_normal(n)
    rtd, gd, sed = to_dev(rt), to_dev(g), to_dev(se)
    out = torch.full((op.nnz,), np.nan, dtype=torch.float64, device=rtd.device)
    L = _native.lib()
    prev = int(L.fp3d_get_gather_chunk())
    try:
        for chunk in (256, 1024):
            assert L.fp3d_set_gather_chunk(ctypes.c_int(chunk)) == 0
            out.fill_(np.nan)
            assert np.array_equal(to_host(op.matrix_dev(rtd, out)), oo.assemble_matrix(rt).data)
            out.fill_(np.nan)
            assert np.array_equal(to_host(op.jacobian_dev(rtd, gd, sed, None, 1.0, out)),
                                  oo.jacobian_terms12(g, rt, se, 1.0))
    finally:
        L.fp3d_set_gather_chunk(ctypes.c_int(prev))

"""Full-size parity of the Newton linearisation on the 4,718,592-tet bump
channel (192x64x64 hexes, jittered, RCM) at M 0.84, against the oracle on
the same device state:

  * the state sweep: velocities and q2 bit for bit, density within 4 ulp
    (device pow);
  * the upstream walk: u(e) bit for bit; nu and rho_tilde bit for bit given
    the same density; the exact sensitivities s_e, s_u within 8 ulp;
  * the residual and the Jacobian Terms 1+2 (with the limiter mask) from
    the node gathers, bit for bit against the oracle's serial element-order
    assembly;
  * the Term-3 build: at this size the active-element scan has more block
    totals than one tile of its second pass holds, so the carried
    multi-tile path runs; the device's Term-3 product (Terms 1+2 values
    zeroed) equals the oracle's Term-3 COO matrix times the same vector."""

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

M, C, MC, FL, G = 0.84, 1.5, 0.95, 0.05, 1.4


def _ulp_close(a, b, n_ulp):
    tol = n_ulp * np.spacing(np.maximum(np.abs(a), np.abs(b)))
    return bool(np.all(np.abs(a - b) <= tol))


@pytest.fixture(scope="module")
def c2(gpu):
    import torch

    from up3d_amd.device import to_dev, to_host
    from up3d_amd.kernels.jacobian import PicardOperator
    from up3d_amd.kernels.upwind import UpwindOperator
    from up3d_amd.mesh.channel import config_mesh
    from up3d_amd.physics.isentropic import q2_at_mach

    mesh = config_mesh("c2")
    print(f"[fullsize] channel 192x64x64 n_tets={mesh.n_tets}", flush=True)
    op = PicardOperator(mesh.nodes, mesh.elements)
    upw = UpwindOperator(mesh.nodes, mesh.elements)
    st = op.store
    nt, dev = st.n_tets, st.device
    x = mesh.nodes[:, 0]
    phi_h = 1.25 * x + 0.35 * np.sin(1.7 * x)
    phi = to_dev(phi_h)
    grad = torch.empty((nt, 3), dtype=torch.float64, device=dev)
    q2n, q2l, rho = (torch.empty(nt, dtype=torch.float64, device=dev) for _ in range(3))
    gb = torch.empty(4 * nt, dtype=torch.float64, device=dev)
    lim = torch.empty(nt, dtype=torch.uint8, device=dev)
    q2cap = q2_at_mach(3.0, M, G)
    op.state_dev(phi, 1.0, M, G, q2cap, grad, q2n, q2l, rho, lim, gdotb=gb)
    se, su = torch.empty_like(q2n), torch.empty_like(q2n)
    upw.rho_tilde_dev(grad, q2l, rho, M, C, MC, G, FL, monitors=False, se_out=se,
                      su_out=su)
    # the oracle's serial assembly fed the device's own geometry (the host
    # set-up's cofactor B_e agrees with the oracle's to ~1e-15, checked on
    # small meshes)
    oo = object.__new__(O.OraclePicardOperator)
    oo.order = "serial"
    oo.elements = np.ascontiguousarray(mesh.elements, dtype=np.int32)
    oo.B = np.ascontiguousarray(st.B_host.reshape(-1, 4, 3))
    oo.V = np.ascontiguousarray(st.V_host)
    oo.n_tets, oo.n_nodes = nt, st.n_nodes
    oo.elem_to_csr = np.ascontiguousarray(st.e2c_host, dtype=np.int32)
    oo.indices = st.indices_host
    oo.rho_ones = np.ones(nt)
    return dict(mesh=mesh, op=op, upw=upw, st=st, phi=phi, phi_h=phi_h, grad=grad,
                q2n=q2n, q2l=q2l, rho=rho, gb=gb, lim=lim, se=se, su=su, oo=oo,
                q2cap=q2cap, h=lambda t: to_host(t))


def test_state_and_walk_full_size(c2):
    h, st, oo = c2["h"], c2["st"], c2["oo"]
    g = h(c2["grad"]).reshape(-1, 3)
    g_o, q2_o = O.element_velocity_q2(oo.elements, oo.B, c2["phi_h"])
    assert np.array_equal(g, g_o) and np.array_equal(h(c2["q2n"]), q2_o)
    q2l = h(c2["q2l"])
    assert np.array_equal(q2l, np.where(q2_o < c2["q2cap"], q2_o, c2["q2cap"]))
    rho = h(c2["rho"])
    assert _ulp_close(rho, O.density_field(q2l, M), 4)
    up_o = O.upstream_elements(st.face_neighbors_host, st.centroids_host, st.nodes_host,
                               oo.elements, g_o)
    up = h(c2["upw"].d_up).astype(np.int64)
    assert np.array_equal(up, up_o)
    assert np.count_nonzero(up != np.arange(len(up))) > len(up) // 2
    nu_o, rt_o = O.rho_tilde_sweep(q2l, rho, up_o, M, MC, C, G, FL)
    assert np.array_equal(h(c2["upw"].d_rt), rt_o)
    assert np.array_equal(h(c2["upw"].d_nu), nu_o)
    n_sup = int(np.count_nonzero(nu_o > 0.0))
    print(f"[fullsize] walk bitwise, {n_sup} elements with nu > 0", flush=True)
    assert n_sup > 100_000
    se_o, su_o = O.rho_tilde_sensitivities(q2l, rho, up_o, M, MC, C, G, FL)
    assert _ulp_close(h(c2["se"]), se_o, 8) and _ulp_close(h(c2["su"]), su_o, 8)


def test_residual_and_jacobian_full_size_bitwise(c2):
    import torch

    h, oo, op = c2["h"], c2["oo"], c2["op"]
    rt = c2["upw"].d_rt
    dev = rt.device
    R = h(op.residual_dev(rt, c2["grad"], torch.empty(oo.n_nodes, dtype=torch.float64,
                                                     device=dev), gdotb=c2["gb"]))
    g = h(c2["grad"]).reshape(-1, 3)
    R_o = oo.assemble_residual(c2["phi_h"], h(rt), g)
    assert np.array_equal(R, R_o)
    out = torch.full((op.nnz,), np.nan, dtype=torch.float64, device=dev)
    J = h(op.jacobian_dev(rt, c2["grad"], c2["se"], c2["lim"], 1.0, out))
    J_o = oo._serial(h(rt), g, h(c2["se"]), h(c2["lim"]), 1.0, 0.0)
    n_diff = int(np.count_nonzero(J != J_o))
    print(f"[fullsize] residual bitwise; Terms 1+2: {len(J)} CSR values, "
          f"{n_diff} differ", flush=True)
    assert n_diff == 0


def test_term3_full_size_matches_oracle(c2):
    import torch

    from up3d_amd.device import to_dev
    from up3d_amd.solve.linear import NewtonOperator

    h, st = c2["h"], c2["st"]
    nt, nn, dev = st.n_tets, st.n_nodes, st.device
    jop = NewtonOperator(st, torch.ones(nn, dtype=torch.uint8, device=dev))
    jop.build_term3(c2["grad"], c2["upw"].d_up, c2["su"], c2["lim"], 1.0, gdotb=c2["gb"])
    jop.data.zero_()
    rng = np.random.default_rng(7)
    xv = rng.standard_normal(nn)
    y = h(jop.matvec(to_dev(xv), torch.empty(nn, dtype=torch.float64, device=dev)))

    up = h(c2["upw"].d_up).astype(np.int64)
    su_eff = h(c2["su"]) * h(c2["lim"])[up]
    n_active = int(np.count_nonzero(su_eff))
    # the element scan's block totals (1024 elements each) exceed one
    # 4096-total tile of the second pass
    assert nt > 4096 * 1024 and n_active > 1_000_000
    assert jop.n_active == n_active
    _, rows, cols, vals = c2["oo"].term3_coo(h(c2["grad"]).reshape(-1, 3), up, su_eff, 1.0)
    y_ref = np.bincount(rows, weights=vals * xv[cols], minlength=nn)
    scale = np.abs(vals).max() * np.abs(xv).max() * 64
    assert np.abs(y - y_ref).max() <= 1e-13 * scale
    assert np.abs(y_ref).max() > 1e-3 * scale / 64

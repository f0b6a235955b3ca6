"""Full-size parity of the Term-3 build on the 4,718,592-tet bump channel
(192x64x64 hexes, jittered, RCM): at this size the active-element scan has
more block totals than one tile of its second pass holds, so the carried
multi-tile path runs.  The device's Term-3 product (Terms 1+2 values zeroed)
is compared with the oracle's Term-3 COO matrix times the same vector, both
built from the same device state (grad, upstream, s_u, limiter), B and V."""

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

M, C, MC, FL, G = 0.84, 1.5, 0.95, 0.05, 1.4


def test_term3_full_size_matches_oracle(gpu):
    import torch

    from up3d_amd.device import to_dev, to_host
    from up3d_amd.kernels.jacobian import PicardOperator
    from up3d_amd.kernels.upwind import UpwindOperator
    from up3d_amd.mesh.channel import config_mesh
    from up3d_amd.physics.isentropic import q2_at_mach
    from up3d_amd.solve.linear import NewtonOperator

    mesh = config_mesh("c2")
    print(f"[fullsize] channel 192x64x64 n_tets={mesh.n_tets}", flush=True)
    op = PicardOperator(mesh.nodes, mesh.elements)
    upw = UpwindOperator(mesh.nodes, mesh.elements)
    st = op.store
    nt, nn, dev = st.n_tets, st.n_nodes, st.device
    x = mesh.nodes[:, 0]
    phi = to_dev(1.25 * x + 0.35 * np.sin(1.7 * x))
    grad = torch.empty((nt, 3), dtype=torch.float64, device=dev)
    q2n, q2l, rho = (torch.empty(nt, dtype=torch.float64, device=dev) for _ in range(3))
    gb = torch.empty(4 * nt, dtype=torch.float64, device=dev)
    lim = torch.empty(nt, dtype=torch.uint8, device=dev)
    op.state_dev(phi, 1.0, M, G, q2_at_mach(3.0, M, G), grad, q2n, q2l, rho, lim,
                 gdotb=gb)
    se, su = torch.empty_like(q2n), torch.empty_like(q2n)
    upw.rho_tilde_dev(grad, q2l, rho, M, C, MC, G, FL, monitors=False, se_out=se,
                      su_out=su)
    jop = NewtonOperator(st, torch.ones(nn, dtype=torch.uint8, device=dev))
    jop.build_term3(grad, upw.d_up, su, lim, 1.0, gdotb=gb)
    jop.data.zero_()
    rng = np.random.default_rng(7)
    xv = rng.standard_normal(nn)
    y = to_host(jop.matvec(to_dev(xv), torch.empty(nn, dtype=torch.float64, device=dev)))

    up = to_host(upw.d_up).astype(np.int64)
    su_eff = to_host(su) * to_host(lim)[up]
    n_active = int(np.count_nonzero(su_eff))
    # the element scan's block totals (1024 elements each) exceed one
    # 4096-total tile of the second pass
    assert nt > 4096 * 1024 and n_active > 1_000_000
    assert jop.n_active == n_active
    oo = object.__new__(O.OraclePicardOperator)
    oo.elements = np.ascontiguousarray(mesh.elements, dtype=np.int32)
    oo.B = np.ascontiguousarray(st.B_host.reshape(-1, 4, 3))
    oo.V = np.ascontiguousarray(st.V_host)
    _, rows, cols, vals = oo.term3_coo(to_host(grad).reshape(-1, 3), up, su_eff, 1.0)
    y_ref = np.bincount(rows, weights=vals * xv[cols], minlength=nn)
    scale = np.abs(vals).max() * np.abs(xv).max() * 64
    assert np.abs(y - y_ref).max() <= 1e-13 * scale
    assert np.abs(y_ref).max() > 1e-3 * scale / 64

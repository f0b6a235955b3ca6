"""AMG set-up from a device matrix (solve/common.py amg_from_device): the
finest level's rho(D^-1 A) estimated on the device agrees with the host
power iteration to rounding, and the hierarchy whose finest level keeps the
device values and pattern applies the V-cycle bit for bit like the one
built from the host copy with the same rho -- in float64 and float32."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def masked_matrix(gpu):
    import torch

    from up3d_amd.device import to_dev
    from up3d_amd.kernels.jacobian import PicardOperator
    from up3d_amd.mesh.channel import boundary_nodes, bump_channel
    from up3d_amd.solve.common import dirichlet_setup, mask_dirichlet

    m = bump_channel(24, 8, 6, jitter=0.15, seed=11, rcm=True)
    op = PicardOperator(m.nodes, m.elements)
    rng = np.random.default_rng(3)
    A = torch.empty(op.store.nnz, dtype=torch.float64, device=gpu)
    op.matrix_dev(to_dev(0.5 + rng.random(m.n_tets)), A)
    dn = boundary_nodes(m, ["inflow", "outflow"])
    fm = dirichlet_setup(m.n_nodes, dn, np.zeros(len(dn)), gpu)[0]
    mask_dirichlet(op.store, fm, A)
    print(f"[amg-setup] channel 24x8x6 n_nodes={m.n_nodes} nnz={op.store.nnz}", flush=True)
    return op.store, A


def test_device_rho_matches_host(masked_matrix):
    from up3d_amd.device import to_dev
    from up3d_amd.solve import amg
    from up3d_amd.solve.common import csr_to_host

    st, A = masked_matrix
    Ah = csr_to_host(st, A)
    d = Ah.diagonal()
    dinv = np.where(d != 0.0, 1.0 / np.where(d != 0.0, d, 1.0), 0.0)
    host = amg.spectral_radius_dinv_a(Ah, dinv, seed=0)
    dev = amg.spectral_radius_dinv_a_dev(st.n_nodes, st.nnz, st.indptr, st.indices, A,
                                         to_dev(dinv), seed=0)
    print(f"[amg-setup] rho host {host!r} device {dev!r}", flush=True)
    assert abs(dev - host) <= 1e-12 * host


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_device_finest_level_equals_host_built(masked_matrix, precision):
    import torch

    from up3d_amd.device import to_dev
    from up3d_amd.solve.amg import AMGHierarchy
    from up3d_amd.solve.common import amg_from_device, csr_to_host

    st, A = masked_matrix
    dev = amg_from_device(st, A, precision, max_coarse=100)
    host = AMGHierarchy(csr_to_host(st, A), precision=precision, seed=0, max_coarse=100,
                        rho0=dev.levels[0].rho)
    assert dev.shapes == host.shapes and len(dev.shapes) >= 2
    a, b = dev.levels[0].A, host.levels[0].A
    assert torch.equal(a.values, b.values) and torch.equal(a.indices16, b.indices16)
    assert [lv.rho for lv in dev.levels] == [lv.rho for lv in host.levels]
    rng = np.random.default_rng(5)
    for rnd in range(2):
        print(f"[amg-setup] {precision} V-cycle round {rnd}", flush=True)
        rhs = to_dev(rng.standard_normal(dev.shapes[0]))
        x1, x2 = torch.empty_like(rhs), torch.empty_like(rhs)
        dev.apply(rhs, x1)
        host.apply(rhs, x2)
        torch.cuda.synchronize()
        assert torch.equal(x1, x2)
    # the finest level owns its values: a later write to the source matrix
    # leaves the hierarchy unchanged
    keep = a.values.clone()
    A.mul_(2.0)
    try:
        assert torch.equal(a.values, keep)
    finally:
        A.mul_(0.5)

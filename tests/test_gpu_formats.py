"""Storage formats of the Krylov kernels: int16 row-relative CSR columns
(square matrices whose bandwidth fits) give bit-identical V-cycles and
Jacobian matvecs in float64; the V-cycle with float32 level vectors (the
FP32-preconditioned solve's) stays within float32 rounding of the float64
one, and its HIP-graph replays equal its eager applications."""

import numpy as np
import pytest

from oracle import oracle as O

from .helpers import cube_boundary_mask, jittered_cube

pytestmark = pytest.mark.gpu


def _matrix():
    nodes, el = jittered_cube(12, seed=4)
    op = O.OraclePicardOperator(nodes, el)
    free = np.nonzero(~cube_boundary_mask(nodes))[0]
    return op.assemble_matrix(np.linspace(0.8, 1.2, len(el)))[free][:, free].tocsr()


def test_int16_columns_are_bitwise_neutral(gpu):
    import torch

    from up3d_amd.device import to_dev, to_host
    from up3d_amd.solve.amg import AMGHierarchy

    A = _matrix()
    n = A.shape[0]
    a16 = AMGHierarchy(A, precision="f64", compact_index=True, use_graph=False)
    a32 = AMGHierarchy(A, precision="f64", compact_index=False, use_graph=False)
    assert a16.levels[0].A.indices16 is not None and a32.levels[0].A.indices16 is None
    b = to_dev(np.random.default_rng(2).standard_normal(n))
    x1 = torch.empty_like(b)
    x2 = torch.empty_like(b)
    a16.apply(b, x1)
    a32.apply(b, x2)
    assert np.array_equal(to_host(x1), to_host(x2))


def test_jacobian_matvec_int16_bitwise(gpu):
    import torch

    from up3d_amd.device import to_dev, to_host
    from up3d_amd.mesh.channel import boundary_nodes, bump_channel
    from up3d_amd.solve.newton import NewtonWorkspace

    m = bump_channel(24, 8, 4, jitter=0.15, seed=5, rcm=True)
    dn = boundary_nodes(m, ["inflow", "outflow"])
    ws = NewtonWorkspace(m.nodes, m.elements, dn, m.nodes[dn, 0])
    ws.set_mach(0.8)
    s = ws.states[0]
    x = m.nodes[:, 0]
    s.phi.copy_(to_dev(1.25 * x + 0.35 * np.sin(1.7 * x)))
    s.phi.index_copy_(0, ws.dn, ws.dv)
    ws.eval_residual(s, 1.5, 0.95, 0.05)
    ws.assemble_coupled(s, 1.5, 0.95, 0.05, None)
    assert ws.st.indices16 is not None and ws.jac.n_active > 0
    v = to_dev(np.random.default_rng(4).standard_normal(ws.n)) * ws.fm.double()
    y16 = to_host(ws.jac.matvec(v, torch.empty_like(v))).copy()
    keep = ws.st.indices16
    try:
        ws.st.indices16 = None
        y32 = to_host(ws.jac.matvec(v, torch.empty_like(v)))
    finally:
        ws.st.indices16 = keep
    assert np.array_equal(y16, y32)


def test_float32_vector_vcycle(gpu):
    import torch

    from up3d_amd.device import to_dev, to_host
    from up3d_amd.solve.amg import AMGHierarchy

    A = _matrix()
    n = A.shape[0]
    ref = AMGHierarchy(A, precision="f64", use_graph=False)
    ml = AMGHierarchy(A, precision="f32")
    assert ml.v32 and ml.levels[1].r.dtype == torch.float32
    rng = np.random.default_rng(1)
    xe = torch.empty(n, dtype=torch.float64, device="cuda")
    xr = torch.empty_like(xe)
    srcs = [torch.empty_like(xe) for _ in range(3)]
    dsts = [torch.empty_like(xe) for _ in range(3)]
    for rep in range(9):
        b, xg = srcs[rep % 3], dsts[(rep // 3) % 3]
        b.copy_(to_dev(rng.standard_normal(n)))
        ml.apply_eager(b, xe)
        ml.apply(b, xg)                  # first call eager, then graphs
        torch.cuda.synchronize()
        assert np.array_equal(to_host(xg), to_host(xe)), rep
        ref.apply(b, xr)
        e, r = to_host(xe), to_host(xr)
        assert np.abs(e - r).max() <= 1e-5 * np.abs(r).max(), rep
        _ = (b * 2.0).sum().item()       # eager work between replays

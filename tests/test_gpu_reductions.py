"""Deterministic device reductions used by GMRES (multi-dot with the fused
||w||^2 row, DGKS-conditional re-orthogonalisation, multi-axpy), checked
against numpy float64 (math.fsum for the reference sums) on sizes that hit
the staged (LDS) and unstaged row slices, ragged tails and a single block."""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,nv", [(1, 1), (37, 3), (1000, 9), (70001, 17),
                                  (815425, 101), (4_500_001, 3)])
def test_multi_dot_matches_exact_sums(gpu, n, nv):
    from up3d_amd import _native
    from up3d_amd.device import ptr, stream_handle, to_dev, to_host
    from up3d_amd.solve.linear import _Red

    print(f"[reduce] multi_dot n={n} nv={nv}", flush=True)
    rng = np.random.default_rng(n + nv)
    Vh = rng.standard_normal((nv + 1, n))
    V = to_dev(Vh)
    red = _Red(nv + 1)
    red.dots(V, nv + 1, V[nv], 2)           # the last row is w itself
    out = to_host(red.out[2:3 + nv])
    w = Vh[nv]
    for k in range(nv + 1):
        exact = math.fsum(Vh[k] * w)
        scale = math.fsum(np.abs(Vh[k] * w))
        assert abs(out[k] - exact) <= 1e-14 * scale + 1e-300, (k, out[k], exact)
    again = to_host(red.out[2:3 + nv]).copy()
    red.dots(V, nv + 1, V[nv], 2)
    assert np.array_equal(to_host(red.out[2:3 + nv]), again)     # bitwise repeatable
    # the fused self row (fp3d_multi_dot_norm): the same V . w and, in
    # out[nv], the bits of a one-row call on w
    import torch

    fused = torch.empty(nv + 1, dtype=torch.float64, device=V.device)
    _native.call("fp3d_multi_dot_norm", n, nv, ptr(V), V.stride(0), ptr(V[nv]),
                 ptr(red.work), ptr(fused), stream_handle())
    single = torch.empty(1, dtype=torch.float64, device=V.device)
    _native.call("fp3d_multi_dot", n, 1, ptr(V[nv]), n, ptr(V[nv]), ptr(red.work),
                 ptr(single), stream_handle())
    fh = to_host(fused)
    assert np.array_equal(fh[:nv], again[:nv]) and fh[nv] == to_host(single)[0]
    V32 = V[:nv].to(torch.float32).contiguous()
    _native.call("fp3d_multi_dot_norm_f32v", n, nv, ptr(V32), V32.stride(0), ptr(V[nv]),
                 ptr(red.work), ptr(fused), stream_handle())
    assert to_host(fused)[nv] == fh[nv]
    # DGKS-conditional pass: no cancellation -> exact zeros
    cond = to_dev(np.concatenate([np.zeros(nv), [1.0]]))
    _native.call("fp3d_multi_dot_reorth", n, nv, ptr(V), V.stride(0), ptr(V[nv]),
                 ptr(cond), 0.5, ptr(red.work), ptr(red.out), stream_handle())
    assert not np.any(to_host(red.out[:nv]))
    # heavy cancellation -> the pass runs
    cond = to_dev(np.concatenate([np.ones(nv), [1.0]]))
    _native.call("fp3d_multi_dot_reorth", n, nv, ptr(V), V.stride(0), ptr(V[nv]),
                 ptr(cond), 0.5, ptr(red.work), ptr(red.out), stream_handle())
    assert np.allclose(to_host(red.out[:nv]), out[:nv], rtol=0, atol=1e-12 * n)
    # multi-axpy: w - sum h_k V_k
    h = rng.standard_normal(nv)
    y = V[nv].clone()
    _native.call("fp3d_multi_axpy", n, nv, ptr(V), V.stride(0), ptr(to_dev(h)), -1.0,
                 ptr(y), stream_handle())
    ref = w - h @ Vh[:nv]
    assert np.allclose(to_host(y), ref, rtol=0, atol=1e-12 * (1 + np.abs(Vh).max() * nv))


def test_gmres_f32_basis_matches_f64(gpu):
    """GMRES with a float32 Krylov basis (float64 arithmetic) reaches the same
    solution as the float64 basis on a nonsymmetric sparse system, to the
    tolerance asked, and the f32-basis kernels agree with float64 sums of the
    rounded basis."""
    import scipy.sparse as sp

    from up3d_amd import _native
    from up3d_amd.device import ptr, stream_handle, to_dev, to_host
    from up3d_amd.solve.amg import DeviceCSR
    from up3d_amd.solve.linear import _Red, gmres_dev

    rng = np.random.default_rng(7)
    n = 20000
    A = sp.random(n, n, density=5e-4, random_state=1, format="csr") + sp.eye(n) * 4.0
    A = (A + sp.diags(rng.uniform(-0.5, 0.5, n - 1), 1)).tocsr()
    D = DeviceCSR(A)
    b = to_dev(rng.standard_normal(n))
    xs = {}
    for basis in ("f64", "f32"):
        x, info, it = gmres_dev(D.matvec, b, None, rtol=1e-11, restart=40,
                                maxiter=50, basis=basis)
        assert info == 0
        xs[basis] = to_host(x)
        r = to_host(b) - A @ xs[basis]
        assert np.linalg.norm(r) <= 1e-11 * np.linalg.norm(to_host(b)) * 1.0001
    assert np.abs(xs["f32"] - xs["f64"]).max() <= 1e-9 * np.abs(xs["f64"]).max()
    # kernel check: float32 basis, float64 sums
    nv = 9
    import torch

    V = torch.randn(nv, n, dtype=torch.float32, device="cuda")
    w = to_dev(rng.standard_normal(n))
    red = _Red(nv + 1)
    _native.call("fp3d_multi_dot_f32v", n, nv, ptr(V), V.stride(0), ptr(w),
                 ptr(red.work), ptr(red.out), stream_handle())
    Vh = to_host(V).astype(np.float64)
    ref = Vh @ to_host(w)
    assert np.allclose(to_host(red.out[:nv]), ref, rtol=0, atol=1e-11 * np.sqrt(n))
    h = rng.standard_normal(nv)
    y = w.clone()
    _native.call("fp3d_multi_axpy_f32v", n, nv, ptr(V), V.stride(0), ptr(to_dev(h)),
                 -1.0, ptr(y), stream_handle())
    assert np.allclose(to_host(y), to_host(w) - h @ Vh, rtol=0, atol=1e-12 * nv * 10)

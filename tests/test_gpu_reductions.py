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

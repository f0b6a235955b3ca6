"""The allocation-free multi-dot (fp3d_multi_dot_norm_noalloc) against the
workspace entry point it mirrors (fp3d_multi_dot_norm), through the C-ABI:
bitwise equal outputs on two calls in a row, the second with a smaller
input (fewer rows and fewer basis vectors), so the library-owned partial
sums left by the first call cannot leak into the second."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(n, nv, seed):
    import torch

    from up3d_amd import _native
    from up3d_amd.device import ptr, stream_handle, to_dev, to_host

    rng = np.random.default_rng(seed)
    V = to_dev(rng.standard_normal((nv + 1, n)))     # rows 0..nv-1 basis, row nv = w
    w = V[nv]
    work = torch.empty(_native.lib().fp3d_multi_dot_work_doubles(nv + 1),
                       dtype=torch.float64, device=V.device)
    ref = torch.full((nv + 1,), np.nan, dtype=torch.float64, device=V.device)
    got = torch.full((nv + 1,), np.nan, dtype=torch.float64, device=V.device)
    _native.call("fp3d_multi_dot_norm", n, nv, ptr(V), V.stride(0), ptr(w),
                 ptr(work), ptr(ref), stream_handle())
    _native.call("fp3d_multi_dot_norm_noalloc", n, nv, ptr(V), V.stride(0), ptr(w),
                 ptr(got), stream_handle())
    return to_host(ref), to_host(got)


def test_noalloc_multi_dot_matches_workspace_entry_point(gpu):
    for n, nv, seed in ((815_425, 101, 1), (70_001, 17, 2)):
        print(f"[noalloc] multi_dot_norm n={n} nv={nv}", flush=True)
        ref, got = _pair(n, nv, seed)
        assert np.all(np.isfinite(ref))
        assert np.array_equal(ref, got), (n, nv, np.abs(ref - got).max())


def test_noalloc_multi_dot_rejects_oversize(gpu):
    from up3d_amd import _native

    L = _native.lib()
    P = 1 << 20      # never dereferenced: the size check runs first
    assert L.fp3d_multi_dot_norm_noalloc(10, 257, P, 10, P, P, None) == _native.FP3D_EINVAL
    assert "256" in _native.last_error()
    assert L.fp3d_multi_dot_norm_noalloc(10, 0, P, 10, P, P, None) == _native.FP3D_EINVAL

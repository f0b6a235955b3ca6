"""Device plumbing: HBM buffers as torch tensors, raw pointers and the HIP
stream for the C-ABI.  torch only allocates and owns memory here; every
computation goes through libfp3d_hip.so.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _native


def require_gpu() -> torch.device:
    """The HIP device the kernels run on; raises if there is none."""
    _native.lib()
    if not torch.cuda.is_available():
        raise RuntimeError(
            "up3d_amd needs a HIP device (MI355X, gfx950): torch.cuda is not "
            "available on this machine")
    return torch.device("cuda", torch.cuda.current_device())


# the current stream's raw handle without building a torch Stream object
# (every C-ABI launch asks for it: torch.cuda.current_stream() costs several
# microseconds per call, this accessor well under one)
_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)
_cur_device = getattr(torch._C, "_cuda_getDevice", None)


def stream_handle() -> int:
    """HIP stream the C-ABI launches go to: torch's current stream of the
    current device (the one `with torch.cuda.stream(...)` selects)."""
    if _raw_stream is not None and _cur_device is not None:
        return int(_raw_stream(_cur_device()))
    return int(torch.cuda.current_stream().cuda_stream)


def ptr(t) -> int:
    """Device address of a tensor (None -> NULL)."""
    if t is None:
        return None
    return int(t.data_ptr())


def to_dev(a, dtype=None) -> torch.Tensor:
    """Copy a numpy array (or tensor) into a contiguous device tensor."""
    dev = require_gpu()
    if isinstance(a, torch.Tensor):
        t = a.to(dev)
        if dtype is not None:
            t = t.to(dtype)
        return t.contiguous()
    arr = np.ascontiguousarray(a)
    t = torch.from_numpy(arr)
    if dtype is not None:
        t = t.to(dtype)
    return t.to(dev).contiguous()


def empty(shape, dtype=torch.float64) -> torch.Tensor:
    return torch.empty(shape, dtype=dtype, device=require_gpu())


def zeros(shape, dtype=torch.float64) -> torch.Tensor:
    return torch.zeros(shape, dtype=dtype, device=require_gpu())


def to_host(t: torch.Tensor) -> np.ndarray:
    return t.detach().cpu().numpy()


def synchronize() -> None:
    torch.cuda.synchronize()

"""Argument plumbing for the drop-in kernel functions.

The reference kernels take numpy arrays and overwrite preallocated
outputs in place.  Here an argument may be a numpy array (copied to HBM,
and for outputs copied back into the caller's array after the launch) or
a device tensor (used in place when its dtype already matches).
Connectivity is canonicalised to int32 on the device.
"""

from __future__ import annotations

import numpy as np
import torch

from ..device import require_gpu, to_dev

_NP2T = {np.float64: torch.float64, np.float32: torch.float32,
         np.int32: torch.int32, np.int64: torch.int64, np.int8: torch.int8,
         np.uint8: torch.uint8, np.bool_: torch.bool}


def din(a, dtype: torch.dtype) -> torch.Tensor:
    """Input on the device with the given dtype (no copy when possible)."""
    if isinstance(a, torch.Tensor) and a.is_cuda and a.dtype == dtype \
            and a.is_contiguous():
        return a
    return to_dev(a, dtype)


class Out:
    """An output buffer: a device tensor to launch into, written back to the
    caller's array by .done()."""

    def __init__(self, a, dtype: torch.dtype):
        self.a = a
        if isinstance(a, torch.Tensor) and a.is_cuda and a.dtype == dtype \
                and a.is_contiguous():
            self.t = a
            self._wb = None
        else:
            shape = tuple(a.shape)
            self.t = torch.empty(shape, dtype=dtype, device=require_gpu())
            if isinstance(a, torch.Tensor):
                self.t.copy_(a.to(self.t.device, dtype))
            else:
                self.t.copy_(torch.from_numpy(np.ascontiguousarray(a)).to(dtype))
            self._wb = a

    def done(self):
        if self._wb is None:
            return self.a
        if isinstance(self._wb, torch.Tensor):
            self._wb.copy_(self.t.to(self._wb.device, self._wb.dtype))
        else:
            host = self.t.cpu().numpy()
            self._wb[...] = host.astype(self._wb.dtype, copy=False)
        return self.a

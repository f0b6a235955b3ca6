"""Project-written autotuner (no reference counterpart): picks the
element-kernel block size, the matrix gather's staged chunk and the
Krylov / AMG parameters of a mesh by timing them on the device, and caches
the choice per mesh shape."""

from .autotune import (TuningCache, apply_element_block, apply_gather_chunk,
                       tune_element_block, tune_gather_chunk, tune_linear_solver,
                       tune_newton_workspace, tuned_newton_kwargs)

__all__ = ["TuningCache", "apply_element_block", "apply_gather_chunk",
           "tune_element_block", "tune_gather_chunk", "tune_linear_solver",
           "tune_newton_workspace", "tuned_newton_kwargs"]

"""Project-written autotuner (no reference counterpart): picks the
element-kernel block size, the matrix gather's staged chunk, the upstream
walks per lane and the Krylov / AMG parameters of a mesh by timing them on
the device, and caches the choice per mesh shape."""

from .autotune import (TuningCache, apply_element_block, apply_gather_chunk,
                       apply_walk_width, tune_element_block, tune_gather_chunk,
                       tune_linear_solver, tune_newton_workspace, tune_walk_width,
                       tuned_newton_kwargs)

__all__ = ["TuningCache", "apply_element_block", "apply_gather_chunk",
           "apply_walk_width", "tune_element_block", "tune_gather_chunk",
           "tune_linear_solver", "tune_newton_workspace", "tune_walk_width",
           "tuned_newton_kwargs"]

"""ctypes binding of libfp3d_hip.so (the C-ABI declared in include/fp3d.h).

The library is loaded from up3d_amd/lib/ (built in-tree by
up3d_amd.csrc.build).  There is no fallback: if the library is missing the
import of any compute entry point raises, and on a machine without a HIP
device every device call fails with a RuntimeError naming the problem.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libfp3d_hip.so"

FP3D_OK = 0
FP3D_EINVAL = -1
FP3D_EHIP = -2
FP3D_ELIMIT = -3

_vp = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int32
_int = C.c_int
_dbl = C.c_double
_pi64 = C.POINTER(C.c_int64)
_pi32 = C.POINTER(C.c_int32)

# name -> (restype, argtypes); every pointer argument is passed as void*
_SIGS = {
    "fp3d_last_error": (C.c_char_p, []),
    "fp3d_version": (_int, []),
    "fp3d_device_count": (_int, [C.POINTER(_int)]),
    "fp3d_stream_sync": (_int, [_vp]),
    "fp3d_density_field": (_int, [_i64, _vp, _dbl, _dbl, _vp, _vp]),
    "fp3d_mach_squared_field": (_int, [_i64, _vp, _dbl, _dbl, _vp, _vp]),
    "fp3d_limit_q2_field": (_int, [_i64, _vp, _dbl, _vp, _vp]),
    "fp3d_density_derivative_field": (_int, [_i64, _vp, _dbl, _dbl, _vp, _vp]),
    "fp3d_element_velocity_q2": (_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fp3d_element_state": (_int, [_i64, _vp, _vp, _vp, _dbl, _dbl, _dbl, _dbl,
                                  _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fp3d_upstream_elements": (_int, [_i64, _vp, _vp, _vp, _vp, _vp, _int, _vp,
                                      _vp]),
    "fp3d_upstream_elements_d": (_int, [_i64, _vp, _vp, _vp, _vp, _vp, _int,
                                        _vp, _vp]),
    "fp3d_upwind_rho_tilde_d": (_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                       _int, _dbl, _dbl, _dbl, _dbl, _dbl, _vp,
                                       _vp, _vp, _vp]),
    "fp3d_upwind_rho_tilde_sens_d": (_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                            _int, _dbl, _dbl, _dbl, _dbl, _dbl, _vp,
                                            _vp, _vp, _vp, _vp, _vp]),
    "fp3d_rho_tilde_sweep": (_int, [_i64, _vp, _vp, _vp, _dbl, _dbl, _dbl, _dbl,
                                    _dbl, _vp, _vp, _vp]),
    "fp3d_upwind_rho_tilde": (_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                     _int, _dbl, _dbl, _dbl, _dbl, _dbl, _vp,
                                     _vp, _vp, _vp]),
    "fp3d_classify_upwind_branches": (_int, [_i64, _vp, _vp, _vp, _dbl, _dbl,
                                             _dbl, _dbl, _dbl, _vp, _vp]),
    "fp3d_rho_tilde_frozen_sweep": (_int, [_i64, _vp, _vp, _vp, _vp, _dbl, _dbl,
                                           _dbl, _dbl, _dbl, _vp, _vp, _vp]),
    "fp3d_rho_tilde_frozen_sensitivities": (_int, [_i64, _vp, _vp, _vp, _vp,
                                                   _dbl, _dbl, _dbl, _dbl, _dbl,
                                                   _vp, _vp, _vp]),
    "fp3d_rho_tilde_sensitivities": (_int, [_i64, _vp, _vp, _vp, _dbl, _dbl,
                                            _dbl, _dbl, _dbl, _vp, _vp, _vp]),
    "fp3d_upwind_monitors": (_int, [_i64, _vp, _vp, _dbl, _vp, _vp, _vp]),
    "fp3d_count_limited": (_int, [_i64, _vp, _vp, _vp, _vp]),
    "fp3d_reduce_work_doubles": (_int, []),
    "fp3d_assemble_residual_colored": (_int, [_int, _pi64, _vp, _vp, _vp, _vp,
                                              _vp, _vp, _i64, _vp, _vp]),
    "fp3d_assemble_matrix_colored": (_int, [_int, _pi64, _vp, _vp, _vp, _vp,
                                            _vp, _vp, _vp]),
    "fp3d_assemble_jacobian_colored": (_int, [_int, _pi64, _vp, _vp, _vp, _vp,
                                              _vp, _vp, _vp, _vp, _dbl, _vp,
                                              _vp]),
    "fp3d_assemble_jacobian_spd_colored": (_int, [_int, _pi64, _vp, _vp, _vp,
                                                  _vp, _vp, _vp, _vp, _vp, _dbl,
                                                  _dbl, _vp, _vp]),
    "fp3d_fill_term3_coo": (_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                   _dbl, _vp, _vp, _vp, _vp]),
    "fp3d_term3_workspace_bytes": (_i64, [_i64, _i64]),
    "fp3d_term3_build": (_int, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                _vp, _vp, _vp, _vp, _vp, _dbl, _vp, _pi64, _vp]),
    "fp3d_term3_active_count": (_int, [_vp, _i64, _i64, _pi64, _vp]),
    "fp3d_csr_spmv": (_int, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fp3d_csr_spmv_f32": (_int, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fp3d_csr_spmv_term3": (_int, [_i64, _i64, _vp, _vp, _vp, _int, _vp, _vp,
                                   _vp, _i64, _i64, _vp, _vp, _vp, _vp]),
    "fp3d_csr_mask_dirichlet": (_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fp3d_csr_to_f32": (_int, [_i64, _vp, _vp, _vp]),
    "fp3d_gather_rows_per_block": (_int, []),
    "fp3d_element_residual_terms": (_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                           _vp]),
    "fp3d_gather_node_terms": (_int, [_i64, _vp, _vp, _vp, _vp, _vp]),
    "fp3d_element_matrix_terms": (_int, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _dbl, _dbl,
                                         _vp, _vp, _vp]),
    "fp3d_gather_csr_terms": (_int, [_i64, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fp3d_multi_dot_reorth": (_int, [_i64, _int, _vp, _i64, _vp, _vp, _dbl, _vp, _vp, _vp]),
    "fp3d_lumped_volume": (_int, [_i64, _vp, _vp, _vp, _vp, _vp]),
    "fp3d_csr_diag_shift": (_int, [_i64, _vp, _vp, _dbl, _vp, _vp, _vp]),
    "fp3d_fill": (_int, [_i64, _dbl, _vp, _vp]),
    "fp3d_axpby": (_int, [_i64, _dbl, _vp, _dbl, _vp, _vp]),
    "fp3d_scale": (_int, [_i64, _dbl, _vp, _vp, _vp]),
    "fp3d_mask_copy": (_int, [_i64, _vp, _vp, _vp, _vp]),
    "fp3d_multi_dot": (_int, [_i64, _int, _vp, _i64, _vp, _vp, _vp, _vp]),
    "fp3d_multi_dot_work_doubles": (_int, [_int]),
    "fp3d_multi_axpy": (_int, [_i64, _int, _vp, _i64, _vp, _dbl, _vp, _vp]),
    "fp3d_multi_dot_f32v": (_int, [_i64, _int, _vp, _i64, _vp, _vp, _vp, _vp]),
    "fp3d_multi_dot_norm": (_int, [_i64, _int, _vp, _i64, _vp, _vp, _vp, _vp]),
    "fp3d_multi_dot_norm_f32v": (_int, [_i64, _int, _vp, _i64, _vp, _vp, _vp, _vp]),
    "fp3d_multi_dot_norm_noalloc": (_int, [_i64, _int, _vp, _i64, _vp, _vp, _vp]),
    "fp3d_batch_job_bytes": (_int, []),
    "fp3d_batch_eval_job_bytes": (_int, []),
    "fp3d_batch_residual_eval": (_int, [_int, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp,
                                        _vp, _vp, _vp, _dbl, _dbl, _dbl, _dbl, _dbl, _int,
                                        _vp]),
    "fp3d_batch_jacobian_matvec": (_int, [_int, _vp, _i64, _i64, _i64, _vp, _vp, _int,
                                          _vp, _i64, _int, _vp]),
    "fp3d_batch_orthonormalise": (_int, [_int, _vp, _int, _i64, _int, _dbl, _vp]),
    "fp3d_multi_dot_reorth_f32v": (_int, [_i64, _int, _vp, _i64, _vp, _vp, _dbl, _vp, _vp, _vp]),
    "fp3d_multi_axpy_f32v": (_int, [_i64, _int, _vp, _i64, _vp, _dbl, _vp, _vp]),
    "fp3d_scale_to_f32": (_int, [_i64, _dbl, _vp, _vp, _vp]),
    "fp3d_f32_to_f64": (_int, [_i64, _vp, _vp, _vp]),
    "fp3d_scale_inv_dev": (_int, [_i64, _vp, _vp, _vp, _vp]),
    "fp3d_set_element_block": (_int, [_int]),
    "fp3d_get_element_block": (_int, []),
    "fp3d_set_gather_chunk": (_int, [_int]),
    "fp3d_get_gather_chunk": (_int, []),
    "fp3d_set_walk_width": (_int, [_int]),
    "fp3d_get_walk_width": (_int, []),
    "fp3d_scale_inv_dev_to_f32": (_int, [_i64, _vp, _vp, _vp, _vp]),
    "fp3d_norms": (_int, [_i64, _vp, _vp, _vp, _vp]),
    "fp3d_amg_cheb_smooth": (_int, [_i64, _i64, _vp, _vp, _vp, _int, _vp, _vp,
                                    _vp, _vp, _vp, _vp, _int, _dbl, _dbl, _int,
                                    _vp]),
    "fp3d_amg_residual": (_int, [_i64, _i64, _vp, _vp, _vp, _int, _vp, _vp, _vp,
                                 _vp]),
    "fp3d_amg_restrict": (_int, [_i64, _i64, _vp, _vp, _vp, _int, _vp, _vp,
                                 _vp]),
    "fp3d_amg_restrict_start": (_int, [_i64, _i64, _vp, _vp, _vp, _int, _vp, _vp,
                                       _vp, _int, _dbl, _dbl, _vp, _vp, _vp, _vp]),
    "fp3d_amg_prolong_add": (_int, [_i64, _i64, _vp, _vp, _vp, _int, _vp, _vp,
                                    _vp]),
    "fp3d_dense_matvec": (_int, [_i64, _vp, _vp, _vp, _vp]),
    "fp3d_dense_matvec_v": (_int, [_int, _i64, _vp, _vp, _vp, _vp]),
    "fp3d_grp_vcycle_op": (_int, [_int, _vp, _int, _i64, _int, _int, _vp]),
    "fp3d_grp_vcycle_run": (_int, [_int, _vp, _int, _vp, _vp]),
    "fp3d_drho_tilde_dmach": (_int, [_i64, _vp, _vp, _vp, _vp, _dbl, _dbl, _dbl,
                                     _dbl, _vp, _vp]),
    "fp3d_adjoint_dx": (_int, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                               _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fp3d_adjoint_dx_work_doubles": (_i64, [_i64]),
    "fp3d_sheet_fill": (_int, [_i64] + [_vp] * 17 + [_vp]),
    "fp3d_sheet_t3_prepare": (_int, [_i64] + [_vp] * 18 + [_dbl] + [_vp] * 4 + [_vp]),
    "fp3d_sheet_t3_apply": (_int, [_i64, _i64] + [_vp] * 8 + [_vp]),
    "fp3d_host_greedy_coloring": (_int, [_i64, _vp, _i64, _vp, _pi32]),
    "fp3d_host_csr_pattern": (_int, [_i64, _vp, _i64, _vp, _vp, _pi64]),
    "fp3d_host_elem_to_csr": (_int, [_i64, _vp, _vp, _vp, _vp]),
    "fp3d_host_spgemm": (_int, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                _pi64]),
    "fp3d_host_spgemm_fill": (_int, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                     _vp, _vp]),
    "fp3d_host_csr_matvec": (_int, [_i64, _vp, _vp, _vp, _vp, _vp]),
    "fp3d_host_sa_prolongator": (_int, [_i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp,
                                        _vp]),
    "fp3d_host_sa_prolongator_fill": (_int, [_i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp,
                                             _vp, _vp, _vp]),
    "fp3d_host_csr_transpose": (_int, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "fp3d_host_apply_givens": (_int, [_int, _vp, _vp]),
    "fp3d_graph_launch": (_int, [_vp, _vp]),
    "fp3d_arnoldi_step": (_int, [_vp, _int, _vp]),
    "fp3d_host_face_adjacency": (_int, [_i64, _vp, _vp]),
    "fp3d_host_element_geometry": (_int, [_i64, _vp, _vp, _vp, _vp, _pi64]),
    "fp3d_host_locate_points": (_int, [_i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp]),
    "fp3d_host_node_min_edge_length": (_int, [_i64, _vp, _i64, _vp, _vp]),
    "fp3d_host_standard_aggregation": (_int, [_i64, _vp, _vp, _vp, _vp, _pi64]),
}

EXPORTED = tuple(_SIGS)
_LIB = None


class NativeMissing(RuntimeError):
    pass


def lib():
    """The loaded library; raises NativeMissing when it has not been built."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise NativeMissing(
                f"{LIB_PATH} is missing: build it with "
                "`python -m up3d_amd.csrc.build` (hipcc, gfx950)")
        # torch ships its own libamdhip64.so.7 (same soname as the system
        # one the library links): load torch first so the process holds one
        # HIP runtime, the one torch's streams and allocations belong to.
        # Loaded the other way round, torch binds to the system runtime and
        # the first launch fails with "no ROCm-capable device".
        import torch  # noqa: F401
        h = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = h
    return _LIB


def last_error() -> str:
    msg = lib().fp3d_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str = "") -> None:
    """Map a C-ABI status to the reference's exception types."""
    if rc == FP3D_OK:
        return
    msg = last_error()
    text = f"{what}: {msg}" if what else msg
    if rc in (FP3D_EINVAL, FP3D_ELIMIT):
        raise ValueError(text)
    raise RuntimeError(text)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)


def device_count() -> int:
    c = _int(0)
    check(lib().fp3d_device_count(C.byref(c)))
    return int(c.value)

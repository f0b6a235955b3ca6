"""C-ABI contract checks that need no GPU: every entry point validates its
arguments before touching the device, so zero sizes return FP3D_OK and bad
arguments return FP3D_EINVAL with a message in fp3d_last_error(), without a
HIP call.  The pointers below are never dereferenced (fake, aligned
addresses); only the validation paths run."""

import pytest

from up3d_amd import _native

OK, EINVAL = 0, -1
P = 1 << 20          # a 256-byte-aligned fake device address
P_ODD = P + 8        # 8-byte but not 16/32-byte aligned


def _lib():
    return _native.lib()


def _err() -> str:
    return _native.lib().fp3d_last_error().decode()


def test_zero_sizes_are_no_ops():
    L = _lib()
    assert L.fp3d_element_state(0, None, None, None, 1.0, 0.8, 1.4, 10.0, None, None,
                                None, None, None, None, None) == OK
    assert L.fp3d_element_residual_terms(0, None, None, None, None, None, None, None,
                                         None) == OK
    assert L.fp3d_gather_node_terms(0, None, None, None, None, None) == OK
    assert L.fp3d_element_matrix_terms(0, None, None, None, None, None, None, 1.0, 0.0,
                                       None, None, None) == OK
    assert L.fp3d_gather_csr_terms(0, 1, None, None, None, None, None, None, None) == OK
    assert L.fp3d_upwind_rho_tilde_sens_d(0, None, None, None, None, None, None, None, 4,
                                          0.8, 0.95, 1.5, 1.4, 0.05, None, None, None,
                                          None, None, None) == OK


@pytest.mark.parametrize("case", [
    "residual_terms_null", "residual_terms_misaligned", "gather_null",
    "matrix_terms_clamp", "csr_gather_max_seg", "csr_gather_rel8_align",
    "state_u2", "walk_sens_null", "term3_empty", "term3_pos_align",
    "term3_gdotb_align"])
def test_bad_arguments_rejected_before_launch(case):
    L = _lib()
    if case == "residual_terms_null":
        rc = L.fp3d_element_residual_terms(10, None, P, P, None, None, P, P, None)
        msg = "NULL"
    elif case == "residual_terms_misaligned":
        rc = L.fp3d_element_residual_terms(10, P, P, P, P, None, P, P_ODD, None)
        msg = "aligned"
    elif case == "gather_null":
        rc = L.fp3d_gather_node_terms(10, None, None, P, P, None)
        msg = "NULL"
    elif case == "matrix_terms_clamp":
        rc = L.fp3d_element_matrix_terms(10, P, P, P, P, P, None, 1.0, 1.5, P, P, None)
        msg = "clamp"
    elif case == "csr_gather_max_seg":
        rc = L.fp3d_gather_csr_terms(10, 0, P, P, None, P, P, P, None)
        msg = "max_seg"
    elif case == "csr_gather_rel8_align":
        rc = L.fp3d_gather_csr_terms(10, 64, P, P, None, P + 2, P, P, None)
        msg = "rel8"
    elif case == "state_u2":
        rc = L.fp3d_element_state(10, P, P, P, 0.0, 0.8, 1.4, 10.0, P, P, P, P, None,
                                  None, None)
        msg = "u_inf"
    elif case == "walk_sens_null":
        rc = L.fp3d_upwind_rho_tilde_sens_d(10, P, P, P, P, P, P, P, 4, 0.8, 0.95, 1.5,
                                            1.4, 0.05, None, P, P, None, P, None)
        msg = "NULL"
    elif case == "term3_empty":
        rc = L.fp3d_term3_build(0, 10, P, P, P, P, P, P, P, None, P, P, None, None, 1.0,
                                P, None, None)
        msg = "empty mesh"
    elif case == "term3_pos_align":
        rc = L.fp3d_term3_build(10, 10, P, P, P, P_ODD, P, P, P, None, P, P, None, None,
                                1.0, P, None, None)
        msg = "inc_pos"
    else:
        rc = L.fp3d_term3_build(10, 10, P, P, P, P, P, P, P, P_ODD, P, P, None, None, 1.0,
                                P, None, None)
        msg = "alignment"
    assert rc == EINVAL
    assert msg in _err()


def test_check_maps_einval_to_value_error():
    L = _lib()
    with pytest.raises(ValueError, match="max_seg"):
        _native.check(L.fp3d_gather_csr_terms(10, 0, P, P, None, P, P, P, None),
                      "gather_csr_terms")


def test_tuning_setters_validate():
    L = _lib()
    prev = int(L.fp3d_get_walk_width())
    for bad in (0, 5):
        assert L.fp3d_set_walk_width(bad) == EINVAL and "walk_width" in _err()
    for w in (1, 2, 3, 4):
        assert L.fp3d_set_walk_width(w) == OK and L.fp3d_get_walk_width() == w
    assert L.fp3d_set_walk_width(prev) == OK

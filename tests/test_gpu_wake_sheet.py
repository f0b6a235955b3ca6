"""Device assembly of the level-set sheet terms (up3d_amd/wake/sheet.py)
against the host restatement it replaces, on the reference's NACA0012 coarse
mesh at a transonic state with per-side upwinding:

  * the Picard sheet terms C (both copies of each cut element, the fan's
    lower copy, the wake least-squares rows, the TE pressure-Kutta rows)
    filled into their fixed pattern equal MultivaluedOperator.correction,
    with and without the Dirichlet mask, for both Kutta forms;
  * the Newton Jacobian's sheet part (D_keep C, exact Kutta Jacobian,
    Term 2 of the special copies) plus the special Term-3 operator equal
    the host COO assembly of the same terms (entries and matvec)."""

import numpy as np
import pytest
import scipy.sparse as sp

from tests.test_gpu_wake_ls import _mesh, _setup

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def state(gpu):
    from up3d_amd.solve.newton_ls import LSNewtonSystem
    from up3d_amd.solve.picard_ls import _farfield_main, solve_multivalued_lifting

    mesh = _mesh("coarse")
    _, mv = _setup(mesh)
    print(f"[sheet] NACA coarse n_tets={len(mesh.elements)} n_aux={mv.n_ext}", flush=True)
    seed = solve_multivalued_lifting(mv, mesh, 0.70, alpha_deg=1.25, upwind_c=1.5,
                                     damping_theta=0.2, n_outer_max=30)
    phi = seed["phi_ext"].copy()
    ff, vals = _farfield_main(mesh.nodes, mesh.boundary_faces["farfield"], 1.25,
                              seed["gamma"], 1.0, (0.25, 0.0), np.sqrt(1 - 0.49))
    phi[ff] = vals
    # a rough potential on the sheet's nodes (main and aux DOFs) makes cut
    # and fan elements supersonic, so their Term 2 / Term 3 are active
    rng = np.random.default_rng(11)
    sheet = np.unique(np.concatenate([mv.el[mv.special].ravel(), mv.cut_aux]))
    sheet = sheet[~np.isin(sheet, ff)]
    phi[sheet] += 0.02 * rng.standard_normal(len(sheet))
    sysm = LSNewtonSystem(mv, 0.70)
    R, sides, C, w = sysm.residual(phi, None, sens=True)
    assert mv.n_nu_active > 0
    return dict(mesh=mesh, mv=mv, phi=phi, ff=ff, sysm=sysm, sides=sides, w=w)


def _close(A, B, tol=1e-13):
    A, B = sp.csr_matrix(A), sp.csr_matrix(B)
    d = abs(A - B)
    scale = max(abs(B).max(), 1e-300)
    return d.max() <= tol * scale if d.nnz else True


@pytest.mark.parametrize("te_kutta", ["pressure", "mass"])
def test_sheet_terms_equal_host_correction(state, te_kutta):
    import torch

    from up3d_amd.device import to_dev

    mv, phi, sides = state["mv"], state["phi"], state["sides"]
    special = (sides[0]["rho_tilde"], sides[1]["rho_tilde"])
    S = mv.sheet_fill(special, te_kutta, phi, user=f"test_{te_kutta}")
    C = mv.correction(special, "wake_ls", te_kutta, phi)
    print(f"[sheet] {te_kutta}: {S.n_slots} slots from {S.n_contrib} contributions, "
          f"host nnz {C.nnz}", flush=True)
    assert _close(S.to_scipy(), C)
    fm = np.ones(mv.n_total, dtype=np.uint8)
    fm[state["ff"]] = 0
    S = mv.sheet_fill(special, te_kutta, phi, fmask=to_dev(fm, torch.uint8),
                      user=f"test_{te_kutta}")
    Dm = sp.diags(fm.astype(np.float64))
    assert _close(S.to_scipy(), Dm @ C @ Dm)
    # the host density triple gives the same values as the device pair
    trip = mv.special_weights_host(special)
    S2 = mv.sheet_fill(trip, te_kutta, phi, user=f"test2_{te_kutta}")
    S = mv.sheet_fill(special, te_kutta, phi, user=f"test_{te_kutta}")
    assert np.array_equal(S2.to_scipy().toarray(), S.to_scipy().toarray())


def test_jacobian_sheet_and_special_term3_equal_host_assembly(state):
    from up3d_amd.device import empty, to_dev, to_host

    mv, phi, sides, sysm = state["mv"], state["phi"], state["sides"], state["sysm"]
    nm, nT = sysm.nm, sysm.nT
    sysm.linearize(phi, sides, None, state["w"], state["ff"])
    fm = np.ones(nT)
    fm[state["ff"]] = 0.0
    # the host assembly of the same terms (as before the device port)
    C = mv.correction((sides[0]["rho_tilde"], sides[1]["rho_tilde"]), "wake_ls",
                      "pressure", phi)
    keep = np.ones(nT)
    keep[sysm.te_aux] = 0.0
    Ck = (sp.diags(keep) @ C).tocoo()
    kr, kc, kd = sysm._kutta_jacobian(phi)
    tr, tc, td = sysm._terms23_host(sides)
    Jh = sp.coo_matrix((np.concatenate([Ck.data, kd, td]),
                        (np.concatenate([Ck.row, kr, tr]),
                         np.concatenate([Ck.col, kc, tc]))), shape=(nT, nT)).tocsr()
    Dm = sp.diags(fm)
    Jh = (Dm @ Jh @ Dm).tocsr()
    n_t3 = sysm._t3s.n_entries
    print(f"[sheet] Jacobian sheet {sysm._jsheet.n_slots} slots, special Term 3 "
          f"{sysm._t3s.n_cand} candidates / {n_t3} row entries", flush=True)
    assert n_t3 > 0
    assert _close(sysm.Jc_host, Jh)
    # operator: plain part + sheet part + special Term 3 == plain part + host
    rng = np.random.default_rng(3)
    x = rng.standard_normal(nT)
    xd = to_dev(x)
    y = to_host(sysm.matvec(xd, empty(nT)))
    yp = empty(nT)
    sysm.nop.matvec(xd[:nm], yp[:nm])
    ref = np.zeros(nT)
    ref[:nm] = to_host(yp[:nm])
    ref += Jh @ x
    assert np.abs(y - ref).max() <= 1e-13 * np.abs(ref).max()
    # the aux block the LU factors is the host block
    A_aux = sysm._jsheet.block_csc(nm, nm)
    assert _close(A_aux, Jh[nm:, nm:])

"""Grouped V-cycles (up3d_amd/solve/amg.py run_grouped /
apply_preconditioners, fp3d_grp_vcycle_op) against each hierarchy's own
V-cycle: several SA hierarchies of different Picard matrices on one channel
(same pattern, so the same level sizes), one launch per V-cycle operation
for all of them, must give every solve the bits of its own V-cycle -- in
float64 and float32, for Chebyshev degrees 1 to 3, and again on a second
round (the records are rebuilt from the cache, the launches re-read them)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hier_inputs(gpu):
    import torch

    from up3d_amd.device import to_dev
    from up3d_amd.kernels.jacobian import PicardOperator
    from up3d_amd.mesh.channel import bump_channel
    from up3d_amd.solve.common import csr_to_host

    m = bump_channel(24, 8, 6, jitter=0.15, seed=11, rcm=True)
    op = PicardOperator(m.nodes, m.elements)
    rng = np.random.default_rng(7)
    mats = []
    for _ in range(3):
        rt = to_dev(0.5 + rng.random(m.n_tets))
        A = torch.empty(op.store.nnz, dtype=torch.float64, device=gpu)
        op.matrix_dev(rt, A)
        mats.append(csr_to_host(op.store, A))
    print(f"[grouped] channel 24x8x6 n_tets={m.n_tets} n_nodes={m.n_nodes}", flush=True)
    return mats


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("degree", [1, 2, 3])
def test_grouped_vcycles_equal_own_vcycles(hier_inputs, precision, degree):
    import torch

    from up3d_amd.device import to_dev
    from up3d_amd.solve.amg import AMGHierarchy, apply_preconditioners, run_grouped

    amgs = [AMGHierarchy(A, precision=precision, degree=degree, max_coarse=60, seed=0)
            for A in hier_inputs]
    sigs = {tuple(a.shapes) for a in amgs}
    assert len(sigs) == 1, f"level sizes differ: {sigs}"
    n = amgs[0].levels[0].n
    rng = np.random.default_rng(degree)
    bs = [to_dev(rng.standard_normal(n)) for _ in amgs]
    solo = [torch.empty_like(b) for b in bs]
    for a, b, x in zip(amgs, bs, solo):
        a.apply_eager(b, x)
    for rnd in range(2):
        print(f"[grouped] {precision} degree {degree} round {rnd}", flush=True)
        xs = [torch.full_like(b, float("nan")) for b in bs]
        progs = [a.program(b, x) for a, b, x in zip(amgs, bs, xs)]
        assert len({p.signature for p in progs}) == 1
        run_grouped(progs)
        torch.cuda.synchronize()
        for k in range(len(amgs)):
            assert torch.equal(xs[k], solo[k]), f"solve {k} differs (round {rnd})"
    # the executor's entry: a hierarchy used twice in one round runs in two
    # groups, a lone one by itself, all with their own bits
    xs = [torch.full_like(b, float("nan")) for b in bs]
    x_dup = torch.full_like(bs[0], float("nan"))
    apply_preconditioners([(amgs[0].apply, bs[0], xs[0]), (amgs[1].apply, bs[1], xs[1]),
                           (amgs[2].apply, bs[2], xs[2]), (amgs[0].apply, bs[0], x_dup)])
    torch.cuda.synchronize()
    for k in range(len(amgs)):
        assert torch.equal(xs[k], solo[k])
    assert torch.equal(x_dup, solo[0])

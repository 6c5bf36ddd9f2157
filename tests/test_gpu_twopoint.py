"""Meson two-point functions on the GPU (contract_twopoint /
contract_twopoint_snk_momentum, cvc_utils.c:3075-3235) against the oracle:
all 16 x 16 source/sink gammas on random fields with and without a sink
momentum (1e-12 relative), and the pion of 12 solved propagators, which is
sum_x |S(x)|^2 per time slice."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
BC = (1.0, 0.0, 0.0, 0.0)


@pytest.mark.parametrize("mom", [None, (1, 2, 0)])
def test_twopoint_matches_oracle(chip, orc, mom):
    dims = (6, 4, 4, 2)
    print(f"two-point mom={mom}", flush=True)
    g = chip.gauge_random(dims, 3, 0.3)
    ctx = chip.Context(dims, 0.15, 0.01, BC, gauge=g)
    L = orc.Lattice(dims, 0.15, 0.01, BC)
    rng = np.random.default_rng(11)
    chi = [rng.standard_normal((L.V, 24)) for _ in range(12)]
    phi = [rng.standard_normal((L.V, 24)) for _ in range(12)]
    src = (2, 1, 3, 1)
    fc, fp = ctx.field(12), ctx.field(12)
    fc.upload(np.stack(chi))
    fp.upload(np.stack(phi))
    for gs in range(16):
        for gk in range(16):
            want = L.contract_twopoint(gs, gk, chi, phi, 3, mom, src)
            got = ctx.field_contract_twopoint(gs, gk, fc, fp, 3, mom, src)
            assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want)), (gs, gk)
    # host-array entry point, n_c = 1
    want = L.contract_twopoint(6, 9, chi[:4], phi[:4], 1, mom, src)
    got = ctx.contract_twopoint(6, 9, chi[:4], phi[:4], 1, mom, src)
    assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want))


def test_pion_of_solved_propagators(chip, orc):
    dims = (8, 4, 4, 4)
    g = chip.gauge_random(dims, 5, 0.3)
    ctx = chip.Context(dims, 0.163, 0.01, BC, gauge=g)
    V = ctx.V
    b = np.stack([chip.point_source(V, 0, a // 3, a % 3) for a in range(12)])
    rc, x, st = ctx.invert_Qtm_her(b, 1e-22, 5000)
    cols = list(x)
    c = ctx.contract_twopoint(5, 5, cols, cols, 3)
    want = sum((f * f).sum(1) for f in cols).reshape(dims[0], -1).sum(1)
    assert np.max(np.abs(c - want)) < 1e-12 * want.max()
    assert np.all(np.diff(np.abs(c[: dims[0] // 2 + 1])) < 0)   # decays away from the source

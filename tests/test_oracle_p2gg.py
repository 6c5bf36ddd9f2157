"""Pin of the oracle's sequential source (p2gg_xspace.c:517-548), CPU only.

The restatement applies the reference's gamma tables and phase in its loop
order; here the same source is formed with explicit 4x4 Dirac matrices (the
chiral basis documented at cvc_utils.c:709-745, products in the reference's
index order 0-3 gamma_mu, 4 identity, 5 gamma5, 6-9 gamma_mu gamma5,
10-15 gamma_0 gamma_1, gamma_0 gamma_2, gamma_0 gamma_3, gamma_1 gamma_2,
gamma_1 gamma_3, gamma_2 gamma_3) and a numpy phase, for every gamma, both
phase signs, a wrapped time slice and a non-cubic lattice; a wrong gamma index,
a transposed Gamma, a conjugated phase or a wrong slice fails it.
"""
import numpy as np
import pytest

from test_oracle import G, G5, coords, from_c, to_c

I4 = np.eye(4, dtype=complex)


def gamma16():
    g = [G[0], G[1], G[2], G[3], I4, G5]
    g += [G[m] @ G5 for m in range(4)]
    g += [G[a] @ G[b] for a, b in ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))]
    return g


@pytest.mark.parametrize("sgn", [1.0, -1.0])
def test_sequential_source_brute_force(orc, sgn):
    dims = (6, 4, 2, 6)
    L = orc.Lattice(dims, 0.15, 0.02)
    rng = np.random.default_rng(31)
    prop = rng.standard_normal((L.V, 24))
    c = coords(dims)
    g = gamma16()
    for gid, tseq, mom in [(5, 0, (0, 0, 0)), (0, 5, (1, 0, 0)), (3, 2, (1, 1, 3)), (7, 1, (0, 1, 2)),
                           (11, 4, (2, 1, 5)), (15, 3, (3, 1, 1))] + [(k, 2, (1, 1, 1)) for k in range(16)]:
        got = orc_src = L.seq_source(prop, tseq, gid, mom, sgn)
        ph = sgn * 2 * np.pi * (c[:, 1] * mom[0] / dims[1] + c[:, 2] * mom[1] / dims[2] + c[:, 3] * mom[2] / dims[3])
        want = np.einsum("ab,vbc->vac", g[gid], to_c(prop)) * np.exp(1j * ph)[:, None, None]
        want[c[:, 0] != tseq] = 0
        assert np.abs(got - from_c(want)).max() < 1e-14, (gid, tseq, mom)
        assert np.all(orc_src[c[:, 0] != tseq] == 0)

"""Pins of the oracle's meson two-point function (CPU), restating
contract_twopoint / contract_twopoint_snk_momentum (cvc_utils.c:3075-3235).

1. For every source and sink gamma (16 x 16), random chi and phi and a sink
   momentum, the oracle equals a brute force with explicit 4x4 matrices:
   C(t) = sum_x e^{ip.(x-y)} sum_{mu,c} s_mu chi_{p(mu),c}(x)^dagger gamma5 Gamma_snk phi_{mu,c}(x)
   (times -i for an imaginary source gamma), where Gamma_src has its non-zero
   g_mu at (mu, p(mu)) and s_mu = (g_mu real ? g_mu : -Im g_mu) times the
   gamma5 sign of spin p(mu) -- the permutation-table mechanics are not used.
2. The pion (gamma5, gamma5) with chi = phi is sum_x |phi(x)|^2 per time slice.
"""
import numpy as np
import pytest

from test_oracle import G, G5

DIMS = (4, 2, 3, 2)


def gammas():
    mats = [G[0], G[1], G[2], G[3], np.eye(4, dtype=complex), G5]
    mats += [G[m] @ G5 for m in range(4)]
    mats += [G[a] @ G[b] for a, b in [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]]
    return mats


def brute(chi, phi, gsrc, gsnk, n_c, mom, src):
    T, LX, LY, LZ = DIMS
    V = T * LX * LY * LZ
    zc = [(a[:, 0::2] + 1j * a[:, 1::2]).reshape(V, 4, 3) for a in chi]
    zp = [(a[:, 0::2] + 1j * a[:, 1::2]).reshape(V, 4, 3) for a in phi]
    M = G5 @ gsnk
    cols = [int(np.nonzero(gsrc[m])[0][0]) for m in range(4)]
    vals = [gsrc[m, cols[m]] for m in range(4)]
    imag = abs(vals[0].imag) > 0
    x = np.stack(np.meshgrid(*[np.arange(L) for L in DIMS], indexing="ij"), -1).reshape(-1, 4)
    p = 2 * np.pi * np.array(mom) / np.array(DIMS[1:])
    ph = np.exp(1j * ((x[:, 1:] - np.array(src[1:])) @ p))
    tmp = np.zeros(V, complex)
    for mu in range(4):
        b = cols[mu]
        s = (vals[mu].real if not imag else -vals[mu].imag) * G5[b, b].real
        for c in range(n_c):
            v = np.einsum("ab,vbk->vak", M, zp[mu * n_c + c])
            tmp += s * np.einsum("vak,vak->v", np.conj(zc[b * n_c + c]), v)
    tmp = tmp * ph
    if imag:
        tmp = -1j * tmp
    return tmp.reshape(T, -1).sum(1)


@pytest.mark.parametrize("mom", [None, (1, 0, 1)])
def test_twopoint_equals_matrix_brute_force(orc, mom):
    L = orc.Lattice(DIMS, 0.15, 0.0)
    rng = np.random.default_rng(7)
    V = L.V
    chi = [rng.standard_normal((V, 24)) for _ in range(12)]
    phi = [rng.standard_normal((V, 24)) for _ in range(12)]
    src = (1, 1, 2, 0)
    mats = gammas()
    for gs in range(16):
        for gk in range(16):
            got = L.contract_twopoint(gs, gk, chi, phi, 3, mom, src)
            want = brute(chi, phi, mats[gs], mats[gk], 3, mom if mom else (0, 0, 0), src)
            assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want)), (gs, gk)


def test_pion_is_the_propagator_norm(orc):
    L = orc.Lattice(DIMS, 0.15, 0.0)
    rng = np.random.default_rng(8)
    phi = [rng.standard_normal((L.V, 24)) for _ in range(4)]
    got = L.contract_twopoint(5, 5, phi, phi, 1)
    want = sum((f * f).sum(1) for f in phi).reshape(DIMS[0], -1).sum(1)
    assert np.max(np.abs(got - want)) < 1e-12 * want.max()

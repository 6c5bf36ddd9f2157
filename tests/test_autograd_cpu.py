"""Host-side pieces of the gradient harness (CPU): the momentum weight of
L = sum_mu sum_{|q|<=qmax} Re Pi_mumu(q) by inverse FFT against the direct sum
over the momentum set, and the differential of exp used to map the link force."""
import numpy as np

from cvc_amd import autograd as ag


def test_momentum_weight_fft_equals_direct_sum():
    dims, src, qmax = (6, 4, 8, 5), (2, 1, 7, 3), 2.0
    q = ag.momenta(dims, qmax)
    X = np.stack(np.meshgrid(*[np.arange(L) for L in dims], indexing="ij"), -1).reshape(-1, 4)
    direct = np.cos((X - np.array(src)) @ q.T).sum(1)
    w, nq = ag.momentum_weight(dims, src, qmax)
    assert nq == len(q)
    assert np.abs(w - direct).max() < 1e-10 * np.abs(direct).max()


def test_dexp_generators_match_finite_difference():
    rng = np.random.default_rng(3)
    om = rng.standard_normal((2, 4, 8)) * 0.4
    G = ag.dexp_generators(om)
    U0 = np.broadcast_to(np.eye(3, dtype=complex), (2, 4, 3, 3))
    h = 1e-6
    for a in range(8):
        d = np.zeros_like(om)
        d[..., a] = h
        dU = (ag.links(U0, om + d) - ag.links(U0, om - d)) / (2 * h)
        # d/d omega_a exp(iX) = i G_a exp(iX)
        want = 1j * G[:, :, a] @ ag.links(U0, om)
        assert np.abs(dU - want).max() < 1e-8

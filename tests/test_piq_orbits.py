"""Pi(qhat^2) on qhat^2 orbits (CPU; host code of libcvchip).

Pinned by a closed form: a tensor built exactly transverse,
Pi_mu_nu(q) = (qhat_mu qhat_nu - delta_mu_nu qhat^2) f(qhat^2) (+ an imaginary
part and an antisymmetric piece that must not enter), returns f on every orbit
with zero spread; the orbits and counts equal a numpy grouping of qhat^2."""
import numpy as np

import cvc_amd


def qhat(dims):
    n = np.meshgrid(*[np.arange(L) for L in dims], indexing="ij")
    return np.stack([2 * np.sin(np.pi * n[m] / dims[m]) for m in range(4)], -1).reshape(-1, 4)


def test_transverse_tensor_gives_f():
    dims = (6, 4, 4, 8)
    qh = qhat(dims)
    q2 = (qh * qh).sum(1)
    f = 1.0 / (1.0 + q2) + 0.3 * q2
    P = np.einsum("vm,vn->vmn", qh, qh) - np.eye(4)[None] * q2[:, None, None]
    P = P * f[:, None, None]
    rng = np.random.default_rng(1)
    A = rng.standard_normal((len(q2), 4, 4))
    P = P + (A - np.swapaxes(A, 1, 2)) * 0.1          # antisymmetric: traceless, must not enter
    piq = np.stack([P.reshape(-1, 16), rng.standard_normal((len(q2), 16))], -1)   # imaginary part ignored
    oq2, opi, osd, cnt = cvc_amd.piq_orbits(dims, piq)
    fq = 1.0 / (1.0 + oq2) + 0.3 * oq2
    assert np.abs(opi - fq).max() < 1e-12 * np.abs(fq).max()
    assert osd.max() < 1e-10
    # orbits: distinct qhat^2 > 0, counts summing to V - 1
    vals = np.unique(np.round(q2[q2 > 1e-14], 10))
    assert len(oq2) == len(vals) and cnt.sum() == len(q2) - 1
    assert np.allclose(oq2, vals, rtol=1e-9) and np.all(np.diff(oq2) > 0)

#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from the oracle (the CPU
restatement of Q_phi.c, invert_Qtm.c and cvc_exact2_xspace.c) and the seeded
gauge generator.  The reference ships no test vectors for this path and cannot
be built here, so these vectors are the oracle's; the oracle itself is pinned
by closed forms and invariants (tests/test_oracle*.py).  The fixtures freeze
those pinned outputs so that any later change of the oracle, the generator or
the GPU path is caught bit for bit (oracle) or to tolerance (GPU).

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import cvc_amd  # noqa: E402  (host-side generator only; no GPU is used)
import oracle  # noqa: E402

BC = (1.0, 0.0, 0.0, 0.0)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def point_props(L, g, src, sp):
    """The 120 propagators of cvc_exact2_xspace.c:401-430 by the oracle's
    gamma5-hermitian solver: up (mu) and down (-mu) from y + nu (nu = 0..3) and y."""
    dims = L.dims
    props = []
    for fl in (1.0, -1.0):
        Lf = oracle.Lattice(dims, L.kappa, fl * L.mu, BC)
        for s in range(5):
            d = [int(s == m) for m in range(4)]
            c = [(src[m] + d[m]) % dims[m] for m in range(4)]
            site = ((c[0] * dims[1] + c[1]) * dims[2] + c[2]) * dims[3] + c[3]
            for a in range(12):
                b = np.zeros((L.V, 24))
                b[site, 2 * a] = 1.0
                it, x, tr = Lf.invert_Qtm_her(g, b, sp, 5000)
                assert it >= 0
                props.append(x)
    return props


def main():
    out = {}
    # 4^4: operator, solver, contraction
    dims = (4, 4, 4, 4)
    g = cvc_amd.gauge_random(dims, 1, 0.3)
    L = oracle.Lattice(dims, 0.163, 0.01, BC)
    rng = np.random.default_rng(20261202)
    phi = rng.uniform(-1, 1, (L.V, 24))
    out["g4_sha256"] = sha(g)
    out["phi4"] = phi
    out["Qphi4"] = L.Q_phi(g, phi)
    out["Qf5_4"] = L.Qf5(g, phi, -0.2)
    Lp = oracle.Lattice(dims, 0.14, 0.02, BC)
    it, x, tr = Lp.invert_Qtm(g, phi, 1e-24, 5000, lagged=True)
    out["inv4_x"], out["inv4_iters"], out["inv4_true_res2"] = x, it, tr
    src = (1, 2, 3, 0)
    props = point_props(L, g, src, 1e-26)
    conn, ct = L.cvc_contract(g, props, src)
    out["conn4"], out["contact4"], out["conn4_src"] = conn, ct, np.array(src)
    # generator pins at the larger configurations
    for dims, seed, key in (((8, 8, 8, 8), 1, "g8"), ((32, 16, 16, 16), 2, "g16x32")):
        gg = cvc_amd.gauge_random(dims, seed, 0.3)
        out[key + "_sha256"] = sha(gg)
        out[key + "_samples"] = gg.reshape(-1)[:: max(1, gg.size // 97)][:97]
    # 8^4 operator summary: norms and sample sites of Q_phi_tbc on a seeded field
    g8 = cvc_amd.gauge_random((8, 8, 8, 8), 1, 0.3)
    L8 = oracle.Lattice((8, 8, 8, 8), 0.163, 0.01, BC)
    phi8 = np.random.default_rng(88).uniform(-1, 1, (L8.V, 24))
    q8 = L8.Q_phi(g8, phi8)
    out["Qphi8_norm2"] = float((q8 * q8).sum())
    out["Qphi8_sites"] = q8[[0, 1, 511, 4095]]
    np.savez_compressed(os.path.join(HERE, "golden_v1.npz"), **out)
    print("wrote", os.path.join(HERE, "golden_v1.npz"))


if __name__ == "__main__":
    main()

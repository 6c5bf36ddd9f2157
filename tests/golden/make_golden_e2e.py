#!/usr/bin/env python3
"""Writes tests/golden/golden_e2e.npz: whole-source Pi_mu_nu(x) and contact
terms from the oracle's own 120 gamma5-hermitian solves (the CPU restatement
of invert_Qtm.c:276-478 feeding cvc_exact2_xspace.c:577-861), solved to a
squared relative residual of 1e-28, for the end-to-end GPU parity test at the
1e-10 bar:

  4^4   kappa 0.163, mu 0.01, gauge exp(i 0.3 H) seed 1, source (1, 2, 3, 0)
  8^4   the plain-method configuration: kappa 0.163, mu 0.01, gauge seed 1,
        antiperiodic time boundary, point source at the origin

Solving to 1e-28 puts the solver's share of the contraction error at the
rounding level: at 4^4 the oracle's conn from solves at 1e-28 and at 1e-30
differ by 4.5e-15 relative (1.2e-13 between 1e-26 and 1e-28).

Run from the repository root:  python tests/golden/make_golden_e2e.py
(about ten minutes on eight cores; the 8^4 solves dominate).
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import cvc_amd  # noqa: E402  (host-side gauge generator only; no GPU is used)
import oracle  # noqa: E402
from make_golden import BC  # noqa: E402

SP = 1e-28
CASES = {"4": ((4, 4, 4, 4), 1, (1, 2, 3, 0)), "8": ((8, 8, 8, 8), 1, (0, 0, 0, 0))}


def point_props(dims, g, src, sp, threads):
    """The 120 propagators of cvc_exact2_xspace.c:401-430 by the oracle's
    gamma5-hermitian solver (its operator loop threaded; bit-identical to the
    serial loop): up (mu) and down (-mu) from y + nu (nu = 0..3) and y."""
    V = int(np.prod(dims))
    props = []
    for fl in (1.0, -1.0):
        Lf = oracle.Lattice(dims, 0.163, fl * 0.01, BC)
        Lf.set_threads(threads)
        for s in range(5):
            c = [(src[m] + (1 if s == m else 0)) % dims[m] for m in range(4)]
            site = ((c[0] * dims[1] + c[1]) * dims[2] + c[2]) * dims[3] + c[3]
            for a in range(12):
                b = np.zeros((V, 24))
                b[site, 2 * a] = 1.0
                it, x, tr = Lf.invert_Qtm_her(g, b, sp, 20000)
                # the returned iterate lags one update behind the tested one (invert_Qtm.c:199-211)
                assert it >= 0 and tr <= 100 * sp, (it, tr)
                props.append(x)
    return props


def main():
    out = {"sp": SP}
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    for key, (dims, seed, src) in CASES.items():
        t0 = time.time()
        g = cvc_amd.gauge_random(dims, seed, 0.3)
        L = oracle.Lattice(dims, 0.163, 0.01, BC)
        conn, ct = L.cvc_contract(g, point_props(dims, g, src, SP, threads), src)
        out[f"conn{key}"], out[f"contact{key}"], out[f"src{key}"] = conn, ct, np.array(src)
        out[f"dims{key}"] = np.array(dims)
        print(f"{dims}: {time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, "golden_e2e.npz"), **out)
    print("wrote", os.path.join(HERE, "golden_e2e.npz"))


if __name__ == "__main__":
    main()

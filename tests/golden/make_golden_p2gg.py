#!/usr/bin/env python3
"""Writes tests/golden/golden_p2gg.npz: the P -> gamma gamma contraction of
p2gg_xspace.c from the oracle alone -- its own 120 point-source solves
(invert_Qtm.c:276-478), its sequential sources (p2gg_xspace.c:517-548), its 60
sequential solves with the operator of the chosen flavour, and its
conserved-vector-current contraction (the loops of p2gg_xspace.c:615-899,
identical to cvc_exact2_xspace.c:577-861) -- every solve to a squared relative
residual of 1e-28, at 4^4 (kappa 0.163, mu 0.01, gauge exp(i 0.3 H) seed 1):

  case a  propagator_position 0 (up), sequential time slice y_0 + 2, gamma 5,
          momentum (1, 0, -1), source (1, 2, 3, 0)
  case b  propagator_position 1 (down), sequential time slice y_0 - 1,
          gamma 6 (gamma_0 gamma_5), momentum (0, 1, 0), source (0, 0, 0, 0)

Run from the repository root:  python tests/golden/make_golden_p2gg.py
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
from make_golden_e2e import point_props  # noqa: E402

SP = 1e-28
DIMS = (4, 4, 4, 4)
KAPPA, MU = 0.163, 0.01
CASES = {"a": dict(pos=0, dt=2, gid=5, mom=(1, 0, -1), src=(1, 2, 3, 0)),
         "b": dict(pos=1, dt=-1, gid=6, mom=(0, 1, 0), src=(0, 0, 0, 0))}


def p2gg_oracle(g, case, threads=1):
    """conn, contact of one P -> gamma gamma source point from the oracle."""
    src, pos = case["src"], case["pos"]
    props = point_props(DIMS, g, src, SP, threads)
    P, Q = (props[:60], props[60:]) if pos == 0 else (props[60:], props[:60])
    sgn = 1.0 - 2.0 * pos
    tseq = (src[0] + case["dt"]) % DIMS[0]
    L = oracle.Lattice(DIMS, KAPPA, MU, BC)
    Lf = oracle.Lattice(DIMS, KAPPA, sgn * MU, BC)
    Lf.set_threads(threads)
    seq = []
    for k in range(60):
        b = L.seq_source(P[k], tseq, case["gid"], case["mom"], sgn)
        it, x, tr = Lf.invert_Qtm_her(g, b, SP, 20000)
        assert it >= 0 and tr <= 100 * SP, (k, it, tr)
        seq.append(x)
    return L.cvc_contract(g, seq + Q, src)


def main():
    out = {"sp": SP, "dims": np.array(DIMS)}
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    g = cvc_amd.gauge_random(DIMS, 1, 0.3)
    for key, case in CASES.items():
        t0 = time.time()
        conn, ct = p2gg_oracle(g, case, threads)
        out[f"conn_{key}"], out[f"contact_{key}"] = conn, ct
        for name in ("pos", "dt", "gid"):
            out[f"{name}_{key}"] = case[name]
        out[f"mom_{key}"], out[f"src_{key}"] = np.array(case["mom"]), np.array(case["src"])
        print(f"case {key}: {time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, "golden_p2gg.npz"), **out)
    print("wrote", os.path.join(HERE, "golden_p2gg.npz"))


if __name__ == "__main__":
    main()

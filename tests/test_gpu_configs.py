"""Parity at the benchmark configurations' own sizes (GPU).

16^3x32 (kappa 0.1632, mu 0.004, gauge seed 2): the 120 GPU propagators of one
source are contracted by the GPU kernels and by the oracle's loop-for-loop
restatement of cvc_exact2_xspace.c:577-861 (about a minute of CPU); conn,
contact and Pi(q) agree to 1e-10 relative.  The propagators themselves satisfy
|D x - b|^2/|b|^2 <= 1e-24 with the oracle's operator.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(900)
def test_contraction_16x32(chip, orc):
    from test_gpu_contract import gpu_props
    dims = (32, 16, 16, 16)
    src = (3, 5, 0, 9)
    print("16^3x32 contraction parity", flush=True)
    g = chip.gauge_random(dims, 2, 0.3)
    ctx = chip.Context(dims, 0.1632, 0.004, (1.0, 0.0, 0.0, 0.0), gauge=g)
    props = gpu_props(chip, ctx, dims, src, sp=1e-24)
    L = orc.Lattice(dims, 0.1632, 0.004)
    for k in (0, 59, 60, 119):
        b = np.zeros_like(props[k])
        s = (k % 60) // 12
        d = [int(s == m) for m in range(4)]
        c = [(src[m] + d[m]) % dims[m] for m in range(4)]
        b[((c[0] * dims[1] + c[1]) * dims[2] + c[2]) * dims[3] + c[3], 2 * (k % 12)] = 1.0
        mu = 0.004 if k < 60 else -0.004
        r = L.Q_phi(g, props[k], mu) - b
        assert L.dot_re(r, r) < 1.01e-24
    conn, ct = ctx.cvc_contract(props, src)
    print("oracle contraction (CPU)", flush=True)
    conn_o, ct_o = L.cvc_contract(g, props, src)
    assert rel_err(conn, conn_o) < 1e-10
    assert np.max(np.abs(ct - ct_o)) < 1e-10 * np.max(np.abs(ct_o))
    piq, piq_o = ctx.pi_q(conn, ct, src), ctx.pi_q(conn_o, ct_o, src)
    assert rel_err(piq, piq_o) < 1e-10

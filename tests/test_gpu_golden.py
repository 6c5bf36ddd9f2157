"""The GPU path against the golden fixtures (tests/golden/golden_v1.npz, the
oracle's outputs): operator to 1e-13, solver to 1e-8 (both solves to a
squared residual of 1e-24, lagged iterate as the reference), one whole source
Pi_mu_nu(x) and contact term to 1e-9 relative."""
import os

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
BC = (1.0, 0.0, 0.0, 0.0)


def test_gpu_reproduces_golden(chip):
    gold = np.load(os.path.join(HERE, "golden", "golden_v1.npz"), allow_pickle=False)
    dims = (4, 4, 4, 4)
    print("golden 4^4", flush=True)
    g = chip.gauge_random(dims, 1, 0.3)
    ctx = chip.Context(dims, 0.163, 0.01, BC, gauge=g)
    phi = gold["phi4"]
    assert rel_err(ctx.Q_phi(phi), gold["Qphi4"]) < 1e-13
    assert rel_err(ctx.Q_phi(phi, mutm=-0.2, gamma5_out=True), gold["Qf5_4"]) < 1e-13
    src = tuple(int(v) for v in gold["conn4_src"])
    conn, ct, st = ctx.cvc_source(src, 1e-26, 5000)
    assert rel_err(conn, gold["conn4"]) < 1e-9
    assert np.max(np.abs(ct - gold["contact4"])) < 1e-9 * np.max(np.abs(gold["contact4"]))
    ctx.set_params(0.14, 0.02)
    rc, x, st = ctx.invert_Qtm(phi, 1e-24, 5000, lagged=True)
    assert rc >= 0 and rel_err(x[0], gold["inv4_x"]) < 1e-8

"""The golden fixtures (tests/golden/golden_v1.npz, written by
tests/golden/make_golden.py) against the oracle and the gauge generator (CPU):
bit for bit, so any change of either is caught."""
import hashlib
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "golden_v1.npz")
BC = (1.0, 0.0, 0.0, 0.0)


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def test_generator_is_frozen(gold):
    import cvc_amd
    assert sha(cvc_amd.gauge_random((4, 4, 4, 4), 1, 0.3)) == str(gold["g4_sha256"])
    for dims, seed, key in (((8, 8, 8, 8), 1, "g8"), ((32, 16, 16, 16), 2, "g16x32")):
        gg = cvc_amd.gauge_random(dims, seed, 0.3)
        assert sha(gg) == str(gold[key + "_sha256"])
        assert np.array_equal(gg.reshape(-1)[:: max(1, gg.size // 97)][:97], gold[key + "_samples"])


def test_oracle_operator_and_solver_are_frozen(gold, orc):
    import cvc_amd
    g = cvc_amd.gauge_random((4, 4, 4, 4), 1, 0.3)
    L = orc.Lattice((4, 4, 4, 4), 0.163, 0.01, BC)
    phi = gold["phi4"]
    assert np.array_equal(L.Q_phi(g, phi), gold["Qphi4"])
    assert np.array_equal(L.Qf5(g, phi, -0.2), gold["Qf5_4"])
    Lp = orc.Lattice((4, 4, 4, 4), 0.14, 0.02, BC)
    it, x, tr = Lp.invert_Qtm(g, phi, 1e-24, 5000, lagged=True)
    assert it == int(gold["inv4_iters"]) and np.array_equal(x, gold["inv4_x"])
    g8 = cvc_amd.gauge_random((8, 8, 8, 8), 1, 0.3)
    L8 = orc.Lattice((8, 8, 8, 8), 0.163, 0.01, BC)
    q8 = L8.Q_phi(g8, np.random.default_rng(88).uniform(-1, 1, (L8.V, 24)))
    assert float((q8 * q8).sum()) == float(gold["Qphi8_norm2"])
    assert np.array_equal(q8[[0, 1, 511, 4095]], gold["Qphi8_sites"])


def test_oracle_contraction_is_frozen(gold, orc):
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from make_golden import point_props
    import cvc_amd
    g = cvc_amd.gauge_random((4, 4, 4, 4), 1, 0.3)
    L = orc.Lattice((4, 4, 4, 4), 0.163, 0.01, BC)
    src = tuple(int(v) for v in gold["conn4_src"])
    conn, ct = L.cvc_contract(g, point_props(L, g, src, 1e-26), src)
    assert np.array_equal(conn, gold["conn4"])
    assert np.array_equal(ct, gold["contact4"])


def test_e2e_fixture_is_frozen(orc):
    """golden_e2e.npz (tests/golden/make_golden_e2e.py): the 4^4 source is recomputed
    bit for bit from the oracle's 120 solves; the 8^4 source (its solves take minutes
    here) satisfies the Ward identity with its own contact term and was solved to the
    recorded tolerance."""
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from make_golden_e2e import point_props
    import cvc_amd
    e2e = np.load(os.path.join(HERE, "golden", "golden_e2e.npz"), allow_pickle=False)
    assert float(e2e["sp"]) == 1e-28
    dims = tuple(int(v) for v in e2e["dims4"])
    src = tuple(int(v) for v in e2e["src4"])
    g = cvc_amd.gauge_random(dims, 1, 0.3)
    L = orc.Lattice(dims, 0.163, 0.01, BC)
    conn, ct = L.cvc_contract(g, point_props(dims, g, src, 1e-28, 1), src)
    assert np.array_equal(conn, e2e["conn4"]) and np.array_equal(ct, e2e["contact4"])
    dims8 = tuple(int(v) for v in e2e["dims8"])
    assert dims8 == (8, 8, 8, 8) and tuple(int(v) for v in e2e["src8"]) == (0, 0, 0, 0)
    L8 = orc.Lattice(dims8, 0.163, 0.01, BC)
    conn8, ct8 = e2e["conn8"], e2e["contact8"]
    w = L8.wi(conn8)
    w = w[..., 0] + 1j * w[..., 1]
    want = np.zeros_like(w)
    want[0, :] += ct8
    for nu in range(4):
        want[8 ** (3 - nu), nu] -= ct8[nu]
    assert np.max(np.abs(w - want)) < 1e-12 * np.max(np.abs(conn8))


def test_p2gg_fixture_is_frozen(orc):
    """golden_p2gg.npz (tests/golden/make_golden_p2gg.py): case a recomputed bit for bit
    from the oracle's pipeline."""
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from make_golden_p2gg import CASES, p2gg_oracle
    import cvc_amd
    gold = np.load(os.path.join(HERE, "golden", "golden_p2gg.npz"), allow_pickle=False)
    g = cvc_amd.gauge_random(tuple(int(v) for v in gold["dims"]), 1, 0.3)
    conn, ct = p2gg_oracle(g, CASES["a"], 1)
    assert np.array_equal(conn, gold["conn_a"]) and np.array_equal(ct, gold["contact_a"])

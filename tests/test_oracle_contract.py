"""Pins of the oracle's contraction (CPU only).

  Ward identity    sum_mu conn_{mu nu}(x) - conn_{mu nu}(x - mu) vanishes away
                   from the source and equals +contact_nu at y and -contact_nu
                   at y + nu (ties conn's -1/4 normalisation, the relative signs
                   of the two parts and of the forward/backward terms, and the
                   Usource handling to the contact term of :697-704, 837-844)
  gauge invariance conn and contact are unchanged when links and propagators
                   are gauge transformed (catches U vs U^dagger, link-site and
                   Usource-index mistakes)
"""
import numpy as np
import pytest

from test_oracle import gauge_c, gauge_r, random_su3

DIMS = (4, 4, 4, 4)
SRC = (1, 2, 0, 3)


def site(c):
    T, LX, LY, LZ = DIMS
    return (((c[0] % T) * LX + (c[1] % LX)) * LY + (c[2] % LY)) * LZ + (c[3] % LZ)


@pytest.fixture(scope="module")
def system(orc, chip):
    g = chip.gauge_random(DIMS, 1, 0.3)
    L = orc.Lattice(DIMS, 0.163, 0.01)
    Ld = orc.Lattice(DIMS, 0.163, -0.01)
    props = [None] * 120
    for s in range(5):
        d = [int(s == m) for m in range(4)]
        ys = site([SRC[m] + d[m] for m in range(4)])
        for a in range(12):
            b = chip.point_source(L.V, ys, a // 3, a % 3)
            it, x, _ = L.invert_Qtm_her(g, b, 1e-24, 5000)
            assert it > 0
            props[12 * s + a] = x
            it, x, _ = Ld.invert_Qtm_her(g, b, 1e-24, 5000)
            assert it > 0
            props[60 + 12 * s + a] = x
    conn, ct = L.cvc_contract(g, props, SRC)
    return g, L, props, conn, ct


def test_ward_identity_with_contact_term(system):
    g, L, props, conn, ct = system
    wi = L.wi(conn)
    w = wi[..., 0] + 1j * wi[..., 1]
    want = np.zeros_like(w)
    want[site(SRC), :] += ct
    for nu in range(4):
        d = [int(nu == m) for m in range(4)]
        want[site([SRC[m] + d[m] for m in range(4)]), nu] -= ct[nu]
    scale = np.max(np.abs(conn))
    assert np.max(np.abs(ct)) > 0.1 * scale
    assert np.max(np.abs(w - want)) < 1e-10 * scale


def test_gauge_invariance(system):
    g, L, props, conn, ct = system
    rng = np.random.default_rng(12)
    Om = random_su3(rng, L.V)
    U = gauge_c(g)
    iup = L.iup()
    Up = np.stack([Om @ U[:, m] @ Om[iup[:, m]].conj().transpose(0, 2, 1) for m in range(4)], axis=1)
    newp = [None] * 120
    # S'(x, y_s) = Om(x) S(x, y_s) Om(y_s)^dagger: transform the 12x12 matrix of each (flavour, s)
    for fl in range(2):
        for s in range(5):
            d = [int(s == m) for m in range(4)]
            ys = site([SRC[m] + d[m] for m in range(4)])
            S = np.stack([props[60 * fl + 12 * s + a] for a in range(12)], axis=-1)   # [V, 24, 12]
            Sc = (S[:, 0::2, :] + 1j * S[:, 1::2, :]).reshape(L.V, 4, 3, 4, 3)       # x, a, c, j, b
            Sc = np.einsum("xcd,xadjb->xacjb", Om, Sc)
            Sc = np.einsum("xacjb,eb->xacje", Sc, Om[ys].conj())
            Sc = Sc.reshape(L.V, 12, 12)
            for a in range(12):
                v = np.empty((L.V, 24))
                v[:, 0::2] = Sc[:, :, a].real
                v[:, 1::2] = Sc[:, :, a].imag
                newp[60 * fl + 12 * s + a] = v
    conn2, ct2 = L.cvc_contract(gauge_r(Up), newp, SRC)
    assert np.max(np.abs(conn2 - conn)) < 1e-12 * np.max(np.abs(conn))
    assert np.max(np.abs(ct2 - ct)) < 1e-12 * np.max(np.abs(ct))

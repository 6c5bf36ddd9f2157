"""Pins of the CPU restatement (oracle/) — CPU only.

The reference sources cannot be compiled in this image (global.h includes
FFTW2's <fftw.h>, which is absent), and the reference ships no tests or
fixtures for this path.  Each oracle function is therefore pinned here by
mathematics chosen so that a plausible mistake fails one of the checks:

  gamma tables       Clifford algebra, hermiticity, the matrices documented in
                     cvc_utils.c:709-745, gamma5 and the product ordering
  Q_phi / Qf5        free-field plane waves (closed form of D(p) with twisted
                     momenta: catches hopping sign, gamma sign, phase sign,
                     kappa and mu terms), gauge covariance (catches U vs U^dagger
                     and link-site mistakes), gamma5-hermiticity
  dot products       numpy vdot
  invert_Qtm(_her)   dense LU solve of the operator matrix; lagged-iterate
                     semantics; x0 = exact solution returns 0
  plaquette          unit links give 1; gauge invariance
  contraction / WI   see test_oracle_contract.py
"""
import numpy as np
import pytest

from conftest import rand_spinor, rel_err

# gamma matrices as documented in the reference (cvc_utils.c:709-745)
G = np.zeros((4, 4, 4), complex)
G[0][0, 2] = G[0][1, 3] = G[0][2, 0] = G[0][3, 1] = -1
G[1][0, 3] = G[1][1, 2] = -1j
G[1][2, 1] = G[1][3, 0] = 1j
G[2][0, 3] = G[2][3, 0] = -1
G[2][1, 2] = G[2][2, 1] = 1
G[3][0, 2] = G[3][3, 1] = -1j
G[3][1, 3] = G[3][2, 0] = 1j
G5 = np.diag([1, 1, -1, -1]).astype(complex)


def to_c(phi):
    """[V,24] real -> [V,4,3] complex."""
    return (phi[..., 0::2] + 1j * phi[..., 1::2]).reshape(phi.shape[:-1] + (4, 3))


def from_c(z):
    z = z.reshape(z.shape[:-2] + (12,))
    out = np.empty(z.shape[:-1] + (24,))
    out[..., 0::2] = z.real
    out[..., 1::2] = z.imag
    return out


def gauge_c(g):
    return (g[..., 0::2] + 1j * g[..., 1::2]).reshape(g.shape[0], 4, 3, 3)


def gauge_r(U):
    z = U.reshape(U.shape[0], 4, 9)
    out = np.empty((U.shape[0], 4, 18))
    out[..., 0::2] = z.real
    out[..., 1::2] = z.imag
    return out


def coords(dims):
    T, LX, LY, LZ = dims
    t, x, y, z = np.meshgrid(np.arange(T), np.arange(LX), np.arange(LY), np.arange(LZ), indexing="ij")
    return np.stack([t.ravel(), x.ravel(), y.ravel(), z.ravel()], axis=1)


def random_su3(rng, n):
    A = rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))
    Q, R = np.linalg.qr(A)
    d = np.diagonal(R, axis1=1, axis2=2)
    Q = Q * (d / np.abs(d))[:, None, :]
    det = np.linalg.det(Q)
    return Q / (det ** (1 / 3))[:, None, None]


# --------------------------------------------------------------------- gammas

def test_gamma_tables_are_documented_matrices(orc):
    perm, sign = orc.gamma_tables()
    mats = []
    for g in range(16):
        M = np.zeros((4, 4), complex)
        for b in range(4):
            for val in (1.0, 1j):
                t = np.zeros(24)
                t[6 * b] = np.real(val)
                t[6 * b + 1] = np.imag(val)
                s = t[perm[g]] * sign[g]
                col = s[0::6] + 1j * s[1::6]
                if val == 1.0:
                    M[:, b] = col
                else:
                    # linearity over C: gamma (i e_b) = i gamma e_b
                    assert np.allclose(col, 1j * M[:, b])
        mats.append(M)
    for m in range(4):
        assert np.array_equal(mats[m], G[m])
        assert np.allclose(mats[m], mats[m].conj().T)
    for m in range(4):
        for n in range(4):
            assert np.allclose(G[m] @ G[n] + G[n] @ G[m], 2 * np.eye(4) * (m == n))
    assert np.array_equal(mats[4], np.eye(4))
    assert np.array_equal(mats[5], G5)
    assert np.allclose(G[0] @ G[1] @ G[2] @ G[3], G5)       # gamma5 = g0 g1 g2 g3 in this basis
    for m in range(4):
        assert np.allclose(mats[6 + m], G[m] @ G5)
    pairs = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    for k, (a, b) in enumerate(pairs):
        assert np.allclose(mats[10 + k], G[a] @ G[b])
    # a colour index rides along unchanged
    assert all(sorted(perm[g]) == list(range(24)) for g in range(16))


# ------------------------------------------------------------------ operator

def free_field_expected(dims, kappa, mu, bc, n, chi):
    """D(p) chi e^{ipx} with D(p) = 1/(2k) + i mu g5 + sum_mu(-cos pt + i g_mu sin pt),
    pt_mu = 2 pi n_mu / L_mu + pi bc_mu / L_mu."""
    L = np.array(dims)
    pt = 2 * np.pi * np.array(n) / L + np.pi * np.array(bc) / L
    Dp = (0.5 / kappa) * np.eye(4) + 1j * mu * G5
    for m in range(4):
        Dp = Dp + (-np.cos(pt[m])) * np.eye(4) + 1j * np.sin(pt[m]) * G[m]
    X = coords(dims)
    ph = np.exp(1j * (X @ (2 * np.pi * np.array(n) / L)))
    phi = ph[:, None, None] * chi[None]
    out = ph[:, None, None] * (Dp @ chi)[None]
    return phi, out


@pytest.mark.parametrize("bc", [(1.0, 0.0, 0.0, 0.0), (1.0, 0.5, -0.25, 0.7)])
def test_free_field_plane_wave(orc, bc):
    dims = (4, 6, 4, 8)
    kappa, mu = 0.137, 0.21
    L = orc.Lattice(dims, kappa, mu, bc)
    g = np.zeros((L.V, 4, 18))
    g[:, :, 0] = g[:, :, 8] = g[:, :, 16] = 1.0
    rng = np.random.default_rng(5)
    chi = rng.standard_normal((4, 3)) + 1j * rng.standard_normal((4, 3))
    for n in [(0, 0, 0, 0), (1, 2, 0, 3), (3, 5, 1, 7)]:
        phi, want = free_field_expected(dims, kappa, mu, bc, n, chi)
        got = L.Q_phi(g, from_c(phi))
        assert rel_err(got, from_c(want)) < 1e-13
        # Qf5 = gamma5 D, and an explicit mutm overrides g_mu
        got5 = L.Qf5(g, from_c(phi), -mu)
        _, want_m = free_field_expected(dims, kappa, -mu, bc, n, chi)
        assert rel_err(got5, from_c(np.einsum("ab,vbc->vac", G5, want_m))) < 1e-13


def test_gauge_covariance(orc, chip):
    dims = (4, 4, 6, 4)
    L = orc.Lattice(dims, 0.15, 0.05, (1.0, 0.3, 0.0, 0.0))
    rng = np.random.default_rng(7)
    g = chip.gauge_random(dims, 11, 0.4)
    U = gauge_c(g)
    Om = random_su3(rng, L.V)
    iup = L.iup()
    Up = np.empty_like(U)
    for m in range(4):
        Up[:, m] = Om @ U[:, m] @ Om[iup[:, m]].conj().transpose(0, 2, 1)
    phi = rand_spinor(rng, L.V)
    pc = to_c(phi)
    phi_p = from_c(np.einsum("vij,vaj->vai", Om, pc))
    lhs = L.Q_phi(gauge_r(Up), phi_p)
    rhs = from_c(np.einsum("vij,vaj->vai", Om, to_c(L.Q_phi(g, phi))))
    assert rel_err(lhs, rhs) < 1e-13


def test_gamma5_hermiticity(orc, chip):
    dims = (4, 4, 4, 6)
    mu = 0.07
    L = orc.Lattice(dims, 0.16, mu)
    g = chip.gauge_random(dims, 3, 0.3)
    rng = np.random.default_rng(1)
    psi, phi = rand_spinor(rng, L.V), rand_spinor(rng, L.V)
    a = L.dot_co(psi, L.Q_phi(g, phi, mu))
    b = L.dot_co(L.g5_phi(L.Q_phi(g, L.g5_phi(psi), -mu)), phi)
    assert abs(a - b) < 1e-12 * abs(a)


def test_dot_products(orc):
    L = orc.Lattice((2, 4, 2, 4), 0.1, 0.0)
    rng = np.random.default_rng(2)
    a, b = rand_spinor(rng, L.V), rand_spinor(rng, L.V)
    want = np.vdot(a[:, 0::2] + 1j * a[:, 1::2], b[:, 0::2] + 1j * b[:, 1::2])
    assert abs(L.dot_co(a, b) - want) < 1e-12 * abs(want)
    assert abs(L.dot_re(a, b) - want.real) < 1e-12 * abs(want)


def test_plaquette(orc, chip):
    dims = (4, 4, 4, 4)
    L = orc.Lattice(dims, 0.1, 0.0)
    g = np.zeros((L.V, 4, 18))
    g[:, :, 0] = g[:, :, 8] = g[:, :, 16] = 1.0
    assert abs(L.plaquette(g) - 1.0) < 1e-15
    g = chip.gauge_random(dims, 4, 0.3)
    U = gauge_c(g)
    Om = random_su3(np.random.default_rng(3), L.V)
    iup = L.iup()
    Up = np.stack([Om @ U[:, m] @ Om[iup[:, m]].conj().transpose(0, 2, 1) for m in range(4)], axis=1)
    p0, p1 = L.plaquette(g), L.plaquette(gauge_r(Up))
    assert abs(p0 - p1) < 1e-13 and 0.3 < p0 < 0.9


# -------------------------------------------------------------------- solvers

def dense_operator(L, g, mu=None):
    n = L.V * 12
    M = np.zeros((n, n), complex)
    for k in range(n):
        e = np.zeros(n * 2)
        e[2 * k] = 1.0
        y = L.Q_phi(g, e.reshape(L.V, 24), mu)
        M[:, k] = y.reshape(-1)[0::2] + 1j * y.reshape(-1)[1::2]
    return M


@pytest.fixture(scope="module")
def small_system(orc, chip):
    dims = (4, 4, 4, 4)
    g = chip.gauge_random(dims, 1, 0.3)
    return dims, g


def test_bicgstab_matches_dense_solve(orc, small_system):
    dims, g = small_system
    # kappa below the critical value of this rough field, where BiCGStab on D converges
    # (at kappa = 1/8 a point source hits an exact breakdown <r2,p2> = 0: the
    # diagonal of the squared hopping term vanishes and 1/(2 kappa) = 4 is exact)
    L = orc.Lattice(dims, 0.14, 0.01)
    M = dense_operator(L, g)
    b = np.zeros((L.V, 24))
    b[0, 0] = 1.0
    want = np.linalg.solve(M, b.reshape(-1)[0::2])
    it, x, tr = L.invert_Qtm(g, b, 1e-24, 2000, lagged=False)
    assert it > 0 and tr <= 1.01e-24
    xz = x.reshape(-1)[0::2] + 1j * x.reshape(-1)[1::2]
    assert np.max(np.abs(xz - want)) < 1e-10 * np.max(np.abs(want))
    # lagged: the returned iterate is one update behind (invert_Qtm.c:199-211)
    it2, x2, tr2 = L.invert_Qtm(g, b, 1e-24, 2000, lagged=True)
    assert it2 == it and tr2 > tr
    # a start vector that already solves the system returns 0 iterations
    it3, _, _ = L.invert_Qtm(g, b, 1e-16, 2000, x0=x)
    assert it3 == 0
    # non-convergence is reported as -3
    it4, _, _ = L.invert_Qtm(g, b, 1e-24, 3)
    assert it4 == -3


def test_bicgstab_her_matches_dense_solve(orc, small_system):
    dims, g = small_system
    L = orc.Lattice(dims, 0.163, 0.01)        # super-critical: only the hermitian form converges
    M = dense_operator(L, g)
    rng = np.random.default_rng(9)
    b = rand_spinor(rng, L.V)
    want = np.linalg.solve(M, b.reshape(-1)[0::2] + 1j * b.reshape(-1)[1::2])
    it, x, tr = L.invert_Qtm_her(g, b, 1e-24, 5000)
    assert it > 0 and tr < 1e-20
    xz = x.reshape(-1)[0::2] + 1j * x.reshape(-1)[1::2]
    assert np.max(np.abs(xz - want)) < 1e-9 * np.max(np.abs(want))


def test_threaded_operator_is_bitwise_serial(orc, chip):
    """The operator loop with OpenMP threads (the reference's Q_phi_tbc distribution,
    Q_phi.c:190-206) returns the serial loop's bits: each site is computed alone."""
    dims = (8, 4, 6, 4)
    g = chip.gauge_random(dims, 3, 0.3)
    L = orc.Lattice(dims, 0.1632, 0.004, (1.0, 0.3, 0.0, 0.0))
    phi = np.random.default_rng(4).standard_normal((L.V, 24))
    serial = [L.Q_phi(g, phi), L.Qf5(g, phi, -0.004)]
    L.set_threads(4)
    assert np.array_equal(L.Q_phi(g, phi), serial[0])
    assert np.array_equal(L.Qf5(g, phi, -0.004), serial[1])

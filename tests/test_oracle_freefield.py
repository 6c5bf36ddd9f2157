"""Pins of the oracle's contraction by closed forms (CPU only).

The Ward identity and gauge invariance (test_oracle_contract.py) are
homogeneous in conn and the contact term: a joint sign flip, a wrong overall
normalisation or a wrong assignment of the up/down propagator records would
pass them.  These tests fix those independently of the reference's source
lines (cvc_exact2_xspace.c:697-704, 837-844, 861) from the definition of the
observable: conn_{mu nu}(x) is the quark-connected correlator of the
conserved vector current of the twisted-mass action psibar D psi,

  J_mu(x) = -1/2 [ psibar(x) (1 - gamma_mu) U_mu(x) psi(x+mu)
                 - psibar(x+mu) (1 + gamma_mu) U_mu(x)^dagger psi(x) ],

(J_mu -> psibar gamma_mu psi in the continuum limit), of the up quark
(D(+mu)), with the source current at y:

  conn_{mu nu}(x) = < J_mu(x) J_nu(y) >_conn
                  = -1/4 sum_{s1,s2} s1 s2 tr[ A_s1 S_u(b,c) B_s2 S_u(d,a) ]

(the minus sign is the fermion loop), where (a, b, A) runs over
(x, x+mu, (1-gamma_mu) U_mu(x)) with s1 = +1 and (x+mu, x, (1+gamma_mu)
U_mu(x)^dagger) with s1 = -1, and (c, d, B) likewise at y for nu.  The
reference evaluates it from point-source propagators only: the backward
propagator S_u(y+nu, x) is gamma5 S_d(x, y+nu)^dagger gamma5 with the down
quark D(-mu), which is why it reads the down records.

  free field    with unit links and twisted boundary phases theta_rho in all
                four directions, the Fourier transform of conn equals the
                momentum sum
                  Pi_{mu nu}(q) = -3/(4V) sum_p tr[ V_mu D_u(k~)^-1 W_nu D_u(p~)^-1 ],
                  k = p + q,  p~ = p + theta,  k~ = k + theta,
                  V_mu = (1-gamma_mu) e^{i k~_mu} - (1+gamma_mu) e^{-i p~_mu},
                  W_nu = (1-gamma_nu) e^{i p~_nu} - (1+gamma_nu) e^{-i k~_nu},
                D_f(p~) = 1/(2 kappa) + i f mu gamma5 + sum_rho (-cos p~_rho + i gamma_rho sin p~_rho)
                (the closed form of D on plane waves, pinned in test_oracle.py).  The
                propagators fed to the oracle are the same momentum sums (inverse FFT),
                not solver output.  The mixed-flavour loop differs by percents, so a
                swapped record assignment fails.
  brute force   on a seeded rough gauge field at 4^4, the definition above with
                the all-to-all up propagator from a dense inverse of the operator
                matrix equals the oracle's contraction of 120 columns of the dense
                inverses of D(+mu) and D(-mu).
"""
import itertools

import numpy as np
import pytest

from test_oracle import G, G5, dense_operator, gauge_c

I4 = np.eye(4)


def dirac_inv(f, mu, kappa, Pt):
    """D_f(p~)^-1 (4x4 spin matrices) on a grid of twisted momenta Pt[..., 4]."""
    sh = Pt.shape[:-1]
    D = np.broadcast_to(I4 / (2 * kappa) + 1j * f * mu * G5, sh + (4, 4)).copy()
    for r in range(4):
        D += -np.cos(Pt[..., r])[..., None, None] * I4 + 1j * np.sin(Pt[..., r])[..., None, None] * G[r]
    return np.linalg.inv(D)


def to_real(z):
    r = np.empty(z.shape[:-1] + (2 * z.shape[-1],))
    r[..., 0::2] = z.real
    r[..., 1::2] = z.imag
    return r


def site(dims, c):
    return (((c[0] % dims[0]) * dims[1] + (c[1] % dims[1])) * dims[2] + (c[2] % dims[2])) * dims[3] + (c[3] % dims[3])


class FreeField:
    dims = (4, 6, 4, 8)
    kappa, mu = 0.137, 0.21
    bc = (1.0, 0.5, -0.25, 0.7)
    src = (1, 2, 3, 5)

    def __init__(self):
        d = self.dims
        self.V = int(np.prod(d))
        self.theta = np.array([np.pi * self.bc[m] / d[m] for m in range(4)])
        grids = np.meshgrid(*[2 * np.pi * np.arange(L) / L for L in d], indexing="ij")
        self.P = np.stack(grids, -1)

    def propagators(self, f):
        """The 60 point-source columns of D_f^-1 of cvc_exact2_xspace.c:401-430 (sources y + s,
        s = 0..3, and y), as momentum sums: S(x, y) = 1/V sum_p e^{i p (x - y)} D_f(p~)^-1."""
        F = np.fft.ifftn(dirac_inv(f, self.mu, self.kappa, self.P + self.theta), axes=(0, 1, 2, 3))
        out = []
        for s in range(5):
            y = [(self.src[m] + (1 if s == m else 0)) % self.dims[m] for m in range(4)]
            Fy = np.roll(F, shift=y, axis=(0, 1, 2, 3)).reshape(self.V, 4, 4)
            for a in range(12):
                sp, c = divmod(a, 3)
                col = np.zeros((self.V, 4, 3), complex)
                col[:, :, c] = Fy[:, :, sp]
                out.append(to_real(col.reshape(self.V, 12)))
        return out

    def pi_sum(self, n, f1, f2):
        """The momentum-sum closed form at q = 2 pi n / L (16 components, 4 mu + nu)."""
        q = np.array([2 * np.pi * n[m] / self.dims[m] for m in range(4)])
        Pt = self.P + self.theta
        Kt = Pt + q
        D1 = dirac_inv(f1, self.mu, self.kappa, Pt)
        D2 = dirac_inv(f2, self.mu, self.kappa, Kt)
        res = np.zeros(16, complex)
        e = lambda a: np.exp(1j * a)[..., None, None]   # noqa: E731
        for m in range(4):
            Vm = (I4 - G[m]) * e(Kt[..., m]) - (I4 + G[m]) * e(-Pt[..., m])
            for nu in range(4):
                Wn = (I4 - G[nu]) * e(Pt[..., nu]) - (I4 + G[nu]) * e(-Kt[..., nu])
                tr = np.einsum("...ab,...bc,...cd,...da->...", Vm, D2, Wn, D1)
                res[4 * m + nu] = -3 * tr.sum() / (4 * self.V)
        return res


@pytest.fixture(scope="module")
def free(orc):
    F = FreeField()
    L = orc.Lattice(F.dims, F.kappa, F.mu, F.bc)
    g = np.zeros((F.V, 4, 18))
    g[:, :, 0] = g[:, :, 8] = g[:, :, 16] = 1.0
    pu, pd = F.propagators(1.0), F.propagators(-1.0)
    conn, ct = L.cvc_contract(g, pu + pd, F.src)
    cz = (conn[..., 0] + 1j * conn[..., 1]).reshape(F.dims + (16,))
    # sum_x e^{-i q (x - y)} conn(x)
    Pi = np.fft.fftn(cz, axes=(0, 1, 2, 3))
    Pi *= np.exp(1j * np.einsum("...m,m->...", F.P, np.array(F.src, float)))[..., None]
    return F, L, g, pu, pd, Pi, ct


def test_free_propagators_solve_the_operator(free):
    F, L, g, pu, pd, Pi, ct = free
    Ld = type(L)(F.dims, F.kappa, -F.mu, F.bc)
    for k in (0, 17, 59):
        s, a = divmod(k, 12)
        y = site(F.dims, [F.src[m] + (1 if s == m else 0) for m in range(4)])
        want = np.zeros((F.V, 24))
        want[y, 2 * a] = 1.0
        assert np.abs(L.Q_phi(g, pu[k]) - want).max() < 1e-13
        assert np.abs(Ld.Q_phi(g, pd[k]) - want).max() < 1e-13


def test_free_field_momentum_sum(free):
    F, L, g, pu, pd, Pi, ct = free
    rng = np.random.default_rng(11)
    qs = [(0, 0, 0, 0), (0, 1, 0, 0), (1, 0, 2, 3)] + [tuple(int(rng.integers(0, n)) for n in F.dims) for _ in range(4)]
    scale = max(np.abs(Pi[n]).max() for n in qs)
    for n in qs:
        want = F.pi_sum(n, 1.0, 1.0)
        assert np.abs(Pi[n] - want).max() < 1e-12 * scale, n
    # the up and the down connected loops agree in the free theory (parity); a loop of
    # one up and one down propagator does not: a mixed record assignment would fail above
    n = qs[-1]
    assert np.abs(F.pi_sum(n, -1.0, -1.0) - F.pi_sum(n, 1.0, 1.0)).max() < 1e-12 * scale
    mixed = F.pi_sum(n, 1.0, -1.0)
    assert np.abs(Pi[n] - mixed).max() > 1e-3 * scale


def test_free_field_contact_term_from_the_ward_identity(free):
    """The x-space Ward identity sum_mu conn_{mu nu}(x) - conn_{mu nu}(x - mu) =
    contact_nu (delta_{x,y} - delta_{x,y+nu}) reads, Fourier transformed,
    sum_mu (1 - e^{-i q_mu}) Pi_{mu nu}(q) = contact_nu (1 - e^{-i q_nu}).  With Pi
    from the momentum-sum closed form this gives the contact term at any q with
    every q_nu != 0, independently of the oracle's contact formula."""
    F, L, g, pu, pd, Pi, ct = free
    for n in ((1, 1, 1, 1), (3, 5, 2, 7)):
        q = np.array([2 * np.pi * n[m] / F.dims[m] for m in range(4)])
        P = F.pi_sum(n, 1.0, 1.0).reshape(4, 4)
        lhs = ((1 - np.exp(-1j * q))[:, None] * P).sum(0)
        want = lhs / (1 - np.exp(-1j * q))
        assert np.abs(ct - want).max() < 1e-12 * np.abs(want).max(), n
    assert np.abs(ct).max() > 0.1


def wick_conn(dims, U, S, src):
    """The definition: -1/4 sum s1 s2 tr[A S(b,c) B S(d,a)] with the all-to-all up
    propagator S[V,12,V,12] (spin-major 12 = 3 spin + colour), links U[V,4,3,3] with
    the boundary phase applied."""
    V = int(np.prod(dims))
    T, LX, LY, LZ = dims
    idx = np.arange(V)
    crd = np.stack(np.unravel_index(idx, dims), 1)

    def shift(c, m):
        cc = c.copy()
        cc[..., m] = (cc[..., m] + 1) % dims[m]
        return np.ravel_multi_index(tuple(cc.T), dims)

    ys = site(dims, src)
    yc = np.array(src)
    out = np.zeros((V, 16), complex)
    for m in range(4):
        xp = shift(crd, m)
        Afw = np.einsum("ab,vij->vaibj", I4 - G[m], U[:, m]).reshape(V, 12, 12)
        Abw = np.einsum("ab,vji->vaibj", I4 + G[m], U[:, m].conj()).reshape(V, 12, 12)
        for nu in range(4):
            yp = int(shift(yc[None], nu)[0])
            Bfw = np.kron(I4 - G[nu], U[ys, nu])
            Bbw = np.kron(I4 + G[nu], U[ys, nu].conj().T)
            acc = np.zeros(V, complex)
            for (s1, a, b, A), (s2, c, d, B) in itertools.product(
                    ((1, idx, xp, Afw), (-1, xp, idx, Abw)), ((1, ys, yp, Bfw), (-1, yp, ys, Bbw))):
                Sbc = S[b, :, c, :]          # [V,12,12]
                Sda = S[d, :, a, :]          # [V,12,12] (d fixed, a varies)
                acc += s1 * s2 * np.einsum("vab,vbc,cd,vda->v", A, Sbc, B, Sda)
            out[:, 4 * m + nu] = -0.25 * acc
    return out


def test_brute_force_definition_on_a_rough_field(orc, chip):
    dims = (4, 4, 4, 4)
    kappa, mu, bc = 0.15, 0.08, (1.0, 0.3, 0.0, -0.5)
    src = (1, 3, 0, 2)
    g = chip.gauge_random(dims, 5, 0.3)
    Lu = orc.Lattice(dims, kappa, mu, bc)
    V = Lu.V
    Su = np.linalg.inv(dense_operator(Lu, g)).reshape(V, 12, V, 12)
    Sd = np.linalg.inv(dense_operator(Lu, g, -mu)).reshape(V, 12, V, 12)
    props = []
    for S in (Su, Sd):
        for s in range(5):
            y = site(dims, [src[m] + (1 if s == m else 0) for m in range(4)])
            for a in range(12):
                props.append(to_real(S[:, :, y, a]))
    conn, ct = Lu.cvc_contract(g, props, src)
    got = conn[..., 0] + 1j * conn[..., 1]
    ph = np.array([np.exp(1j * np.pi * bc[m] / dims[m]) for m in range(4)])
    U = gauge_c(g) * ph[None, :, None, None]
    want = wick_conn(dims, U, Su, src)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
    # the loop of the down quark differs on a single rough configuration
    other = wick_conn(dims, U, Sd, src)
    assert np.abs(other - want).max() > 1e-3 * np.abs(want).max()

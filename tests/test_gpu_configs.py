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


def piq_errors(piq, ref, dims, qmax=2.0):
    """Per-momentum error of Pi_mu_nu(q) for every q with 0 < |q| <= qmax: max over
    (mu, nu) of |Pi - Pi_ref| over the largest |Pi_ref_mu_nu(q)| of that q.  Also the
    error at q = 0, where the Ward identity makes Pi(0) vanish up to rounding (so
    it is measured against the largest |Pi_ref(q)| of all q instead)."""
    z = (piq[..., 0] + 1j * piq[..., 1]).reshape(dims + (16,))
    zr = (ref[..., 0] + 1j * ref[..., 1]).reshape(dims + (16,))
    n = np.stack(np.meshgrid(*[np.fft.fftfreq(L, 1.0 / L) for L in dims], indexing="ij"), -1)
    qq = ((2 * np.pi * n / np.array(dims)) ** 2).sum(-1)
    sel = (qq <= qmax * qmax + 1e-12) & (qq > 0)
    err = np.abs(z - zr).max(-1)[sel] / np.abs(zr).max(-1)[sel]
    err0 = np.abs(z - zr)[(0,) * 4].max() / np.abs(zr).max()
    return err, int(sel.sum()), err0


@pytest.mark.timeout(900)
def test_piq_mixed_precision_24x48(chip):
    """24^3x48, kappa 0.1632, mu 0.004, gauge seed 3: Pi_mu_nu(q) from 120 propagators
    solved with FP32 and FP16+norm inner solvers (FP64 reliable updates) against the
    FP64 solves, per q with |q| <= 2, all to |r|^2/|b|^2 <= 1e-20.  The FP64 GPU path
    is the oracle's equal on the same propagators (test_contraction_16x32)."""
    dims = (48, 24, 24, 24)
    src = (0, 0, 0, 0)
    print("24^3x48 Pi(q) mixed vs FP64", flush=True)
    g = chip.gauge_random(dims, 3, 0.3)
    ctx = chip.Context(dims, 0.1632, 0.004, (1.0, 0.0, 0.0, 0.0), gauge=g)
    res = {}
    for prec, delta in ((chip.PREC_F64, 0.1), (chip.PREC_F32, 0.01), (chip.PREC_F16, 0.1)):
        conn, ct, st = ctx.cvc_source(src, 1e-20, 4000, prec, delta)
        res[prec] = ctx.pi_q(conn, ct, src)
        print(f"prec {prec}: {st['seconds']:.2f} s of solves, {st['iterations']} iterations summed, "
              f"{st['reliable_updates']} reliable updates, max true |r|^2/|b|^2 {st['true_res2_max']:.2e}", flush=True)
    for prec in (chip.PREC_F32, chip.PREC_F16):
        err, nq, err0 = piq_errors(res[prec], res[chip.PREC_F64], dims)
        print(f"prec {prec}: per-q relative error of Pi_mu_nu(q) over {nq} momenta 0 < |q| <= 2: "
              f"max {err.max():.2e}, median {np.median(err):.2e}; q = 0: {err0:.2e} of max |Pi|", flush=True)
        assert err.max() < 1e-7
        assert err0 < 1e-9
        # the scalar vacuum polarisation on qhat^2 orbits, qhat^2 <= 4
        q2, p64, _, cnt = chip.piq_orbits(dims, res[chip.PREC_F64])
        q2m, pm, _, _ = chip.piq_orbits(dims, res[prec])
        sel = q2 <= 4.0
        assert np.array_equal(q2, q2m)
        eo = np.abs(pm[sel] - p64[sel]) / np.abs(p64[sel])
        print(f"prec {prec}: Pi(qhat^2) on {int(sel.sum())} orbits qhat^2 <= 4 ({int(cnt[sel].sum())} momenta): "
              f"max relative error {eo.max():.2e}", flush=True)
        assert eo.max() < 1e-7


@pytest.mark.timeout(900)
def test_gradient_16x32(chip):
    """16^3x32, kappa 0.1632, mu 0.004, gauge seed 2: the gradient of
    L = sum_mu sum_{|q|<=2} Re Pi_mu_mu(q) by the adjoint kernels against central
    finite differences (h = 1e-5) on mu and on 16 random link directions plus the
    four source links, every L from FP64 solves to |r|^2/|b|^2 <= 1e-24 (even/odd
    preconditioned); the FP32-inner gradient (FP64 accumulation) against the FP64 one."""
    from cvc_amd import autograd as ag
    dims = (32, 16, 16, 16)
    src = (5, 3, 7, 1)
    kappa, mu, qmax, sp = 0.1632, 0.004, 2.0, 1e-24
    g = chip.gauge_random(dims, 2, 0.3)
    ctx = chip.Context(dims, kappa, mu, (1.0, 0.0, 0.0, 0.0), gauge=g)
    ctx.set_solver_flags(chip.SOLVE_EO)
    print("16^3x32 gradient FP64", flush=True)
    L, dmu, F, st = ctx.cvc_gradient(src, qmax, sp, 5000)
    print(f"L = {L:.12e}, dL/dmu = {dmu:.12e}, {st['seconds']:.1f} s of solves", flush=True)

    def loss(gg=None, m=mu):
        ctx.set_params(kappa, m)
        ctx.set_gauge(g if gg is None else gg)
        conn, ct, _ = ctx.cvc_source(src, sp, 5000)
        return ag.loss_from_conn(conn, ct, dims, src, qmax)

    h = 1e-5
    fd = (loss(m=mu + h) - loss(m=mu - h)) / (2 * h)
    print(f"dL/dmu adjoint {dmu:.10e} fd {fd:.10e}", flush=True)
    assert abs(dmu - fd) < 1e-6 * abs(fd) + 1e-9
    U0 = ag.to_complex(g)
    V = int(np.prod(dims))
    y = ((src[0] * dims[1] + src[1]) * dims[2] + src[2]) * dims[3] + src[3]
    rng = np.random.default_rng(23)
    dirs = [(int(rng.integers(V)), int(rng.integers(4))) for _ in range(16)] + [(y, r) for r in range(4)]
    scale = np.max(np.abs(F))
    for z, rho in dirs:
        a = rng.standard_normal(8)
        T = np.einsum("a,aij->ij", a / np.linalg.norm(a), ag.GELL_MANN) / 2
        vals = []
        for eps in (h, -h):
            U = U0.copy()
            U[z, rho] = ag.expi(eps * T) @ U[z, rho]
            vals.append(loss(ag.to_real(U)))
        fdl = (vals[0] - vals[1]) / (2 * h)
        an = np.trace(T @ F[z, rho]).real
        print(f"link ({z},{rho}): adjoint {an:.10e} fd {fdl:.10e}", flush=True)
        assert abs(an - fdl) < 1e-6 * scale, (z, rho, an, fdl)
    ctx.set_params(kappa, mu)
    ctx.set_gauge(g)
    L32, dmu32, F32, st32 = ctx.cvc_gradient(src, qmax, sp, 5000, prec=chip.PREC_F32, delta=0.01)
    print(f"FP32 inner: L rel {abs(L32 - L) / abs(L):.2e}, dmu rel {abs(dmu32 - dmu) / abs(dmu):.2e}, "
          f"F rel {rel_err(F32, F):.2e}, {st32['reliable_updates']} reliable updates", flush=True)
    assert st32["reliable_updates"] > 0
    assert abs(L32 - L) < 1e-9 * abs(L)
    assert abs(dmu32 - dmu) < 1e-7 * abs(dmu)
    assert rel_err(F32, F) < 1e-7


@pytest.mark.timeout(900)
def test_config_32x64_48rhs_mixed(chip, orc):
    """32^3x64, kappa 0.1632, mu 0.004, gauge seed 4: 48 rhs (12 spin-colour x 4 source
    positions drawn from seed 4) solved together in one batched gamma5-hermitian BiCGStab
    (one operator launch per application for all 48, links shared), FP32 inner iteration
    with FP64 reliable updates (delta 0.01), to |r|^2/|b|^2 <= 1e-20.  Every rhs meets the
    tolerance by the library's FP64 true residual; two are checked with the oracle's
    operator.  The four sources' pion correlators from the batched propagators: source 0
    equals the oracle's contract_twopoint of the same propagators to 1e-10, and every
    correlator is real and positive (C(t) = sum |S|^2)."""
    dims = (64, 32, 32, 32)
    V = int(np.prod(dims))
    rng = np.random.default_rng(4)
    srcs = [tuple(int(rng.integers(0, L)) for L in dims) for _ in range(4)]
    print(f"32^3x64 48-rhs mixed solve, sources {srcs}", flush=True)
    g = chip.gauge_random(dims, 4, 0.3)
    ctx = chip.Context(dims, 0.1632, 0.004, gauge=g)
    b, x = ctx.field(48), ctx.field(48)
    b.point_sources(srcs)
    x.zero()
    rc, st = ctx.invert_her(x, b, 1e-20, 4000, prec=chip.PREC_F32, delta=0.01)
    print(f"solve rc={rc} {st}", flush=True)
    assert rc > 0 and st["converged"] == 48 and st["true_res2_max"] <= 1e-20
    L = orc.Lattice(dims, 0.1632, 0.004)
    for col in (0, 47):
        k, a = divmod(col, 12)
        t, xx, y, z = srcs[k]
        bb = np.zeros((V, 24))
        bb[((t * dims[1] + xx) * dims[2] + y) * dims[3] + z, 2 * a] = 1.0
        assert np.array_equal(b.download(col, 1)[0], bb)
        r = L.Q_phi(g, x.download(col, 1)[0]) - bb
        assert L.dot_re(r, r) < 1.01e-20
    one = ctx.field(12)
    for k in range(4):
        one.copy_columns(0, x, 12 * k, 12)
        c = ctx.field_contract_twopoint(5, 5, one, one, src=srcs[k])
        assert np.all(c.real > 0) and np.max(np.abs(c.imag)) < 1e-12 * np.max(c.real)
        if k == 0:
            props = list(one.download())
            co = L.contract_twopoint(5, 5, props, props, src=srcs[0])
            assert rel_err(c, co) < 1e-10
            del props

"""Parity of the HIP contraction, Ward identity, Pi(q) and whole-source
pipeline with the oracle (GPU).

conn and the contact term from the same propagators agree with the oracle's
loop-for-loop restatement of cvc_exact2_xspace.c:577-861 to 1e-12 relative;
the whole-source pipeline (its own 120 solves) agrees to 1e-9 relative with
the oracle contraction of separately solved propagators (both at a squared
residual of 1e-26).  Pi(q) is checked against a brute-force numpy DFT and
for transversality.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

BC = (1.0, 0.0, 0.0, 0.0)


def site(dims, c):
    T, LX, LY, LZ = dims
    return (((c[0] % T) * LX + (c[1] % LX)) * LY + (c[2] % LY)) * LZ + (c[3] % LZ)


def gpu_props(chip, ctx, dims, src, sp=1e-26):
    """120 propagators in the reference's g_spinor_field order, solved on the GPU."""
    V = ctx.V
    props = [None] * 120
    mu0 = ctx.mu
    for s in range(5):
        d = [int(s == m) for m in range(4)]
        ys = site(dims, [src[m] + d[m] for m in range(4)])
        b = np.stack([chip.point_source(V, ys, a // 3, a % 3) for a in range(12)])
        for fl, mu in ((0, mu0), (1, -mu0)):
            ctx.set_params(ctx.kappa, mu)
            rc, x, st = ctx.invert_Qtm_her(b, sp, 5000)
            assert st["converged"] == 12
            for a in range(12):
                props[60 * fl + 12 * s + a] = x[a]
    ctx.set_params(ctx.kappa, mu0)
    return props


CASES = [((4, 4, 4, 4), (1, 2, 0, 3)), ((6, 4, 4, 8), (5, 0, 3, 7))]


@pytest.fixture(scope="module", params=CASES, ids=["4^4", "6x4x4x8"])
def setup(request, chip, orc):
    dims, src = request.param
    print(f"contraction setup dims={dims} src={src}", flush=True)
    g = chip.gauge_random(dims, 21, 0.3)
    ctx = chip.Context(dims, 0.163, 0.01, BC, gauge=g)
    L = orc.Lattice(dims, 0.163, 0.01, BC)
    props = gpu_props(chip, ctx, dims, src)
    conn_o, ct_o = L.cvc_contract(g, props, src)
    return dict(dims=dims, src=src, g=g, ctx=ctx, L=L, props=props, conn_o=conn_o, ct_o=ct_o)


def test_contract_matches_oracle(setup):
    ctx, props, src = setup["ctx"], setup["props"], setup["src"]
    conn, ct = ctx.cvc_contract(props, src)
    assert rel_err(conn, setup["conn_o"]) < 1e-12
    assert np.max(np.abs(ct - setup["ct_o"])) < 1e-12 * np.max(np.abs(ct))


def test_ward_identity(setup):
    ctx, L, dims, src = setup["ctx"], setup["L"], setup["dims"], setup["src"]
    conn, ct = setup["conn_o"], setup["ct_o"]
    wi = ctx.wi(conn)
    assert rel_err(wi, L.wi(conn)) < 1e-13
    w = wi[..., 0] + 1j * wi[..., 1]
    want = np.zeros_like(w)
    y = site(dims, src)
    want[y, :] += ct
    for nu in range(4):
        d = [int(nu == m) for m in range(4)]
        want[site(dims, [src[m] + d[m] for m in range(4)]), nu] -= ct[nu]
    assert np.max(np.abs(w - want)) < 1e-12 * np.max(np.abs(conn))


def test_pi_q(setup):
    ctx, dims, src = setup["ctx"], setup["dims"], setup["src"]
    conn, ct = setup["conn_o"], setup["ct_o"]
    piq = ctx.pi_q(conn, ct, src)
    P = piq[..., 0] + 1j * piq[..., 1]
    # brute force: Pi(q) = sum_x e^{-iq(x-y)} e^{-i(q_mu-q_nu)/2} conn(x) - delta contact
    c = (conn[..., 0] + 1j * conn[..., 1]).reshape(dims + (4, 4))
    F = np.fft.fftn(c, axes=(0, 1, 2, 3))
    ns = np.meshgrid(*[np.arange(L) for L in dims], indexing="ij")
    q = [2 * np.pi * ns[d] / dims[d] for d in range(4)]
    ey = np.exp(1j * sum(q[d] * src[d] for d in range(4)))
    want = np.empty_like(F)
    for mu in range(4):
        for nu in range(4):
            want[..., mu, nu] = ey * np.exp(-0.5j * (q[mu] - q[nu])) * F[..., mu, nu] - (mu == nu) * ct[nu]
    want = want.reshape(-1, 16)
    assert np.max(np.abs(P - want)) < 1e-12 * np.max(np.abs(want))
    # transversality: sum_mu qhat_mu Pi_{mu nu}(q) = 0
    qh = np.stack([2 * np.sin(qd / 2) for qd in q], axis=-1).reshape(-1, 4)
    Pm = P.reshape(-1, 4, 4)
    div = np.einsum("vm,vmn->vn", qh, Pm)
    assert np.max(np.abs(div)) < 1e-12 * np.max(np.abs(Pm))


def test_whole_source_pipeline(setup, chip):
    ctx, src = setup["ctx"], setup["src"]
    print("cvc_source pipeline", flush=True)
    conn, ct, st = ctx.cvc_source(src, 1e-26, 5000)
    assert st["converged"] == 120
    assert rel_err(conn, setup["conn_o"]) < 1e-9
    assert np.max(np.abs(ct - setup["ct_o"])) < 1e-9 * np.max(np.abs(ct))
    # mixed precision pipeline: per-component error against the FP64 result
    conn32, ct32, st32 = ctx.cvc_source(src, 1e-26, 5000, prec=chip.PREC_F32, delta=0.1)
    assert st32["converged"] == 120
    assert rel_err(conn32, setup["conn_o"]) < 1e-9

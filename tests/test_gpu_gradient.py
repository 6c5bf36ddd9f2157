"""Adjoint gradient of L = sum_mu sum_{|q|<=2} Re Pi_{mu mu}(q) (GPU).

Checks: L from the gradient call equals the sum over Pi(q) of the oracle's
contraction; dL/dmu and tr(T F) on 16 random link directions plus the four
source links agree with central finite differences (h = 1e-5, forward solves
to a squared residual of 1e-26, agreement to 1e-6 relative); the FP32-inner
gradient agrees with the FP64 one; the torch.autograd harness returns the same
gradients.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

DIMS = (4, 4, 4, 8)
SRC = (1, 2, 0, 3)
KAPPA, MU, QMAX = 0.163, 0.01, 2.0
SP = 1e-26


@pytest.fixture(scope="module")
def prob(chip):
    from cvc_amd import autograd as ag
    g = chip.gauge_random(DIMS, 31, 0.3)
    ctx = chip.Context(DIMS, KAPPA, MU, (1.0, 0.0, 0.0, 0.0), gauge=g)
    print("gradient f64", flush=True)
    L, dmu, F, st = ctx.cvc_gradient(SRC, QMAX, SP, 5000)
    return dict(g=g, ctx=ctx, L=L, dmu=dmu, F=F, ag=ag)


def forward_L(prob, g=None, mu=MU):
    ctx, ag = prob["ctx"], prob["ag"]
    ctx.set_params(KAPPA, mu)
    ctx.set_gauge(prob["g"] if g is None else g)
    conn, ct, st = ctx.cvc_source(SRC, SP, 5000)
    ctx.set_params(KAPPA, MU)
    ctx.set_gauge(prob["g"])
    return ag.loss_from_conn(conn, ct, DIMS, SRC, QMAX)


def test_loss_value(prob, orc, chip):
    from test_gpu_contract import gpu_props
    ctx = prob["ctx"]
    props = gpu_props(chip, ctx, DIMS, SRC, SP)
    L_o = orc.Lattice(DIMS, KAPPA, MU)
    conn, ct = L_o.cvc_contract(prob["g"], props, SRC)
    piq = ctx.pi_q(conn, ct, SRC)
    q = prob["ag"].momenta(DIMS, QMAX)
    # sum of Re Pi_mumu over the momentum set from the DFT output
    n = np.round(q * np.array(DIMS) / (2 * np.pi)).astype(int) % np.array(DIMS)
    idx = ((n[:, 0] * DIMS[1] + n[:, 1]) * DIMS[2] + n[:, 2]) * DIMS[3] + n[:, 3]
    want = sum(piq[i, 5 * m, 0] for i in idx for m in range(4))
    assert abs(prob["L"] - want) < 1e-10 * abs(want)


def test_dmu_finite_difference(prob):
    h = 1e-5
    fd = (forward_L(prob, mu=MU + h) - forward_L(prob, mu=MU - h)) / (2 * h)
    print("dL/dmu", prob["dmu"], "fd", fd)
    assert abs(prob["dmu"] - fd) < 1e-6 * abs(fd) + 1e-9


def perturbed(prob, z, rho, T, eps):
    ag = prob["ag"]
    U = ag.to_complex(prob["g"]).copy()
    U[z, rho] = ag.expi(eps * T) @ U[z, rho]
    return ag.to_real(U)


def test_link_finite_differences(prob):
    ag = prob["ag"]
    rng = np.random.default_rng(17)
    V = int(np.prod(DIMS))
    y = ((SRC[0] * DIMS[1] + SRC[1]) * DIMS[2] + SRC[2]) * DIMS[3] + SRC[3]
    dirs = [(int(rng.integers(V)), int(rng.integers(4))) for _ in range(16)] + [(y, r) for r in range(4)]
    h = 1e-5
    F = prob["F"]
    scale = np.max(np.abs(F))
    for z, rho in dirs:
        a = rng.standard_normal(8)
        T = np.einsum("a,aij->ij", a / np.linalg.norm(a), ag.GELL_MANN) / 2
        fd = (forward_L(prob, perturbed(prob, z, rho, T, h)) - forward_L(prob, perturbed(prob, z, rho, T, -h))) / (2 * h)
        an = np.trace(T @ F[z, rho]).real
        print(z, rho, an, fd)
        assert abs(an - fd) < 1e-6 * scale, (z, rho, an, fd)
    # F is Hermitian and traceless
    assert np.abs(F - np.conj(np.swapaxes(F, -1, -2))).max() < 1e-12 * scale
    assert np.abs(np.trace(F, axis1=-2, axis2=-1)).max() < 1e-12 * scale


def test_fp32_inner_gradient(prob, chip):
    ctx = prob["ctx"]
    L, dmu, F, st = ctx.cvc_gradient(SRC, QMAX, SP, 5000, prec=chip.PREC_F32, delta=0.1)
    assert st["reliable_updates"] > 0
    assert abs(L - prob["L"]) < 1e-9 * abs(prob["L"])
    assert abs(dmu - prob["dmu"]) < 1e-7 * abs(prob["dmu"])
    assert rel_err(F, prob["F"]) < 1e-7


def test_autograd_harness(prob, chip):
    import torch
    ag = prob["ag"]
    problem = dict(ctx=prob["ctx"], U0=ag.to_complex(prob["g"]), src=SRC, qmax=QMAX, sp=SP, niter=5000,
                   prec=chip.PREC_F64, delta=0.1)
    mu = torch.tensor(MU, dtype=torch.float64, requires_grad=True)
    om = torch.zeros((int(np.prod(DIMS)), 4, 8), dtype=torch.float64, requires_grad=True)
    L = ag.CvcLoss.apply(mu, om, problem)
    L.backward()
    assert abs(L.item() - prob["L"]) < 1e-10 * abs(prob["L"])
    assert abs(mu.grad.item() - prob["dmu"]) < 1e-10 * abs(prob["dmu"])
    want = np.einsum("aij,vrji->vra", ag.GELL_MANN / 2, prob["F"]).real
    assert np.abs(om.grad.numpy() - want).max() < 1e-10 * np.abs(want).max()

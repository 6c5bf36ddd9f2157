"""torch.autograd harness around the hand-written adjoint (cvchip_cvc_gradient).

The scalar L(mu, omega) = sum_mu sum_{|q|<=qmax} Re Pi_{mu mu}(q) is computed on
the GPU for gauge links U_rho(x) = exp(i sum_a omega_a(x, rho) lambda_a / 2) U0_rho(x)
(lambda_a the Gell-Mann matrices, omega real [V, 4, 8]).  Its backward pass calls
the adjoint kernels once and maps the link force F (dL/d eps = tr(T F) for
U -> exp(i eps T) U) to omega through the differential of exp:
dL/d omega_a = tr(G_a F), G_a = sum_n ad_{iX}^n(lambda_a/2)/(n+1)!, X = omega.lambda/2.
The forward solves are recomputed in the backward pass rather than stored.
"""
from __future__ import annotations

import numpy as np
import torch

import cvc_amd

GELL_MANN = np.zeros((8, 3, 3), complex)
GELL_MANN[0][0, 1] = GELL_MANN[0][1, 0] = 1
GELL_MANN[1][0, 1], GELL_MANN[1][1, 0] = -1j, 1j
GELL_MANN[2][0, 0], GELL_MANN[2][1, 1] = 1, -1
GELL_MANN[3][0, 2] = GELL_MANN[3][2, 0] = 1
GELL_MANN[4][0, 2], GELL_MANN[4][2, 0] = -1j, 1j
GELL_MANN[5][1, 2] = GELL_MANN[5][2, 1] = 1
GELL_MANN[6][1, 2], GELL_MANN[6][2, 1] = -1j, 1j
GELL_MANN[7] = np.diag([1, 1, -2]) / np.sqrt(3)


def momenta(dims, qmax):
    """The momentum set Q = {q : |q| <= qmax}, n_mu in (-L/2, L/2]."""
    rng = [np.arange(-(L - 1) // 2, L // 2 + 1) for L in dims]
    n = np.stack(np.meshgrid(*rng, indexing="ij"), -1).reshape(-1, 4)
    q = 2 * np.pi * n / np.array(dims)
    return q[(q * q).sum(1) <= qmax * qmax + 1e-12]


def momentum_weight(dims, src, qmax):
    """c(x) = sum_{q in Q} cos q.(x - y) for all sites x (flat, lexicographic), and |Q|,
    by one inverse FFT of the indicator of Q on the momentum grid."""
    n = np.meshgrid(*[np.fft.fftfreq(L, 1.0 / L) for L in dims], indexing="ij")
    qq = sum((2 * np.pi * nd / L) ** 2 for nd, L in zip(n, dims))
    mask = (qq <= qmax * qmax + 1e-12).astype(np.float64)
    w = np.fft.ifftn(mask).real * mask.size           # w[d] = sum_q cos(q.d)
    w = np.roll(w, shift=tuple(src), axis=(0, 1, 2, 3))   # w[x] -> sum_q cos(q.(x - y))
    return w.reshape(-1), int(mask.sum())


def loss_from_conn(conn, contact, dims, src, qmax):
    """L = sum_mu sum_{q in Q} Re Pi_{mu mu}(q) from conn[V,16,2] and the contact term."""
    w, nq = momentum_weight(dims, src, qmax)
    diag = conn[:, [0, 5, 10, 15], 0].sum(1)
    return float((w * diag).sum() - nq * np.real(np.sum(contact)))


def expi(H):
    """exp(i H) for Hermitian 3x3 matrices [..., 3, 3]."""
    e, v = np.linalg.eigh(H)
    return (v * np.exp(1j * e)[..., None, :]) @ np.conj(np.swapaxes(v, -1, -2))


def links(U0, omega):
    """U = exp(i omega.lambda/2) U0 (complex [V,4,3,3])."""
    X = np.einsum("vra,aij->vrij", omega, GELL_MANN) / 2
    return expi(X) @ U0


def dexp_generators(omega, nterms=14):
    """G_a(x, rho) = sum_n ad_{iX}^n (lambda_a / 2) / (n+1)!  -> [V, 4, 8, 3, 3]."""
    X = np.einsum("vra,aij->vrij", omega, GELL_MANN) / 2
    out = np.empty(omega.shape[:2] + (8, 3, 3), complex)
    for a in range(8):
        term = np.broadcast_to(GELL_MANN[a] / 2, X.shape).copy()
        acc = term.copy()
        fact = 1.0
        for n in range(1, nterms):
            term = 1j * (X @ term - term @ X)
            fact *= n + 1
            acc = acc + term / fact
        out[:, :, a] = acc
    return out


def to_real(U):
    z = U.reshape(U.shape[0], 4, 9)
    g = np.empty((U.shape[0], 4, 18))
    g[..., 0::2] = z.real
    g[..., 1::2] = z.imag
    return g


def to_complex(g):
    return (g[..., 0::2] + 1j * g[..., 1::2]).reshape(g.shape[0], 4, 3, 3)


class CvcLoss(torch.autograd.Function):
    """L(mu, omega) on the GPU.  The forward pass solves the 120 propagators and
    contracts; the backward pass recomputes them inside the adjoint call (the
    propagators are not kept between the passes) and returns dL/dmu and
    dL/domega (FP64 accumulation, inner solves in the precision `prec`)."""

    @staticmethod
    def _set(problem, mu_v, om):
        ctx = problem["ctx"]
        ctx.set_params(ctx.kappa, mu_v)
        ctx.set_gauge(to_real(links(problem["U0"], om)))
        return ctx

    @staticmethod
    def forward(fctx, mu, omega, problem):
        mu_v = float(mu.detach().cpu())
        om = omega.detach().cpu().double().numpy()
        ctx = CvcLoss._set(problem, mu_v, om)
        conn, contact, _ = ctx.cvc_source(problem["src"], problem["sp"], problem["niter"], problem["prec"],
                                          problem["delta"])
        fctx.problem = problem
        fctx.save_for_backward(mu.detach(), omega.detach())
        return torch.tensor(loss_from_conn(conn, contact, ctx.dims, problem["src"], problem["qmax"]),
                            dtype=torch.float64)

    @staticmethod
    def backward(fctx, gL):
        mu, omega = fctx.saved_tensors
        problem = fctx.problem
        om = omega.cpu().double().numpy()
        ctx = CvcLoss._set(problem, float(mu.cpu()), om)
        _, dmu, F, _ = ctx.cvc_gradient(problem["src"], problem["qmax"], problem["sp"], problem["niter"],
                                        problem["prec"], problem["delta"], force=True)
        G = dexp_generators(om)
        dom = np.einsum("vrakl,vrlk->vra", G, F).real
        return gL * torch.tensor(dmu, dtype=torch.float64), gL * torch.from_numpy(dom), None

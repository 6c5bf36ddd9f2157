"""Edge cases of the batched solvers (GPU): a zero right-hand side inside a batch
(the reference returns at once when the start vector solves the system,
invert_Qtm.c:138-144, so that column's solution stays exactly zero and the others
are unaffected), a column whose initial guess is already the solution, a batch whose
columns differ in scale by 1e12, and an all-zero batch, for the FP64 and mixed solvers,
with and without even/odd preconditioning.  Solutions are checked with the oracle's
operator."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = (8, 4, 4, 8)
KAPPA, MU = 0.16, 0.01


@pytest.fixture(scope="module")
def prob(chip, orc):
    g = chip.gauge_random(DIMS, 5, 0.3)
    ctx = chip.Context(DIMS, KAPPA, MU, gauge=g)
    L = orc.Lattice(DIMS, KAPPA, MU)
    return g, ctx, L


def res2(L, g, x, b):
    r = L.Q_phi(g, x) - b
    return L.dot_re(r, r) / L.dot_re(b, b)


@pytest.mark.parametrize("eo", [False, True])
@pytest.mark.parametrize("prec", ["f64", "f32", "f16"])
def test_zero_and_scaled_columns(chip, prob, prec, eo):
    g, ctx, L = prob
    V = ctx.V
    P = {"f64": chip.PREC_F64, "f32": chip.PREC_F32, "f16": chip.PREC_F16}[prec]
    print(f"edge: zero/scaled columns prec={prec} eo={eo}", flush=True)
    rng = np.random.default_rng(8)
    b = np.zeros((4, V, 24))
    b[0] = chip.point_source(V, 0, 1, 2)
    b[2] = rng.standard_normal((V, 24))
    b[3] = 1e-12 * chip.point_source(V, 5, 3, 0)
    rc, x, st = ctx.invert_Qtm_her(b, 1e-20, 4000, prec=P, delta=0.1 if prec == "f16" else 0.01, eo=eo)
    assert rc > 0 and st["converged"] == 4, (rc, st)
    assert np.all(np.isfinite(x))
    assert not np.any(x[1])                       # the zero column's solution is exactly zero
    for k in (0, 2, 3):
        assert res2(L, g, x[k], b[k]) < 1.01e-20, k


@pytest.mark.parametrize("eo", [False, True])
def test_solved_start_column(chip, orc, eo):
    """invert_Qtm (kappa 0.14, where the plain BiCGStab converges) with the initial guess of
    one column already its solution: that column is returned unchanged, the others solve."""
    g = chip.gauge_random(DIMS, 5, 0.3)
    ctx = chip.Context(DIMS, 0.14, MU, gauge=g)
    L = orc.Lattice(DIMS, 0.14, MU)
    V = ctx.V
    print(f"edge: solved start column eo={eo}", flush=True)
    b = np.stack([chip.point_source(V, 0, a, 0) for a in range(3)])
    rc, x, _ = ctx.invert_Qtm(b, 1e-24, 4000, eo=eo)
    assert rc > 0
    x0 = np.zeros_like(x)
    x0[1] = x[1]
    rc2, x2, st = ctx.invert_Qtm(b, 1e-20, 4000, x0=x0, eo=eo)
    assert rc2 > 0 and st["converged"] == 3
    assert np.array_equal(x2[1], x[1])
    for k in range(3):
        assert res2(L, g, x2[k], b[k]) < 1.01e-20


def test_all_zero_batch(chip, prob):
    g, ctx, L = prob
    b = np.zeros((2, ctx.V, 24))
    for P in (chip.PREC_F64, chip.PREC_F32, chip.PREC_F16):
        rc, x, st = ctx.invert_Qtm_her(b, 1e-20, 100, prec=P, delta=0.1)
        assert rc == 0 and not np.any(x)
        rc, x, st = ctx.invert_Qtm(b, 1e-20, 100, prec=P, delta=0.1)
        assert rc == 0 and not np.any(x)

"""Even/odd-preconditioned solves (GPU): the Schur complement on the odd
sites, M = A_oo - H_oe A_ee^-1 H_eo / 4, with A = 1/(2 kappa) + i mu gamma5 and
H the hopping sum of Q_phi.c:58-174 (D = A - H/2).

The preconditioned solver must return a solution of the reference's own
system D x = b: every solution is checked with the oracle's operator
(|D x - b|^2/|b|^2 below the requested squared tolerance) and against the
unpreconditioned FP64 solve (1e-8 relative, both to 1e-24).  A wrong Schur
complement, sign or parity map cannot pass these checks.
"""
import numpy as np
import pytest

from conftest import rand_spinor, rel_err

pytestmark = pytest.mark.gpu

BC = (1.0, 0.0, 0.0, 0.0)


@pytest.fixture(scope="module")
def lat(chip, orc):
    dims = (8, 4, 6, 8)
    g = chip.gauge_random(dims, 9, 0.3)
    return dims, g


@pytest.mark.parametrize("her", [False, True])
@pytest.mark.parametrize("prec", ["f64", "f32", "f16"])
def test_eo_solution_is_the_reference_solution(chip, orc, lat, her, prec):
    dims, g = lat
    print(f"eo solve her={her} prec={prec}", flush=True)
    P = {"f64": chip.PREC_F64, "f32": chip.PREC_F32, "f16": chip.PREC_F16}[prec]
    # plain BiCGStab on D (or M) needs a spectrum in the right half plane: kappa 0.14
    # (as in test_invert_plain_matches_oracle); the gamma5-hermitian system runs near
    # the critical kappa
    kappa, mu = (0.163, 0.01) if her else (0.14, 0.02)
    ctx = chip.Context(dims, kappa, mu, BC, gauge=g)
    L = orc.Lattice(dims, kappa, mu, BC)
    V = ctx.V
    rng = np.random.default_rng(5)
    b = np.concatenate([rand_spinor(rng, V, 2), np.stack([chip.point_source(V, 7, 1, 2)])])
    solve = ctx.invert_Qtm_her if her else ctx.invert_Qtm
    rc, x, st = solve(b, 1e-22, 5000, prec=P, delta=0.1, eo=True)
    assert rc > 0 and st["converged"] == 3
    assert st["true_res2_max"] < 1e-22
    for a in range(3):
        r = L.Q_phi(g, x[a]) - b[a]
        assert L.dot_re(r, r) < 1e-22 * L.dot_re(b[a], b[a])
    if prec == "f64":
        rc0, x0, st0 = solve(b, 1e-24, 5000)
        rc1, x1, st1 = solve(b, 1e-24, 5000, eo=True)
        assert rel_err(x1, x0) < 1e-8
        # the Schur complement is better conditioned: fewer iterations
        assert st1["iterations"] < st0["iterations"]
        print(f"iterations: unpreconditioned {st0['iterations']}, even/odd {st1['iterations']}", flush=True)


def test_eo_initial_guess_and_flags(chip, orc, lat):
    dims, g = lat
    ctx = chip.Context(dims, 0.14, 0.02, BC, gauge=g)
    V = ctx.V
    b = np.stack([chip.point_source(V, 0, 0, 0)])
    rc, x, _ = ctx.invert_Qtm(b, 1e-24, 5000, eo=True)
    # the converged solution as initial guess: the odd system starts converged
    rc2, x2, _ = ctx.invert_Qtm(b, 1e-20, 5000, x0=x, eo=True)
    assert rc2 == 0 and rel_err(x2, x) < 1e-10
    # context-wide flag: the plain call now runs preconditioned
    ctx.set_solver_flags(chip.SOLVE_EO)
    rc3, x3, st3 = ctx.invert_Qtm(b, 1e-24, 5000)
    assert rc3 == rc and rel_err(x3, x) < 1e-12
    ctx.set_solver_flags(0)
    with pytest.raises(chip.CvchipError):
        ctx.set_solver_flags(64)
    odd = chip.Context((4, 4, 4, 5), 0.16, 0.02, BC, gauge=chip.gauge_random((4, 4, 4, 5), 1, 0.3))
    with pytest.raises(chip.CvchipError):
        odd.set_solver_flags(chip.SOLVE_EO)        # LZ must be even


def test_eo_source_pipeline(chip, orc):
    """One whole source through the preconditioned solves: conn equals the
    unpreconditioned pipeline's."""
    dims = (6, 4, 4, 8)
    src = (1, 2, 0, 3)
    g = chip.gauge_random(dims, 4, 0.3)
    ctx = chip.Context(dims, 0.163, 0.01, BC, gauge=g)
    conn0, ct0, _ = ctx.cvc_source(src, 1e-26, 5000)
    ctx.set_solver_flags(chip.SOLVE_EO)
    conn1, ct1, st1 = ctx.cvc_source(src, 1e-26, 5000)
    conn2, ct2, st2 = ctx.cvc_source(src, 1e-22, 5000, prec=chip.PREC_F16, delta=0.1)
    assert st1["converged"] == 120 and st2["converged"] == 120
    assert rel_err(conn1, conn0) < 1e-9
    assert np.max(np.abs(ct1 - ct0)) < 1e-9 * np.max(np.abs(ct0))
    assert rel_err(conn2, conn0) < 1e-7

"""Misuse of the C ABI (GPU context): every case returns its error code with a
message from cvchip_last_error() and leaves the context usable (an operator
call after the errors still gives the oracle's result)."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def test_api_errors_leave_context_usable(chip, orc):
    dims = (4, 4, 4, 5)            # odd LZ: no even/odd preconditioning
    print("API misuse errors", flush=True)
    with pytest.raises(chip.CvchipError):
        chip.Context((4, 0, 4, 4), 0.15, 0.01)
    ctx = chip.Context(dims, 0.15, 0.01)
    V = ctx.V
    phi = np.random.default_rng(2).standard_normal((V, 24))
    with pytest.raises(chip.CvchipError) as e:
        ctx.Q_phi(phi)                              # no gauge field yet
    assert e.value.rc == chip.ERR_STATE
    g = chip.gauge_random(dims, 9, 0.3)
    ctx.set_gauge(g)
    for bad in (0, 257):
        with pytest.raises(chip.CvchipError):
            ctx.field(bad)
    with pytest.raises(chip.CvchipError):
        ctx.field(2, 7)                             # unknown precision
    with pytest.raises(chip.CvchipError):
        ctx.set_link_recon(13)
    a, b = ctx.field(2), ctx.field(3)
    with pytest.raises(chip.CvchipError):
        ctx.dslash(b, a)                            # rhs counts differ
    h = ctx.field(2, chip.PREC_F16)
    with pytest.raises(chip.CvchipError):
        h.upload(np.zeros((2, V, 24)))              # host data enters as FP64 or FP32
    bb = np.zeros((1, V, 24))
    bb[0, 0, 0] = 1.0
    with pytest.raises(chip.CvchipError):
        ctx.invert_Qtm(bb, 1e-20, 100, eo=True)     # odd LZ
    with pytest.raises(chip.CvchipError):
        ctx.invert_Qtm_her(bb, 1e-20, 100, prec=chip.PREC_F32, delta=0.0)
    with pytest.raises(chip.CvchipError):
        a.copy_columns(0, b, 2, 2)                  # source range beyond its 3 columns
    L = orc.Lattice(dims, 0.15, 0.01)
    assert rel_err(ctx.Q_phi(phi), L.Q_phi(g, phi)) < 1e-13

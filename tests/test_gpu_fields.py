"""Device-resident set utilities (GPU): zeroing, point sources of several source
points (cvc_exact2_xspace.c:521-524 convention, one 12-column block per point)
and column copies between sets, checked against host arrays at 4^4."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_zero_point_sources_and_columns(chip):
    dims = (4, 4, 4, 4)
    V = 256
    ctx = chip.Context(dims, 0.15, 0.01, gauge=chip.gauge_random(dims, 1, 0.3))
    print("field utilities 4^4", flush=True)
    f = ctx.field(24)
    f.fill_random(3)
    f.zero()
    assert not np.any(f.download())
    h = ctx.field(12, chip.PREC_F16)
    h.zero()                                   # FP16 sets: halves and scales
    srcs = [(1, 2, 3, 0), (3, 0, 1, 2)]
    f.point_sources(srcs)
    want = np.zeros((24, V, 24))
    for k, (t, x, y, z) in enumerate(srcs):
        for a in range(12):
            want[12 * k + a, ((t * 4 + x) * 4 + y) * 4 + z, 2 * a] = 1.0
    assert np.array_equal(f.download(), want)
    g = ctx.field(12)
    g.fill_random(5)
    before = g.download()
    g.copy_columns(2, f, 14, 5)                # columns 2..6 <- 14..18
    after = g.download()
    assert np.array_equal(after[2:7], want[14:19])
    assert np.array_equal(after[:2], before[:2]) and np.array_equal(after[7:], before[7:])
    with pytest.raises(chip.CvchipError):
        g.copy_columns(10, f, 0, 5)            # beyond the destination
    with pytest.raises(chip.CvchipError):
        f.copy_columns(0, f, 3, 5)             # overlapping ranges of one set
    with pytest.raises(chip.CvchipError):
        g.point_sources(srcs)                  # 12 columns for two points
    with pytest.raises(chip.CvchipError):
        f.point_sources([(4, 0, 0, 0), (0, 0, 0, 0)])   # t outside the lattice

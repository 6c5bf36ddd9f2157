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


def test_large_transfers_round_trip(chip, orc):
    """Host transfers of 16 MB and more stream through two pinned 64 MB buffers
    in chunks: 12 columns at 16^3x32 (302 MB, five chunks, the last one short),
    one column (25 MB, one short chunk) and a column range inside the set come
    back bit for bit, FP32 sets within their rounding; the host-array operator
    writes into a caller-owned output and equals the oracle's."""
    dims = (32, 16, 16, 16)
    V = 32 * 16 ** 3
    print("large host transfers 16^3x32", flush=True)
    g = chip.gauge_random(dims, 2, 0.3)
    ctx = chip.Context(dims, 0.1632, 0.004, gauge=g)
    rng = np.random.default_rng(17)
    host = rng.standard_normal((12, V, 24))
    f = ctx.field(12)
    f.upload(host)
    assert np.array_equal(f.download(), host)
    assert np.array_equal(f.download(5, 3), host[5:8])
    one = ctx.field(1)
    one.upload(host[7])
    assert np.array_equal(one.download()[0], host[7])
    f.upload(host[2:4], rhs0=9)                 # a column range inside the set
    back = f.download()
    assert np.array_equal(back[9:11], host[2:4]) and np.array_equal(back[:9], host[:9])
    lo = ctx.field(12, chip.PREC_F32)
    lo.upload(host)
    assert np.max(np.abs(lo.download() - host)) <= 1e-6 * np.max(np.abs(host))
    out = np.zeros((V, 24))
    ret = ctx.Q_phi(host[0], out=out)
    assert ret is out
    L = orc.Lattice(dims, 0.1632, 0.004, (1.0, 0.0, 0.0, 0.0))
    want = L.Q_phi(g, host[0])
    assert np.max(np.abs(out - want)) / np.max(np.abs(want)) < 1e-13

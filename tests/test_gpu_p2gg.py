"""P -> gamma gamma (p2gg_xspace.c) on the GPU against the oracle.

The sequential sources of the device kernel equal the oracle's restatement of
p2gg_xspace.c:517-548 (pinned by an explicit-matrix brute force in
test_oracle_p2gg.py) for all 16 gammas, both phase signs and column offsets
(1e-15 relative; the phases come from different sin/cos implementations).
The whole source point -- 120 point solves, 60 sequential solves, the
contraction -- agrees with the oracle's own pipeline (tests/golden/golden_p2gg.npz,
every solve to a squared residual of 1e-28 on both sides) to 1e-10 relative in
conn and the contact term, for the up and the down flavour.
"""
import os

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
BC = (1.0, 0.0, 0.0, 0.0)


def test_sequential_sources_match_oracle(chip, orc):
    dims = (6, 4, 4, 8)
    print(f"sequential sources dims={dims}", flush=True)
    ctx = chip.Context(dims, 0.15, 0.02, BC)
    L = orc.Lattice(dims, 0.15, 0.02, BC)
    rng = np.random.default_rng(5)
    cols = rng.standard_normal((24, L.V, 24))
    fin, fout = ctx.field(24), ctx.field(12)
    fin.upload(cols)
    for gid, tseq, mom, sgn, col0 in [(5, 0, (0, 0, 0), 1.0, 0), (0, 5, (1, 0, 0), -1.0, 12), (3, 2, (1, 1, 3), 1.0, 7),
                                      (11, 4, (2, 1, 5), -1.0, 12)] + [(k, 3, (1, 2, 1), 1.0, 12) for k in range(16)]:
        ctx.field_sequential_sources(fout, fin, col0, tseq, gid, mom, sgn)
        got = fout.download()
        for r in (0, 5, 11):
            want = L.seq_source(cols[col0 + r], tseq, gid, mom, sgn)
            assert rel_err(got[r], want) < 1e-15, (gid, tseq, mom, sgn, col0, r)
    with pytest.raises(chip.CvchipError):
        ctx.field_sequential_sources(fout, fin, 13, 0, 5)     # columns beyond the input set
    with pytest.raises(chip.CvchipError):
        ctx.field_sequential_sources(fout, fin, 0, 6, 5)      # time slice beyond T


@pytest.mark.parametrize("key", ["a", "b"])
def test_p2gg_source_matches_oracle_pipeline(chip, key):
    gold = np.load(os.path.join(HERE, "golden", "golden_p2gg.npz"), allow_pickle=False)
    dims = tuple(int(v) for v in gold["dims"])
    src = tuple(int(v) for v in gold[f"src_{key}"])
    print(f"p2gg case {key} dims={dims}", flush=True)
    ctx = chip.Context(dims, 0.163, 0.01, BC, gauge=chip.gauge_random(dims, 1, 0.3))
    conn, ct, st = ctx.p2gg_source(src, int(gold[f"dt_{key}"]), int(gold[f"gid_{key}"]),
                                   tuple(int(v) for v in gold[f"mom_{key}"]), int(gold[f"pos_{key}"]),
                                   float(gold["sp"]), 20000)
    assert st["converged"] == 180, st
    e_conn = rel_err(conn, gold[f"conn_{key}"])
    e_ct = float(np.max(np.abs(ct - gold[f"contact_{key}"])) / np.max(np.abs(gold[f"contact_{key}"])))
    print(f"  conn {e_conn:.2e} contact {e_ct:.2e}", flush=True)
    assert e_conn < 1e-10 and e_ct < 1e-10

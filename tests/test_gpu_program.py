"""The drop-in contraction program cvc_amd/dropin/cvc_exact2_xspace (GPU):
reference-style input file, ILDG gauge file, output in the LIME format of
write_lime_contraction.  At 4^4 (gauge seed 1, kappa 0.163, mu 0.01, source
(1,2,3,0)) its Pi_mu_nu(x) and contact term equal the golden fixture (the
oracle's loop-for-loop contraction of 120 oracle solves) to 1e-9, its plaquette
the oracle's, and with -w the Ward identity file holds the contact term at the
source and zero elsewhere."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
PROG = os.path.join(os.path.dirname(HERE), "cvc_amd", "dropin", "cvc_exact2_xspace")


@pytest.mark.parametrize("variant", ["f64", "eo_f16"])
def test_program_reproduces_golden(chip, orc, tmp_path, variant):
    gold = np.load(os.path.join(HERE, "golden", "golden_v1.npz"), allow_pickle=False)
    dims = (4, 4, 4, 4)
    g = chip.gauge_random(dims, 1, 0.3)
    chip.lime_write_gauge(tmp_path / "conf.0001", dims, g)
    src = tuple(int(v) for v in gold["conn4_src"])
    sl = ((src[0] * 4 + src[1]) * 4 + src[2]) * 4 + src[3]
    extra = "gpu_precision = 16\ngpu_even_odd = 1\n" if variant == "eo_f16" else ""
    (tmp_path / "cvc.input").write_text(
        "T = 4\nL = 4\nkappa = 0.163\nmu = 0.01\nNconf = 1\n"
        f"gaugefilename_prefix = {tmp_path}/conf\nsource_location = {sl}\n"
        "solver_precision = 1.e-26\nniter_max = 5000\nBCangleT = 1.\n# a comment\nformat = 0\n"
        "read_propagator = no\n" + extra)
    print(f"program {variant}", flush=True)
    p = subprocess.run([PROG, "-f", "cvc.input", "-w"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    out = p.stdout
    assert f"global source coordinates: ({src[0]:3d},{src[1]:3d},{src[2]:3d},{src[3]:3d})" in out
    plaq = float(re.search(r"measured plaquette value:\s+(\S+)", out).group(1))
    assert abs(plaq - orc.Lattice(dims, 0.163, 0.01).plaquette(g)) < 1e-12
    conn, ck = chip.lime_read_contraction(tmp_path / "cvc2_v_x.0001", dims)
    tol = 1e-9 if variant == "f64" else 1e-7
    assert rel_err(conn, gold["conn4"]) < tol
    ct = np.array([[float(v) for v in l.split()[1:3]] for l in out.split("contact term:")[1].strip().splitlines()[:4]])
    assert np.max(np.abs(ct[:, 0] + 1j * ct[:, 1] - gold["contact4"])) < tol * np.max(np.abs(gold["contact4"]))
    assert f"({ck[0]:#x}  {ck[1]:#x})" in out
    # Ward identity file: sum_mu conn(x) - conn(x - mu) = contact (delta_{x,y} - delta_{x,y+nu})
    wi = np.array([[float(v) for v in l.split()[1:3]] for l in (tmp_path / "WI_X.0001.0000").read_text().splitlines()
                   if not l.startswith("#")]).reshape(256, 4, 2)
    w = wi[..., 0] + 1j * wi[..., 1]
    want = np.zeros((256, 4), complex)
    for nu in range(4):
        want[sl, nu] += gold["contact4"][nu]
        c = list(src)
        c[nu] = (c[nu] + 1) % 4
        want[((c[0] * 4 + c[1]) * 4 + c[2]) * 4 + c[3], nu] -= gold["contact4"][nu]
    assert np.max(np.abs(w - want)) < 1e-8 * np.max(np.abs(want))


def prop_name(fmt, s, a, src):
    """get_filename(.., s, a, +1) of cvc_utils.c:2143-2197 at 4^4, prefix "prop." (the default)
    and filename_prefix2 "lat4"."""
    if fmt == 0:
        return f"prop..0001.{s:02d}.{a:02d}.inverted"
    c = list(src)
    if s != 4:
        c[s] = (c[s] + 1) % 4
    return f"prop..0001.t{c[0]:02d}x{c[1]:02d}y{c[2]:02d}z{c[3]:02d}s{a // 3:02d}c{a % 3:02d}.pmass.lat4.inverted"


@pytest.mark.parametrize("fmt", [0, 2])
def test_program_propagator_files(chip, orc, tmp_path, fmt):
    gold = np.load(os.path.join(HERE, "golden", "golden_v1.npz"), allow_pickle=False)
    dims = (4, 4, 4, 4)
    V = 256
    g = chip.gauge_random(dims, 1, 0.3)
    chip.lime_write_gauge(tmp_path / "conf.0001", dims, g)
    src = tuple(int(v) for v in gold["conn4_src"])
    sl = ((src[0] * 4 + src[1]) * 4 + src[2]) * 4 + src[3]
    base = ("T = 4\nL = 4\nkappa = 0.163\nmu = 0.01\nNconf = 1\n"
            f"gaugefilename_prefix = {tmp_path}/conf\nsource_location = {sl}\n"
            f"solver_precision = 1.e-26\nniter_max = 5000\nformat = {fmt}\nfilename_prefix2 = lat4\n")
    (tmp_path / "w.input").write_text(base + "read_propagator = no\nwrite_propagator = yes\n")
    (tmp_path / "r.input").write_text(base)
    print(f"program write_propagator format {fmt}", flush=True)
    p = subprocess.run([PROG, "-f", "w.input"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    conn_w, _ = chip.lime_read_contraction(tmp_path / "cvc2_v_x.0001", dims)
    names = sorted(f.name for f in tmp_path.iterdir() if f.name.endswith(".inverted"))
    assert names == sorted(prop_name(fmt, s, a, src) for s in range(5) for a in range(12))
    (tmp_path / "cvc2_v_x.0001").unlink()
    print(f"program read_propagator format {fmt}", flush=True)
    p = subprocess.run([PROG, "-f", "r.input"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert p.stdout.count("Reading Dirac-fermion field") == 120
    conn_r, _ = chip.lime_read_contraction(tmp_path / "cvc2_v_x.0001", dims)
    assert np.array_equal(conn_r, conn_w)
    assert rel_err(conn_r, gold["conn4"]) < 1e-9
    lat = orc.Lattice(dims, 0.163, 0.01)
    for s, a in ((4, 0), (1, 7), (3, 11)):
        c = list(src)
        if s != 4:
            c[s] = (c[s] + 1) % 4
        b = np.zeros((V, 24))
        b[((c[0] * 4 + c[1]) * 4 + c[2]) * 4 + c[3], 2 * a] = 1.0
        up, _ = chip.lime_read_spinor(tmp_path / prop_name(fmt, s, a, src), dims, 0)
        dn, _ = chip.lime_read_spinor(tmp_path / prop_name(fmt, s, a, src), dims, 1)
        assert np.linalg.norm(lat.Q_phi(g, up, 0.01) - b) < 1e-10
        assert np.linalg.norm(lat.Q_phi(g, dn, -0.01) - b) < 1e-10


def test_cvc_propagators_match_cvc_source(chip):
    """cvchip_cvc_propagators returns the propagators cvchip_cvc_source contracts: contracting
    them with cvchip_cvc_contract gives the same Pi(x) and contact term bit for bit."""
    dims = (4, 4, 4, 4)
    ctx = chip.Context(dims, 0.163, 0.01, gauge=chip.gauge_random(dims, 2, 0.3))
    src = (2, 1, 0, 3)
    props, st = ctx.cvc_propagators(src, 1e-24, 5000)
    assert st["converged"] == 120
    conn_a, ct_a = ctx.cvc_contract(list(props), src)
    conn_b, ct_b, _ = ctx.cvc_source(src, 1e-24, 5000)
    assert np.array_equal(conn_a, conn_b) and np.array_equal(ct_a, ct_b)

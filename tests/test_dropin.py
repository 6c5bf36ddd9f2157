"""The drop-in shim (cvc_amd/dropin/cvc_dropin.cc) exports the reference's
C++ symbols, and a stand-in reference program linked against it computes the
reference's results (GPU part)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_err

OBJ = os.path.join(ROOT, "cvc_amd", "dropin", "cvc_dropin.o")
DRIVER = os.path.join(ROOT, "tests", "dropin", "driver")

# mangled names of the reference entry points (namespace cvc, Q_phi.h / invert_Qtm.h)
REFERENCE_SYMBOLS = [
    "_ZN3cvc9Q_phi_tbcEPdS0_", "_ZN3cvc5Q_phiEPdS0_d", "_ZN3cvc3Qf5EPdS0_d", "_ZN3cvc6g5_phiEPd",
    "_ZN3cvc10invert_QtmEPdS0_i", "_ZN3cvc14invert_Qtm_herEPdS0_i",
    "_ZN3cvc24spinor_scalar_product_reEPdS0_S0_i", "_ZN3cvc24spinor_scalar_product_coEPNS_7complexEPdS2_i",
]


def test_shim_exports_reference_symbols():
    assert os.path.exists(OBJ), "build the package first (python -m cvc_amd.build)"
    out = subprocess.run(["nm", OBJ], capture_output=True, text=True, check=True).stdout
    defined = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert set(REFERENCE_SYMBOLS) <= defined


@pytest.mark.gpu
def test_stand_in_program(tmp_path, chip, orc):
    dims = (4, 4, 4, 8)
    kappa, mu = 0.163, 0.01
    g = chip.gauge_random(dims, 9, 0.3)
    L = orc.Lattice(dims, kappa, mu)
    phi = np.random.default_rng(3).standard_normal((L.V, 24))
    g.tofile(tmp_path / "gauge.bin")
    phi.tofile(tmp_path / "phi.bin")
    r = subprocess.run([DRIVER, str(tmp_path), *map(str, dims), str(kappa), str(mu)], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    qphi = np.fromfile(tmp_path / "qphi.bin").reshape(L.V, 24)
    qf5 = np.fromfile(tmp_path / "qf5.bin").reshape(L.V, 24)
    x = np.fromfile(tmp_path / "x.bin").reshape(L.V, 24)
    assert rel_err(qphi, L.Q_phi(g, phi)) < 1e-13
    assert rel_err(qf5, L.Qf5(g, phi, -mu)) < 1e-13
    ret, ret2, wre, wim = open(tmp_path / "ret.txt").read().split()
    w = L.dot_co(phi, qf5)
    assert abs(complex(float(wre), float(wim)) - w) < 1e-12 * abs(w)
    assert int(ret) > 0 and int(ret2) == -2
    res = L.Q_phi(g, x) - phi
    # lagged iterate (reference semantics): within a factor of the tolerance 1e-20
    assert L.dot_re(res, res) / L.dot_re(phi, phi) < 1e-18
    assert "BiCGStab converged" in r.stdout
    # in-place edits of g_gauge_field between calls reach the device gauge field
    g1 = g.copy().reshape(-1)
    g1[72 * 5 + 18 * 1 + 0] += 1e-3
    e1 = np.fromfile(tmp_path / "qphi_edit.bin").reshape(L.V, 24)
    assert rel_err(e1, L.Q_phi(g1.reshape(g.shape), phi)) < 1e-13
    assert rel_err(e1, qphi) > 1e-8
    g1[72 * 5 + 18 * 1 + 1] -= 1e-3
    e2 = np.fromfile(tmp_path / "qphi_edit2.bin").reshape(L.V, 24)
    assert rel_err(e2, L.Q_phi(g1.reshape(g.shape), phi)) < 1e-13


def test_shim_exports_its_header():
    """Every function include/cvc_dropin.h declares is defined by the shim object."""
    import re
    text = open(os.path.join(ROOT, "include", "cvc_dropin.h")).read()
    decl = set(re.findall(r"^\s*\w+\s*\*?\s*(cvchip_\w+)\s*\(", text, re.M))
    assert decl == {"cvchip_gauge_changed"}
    out = subprocess.run(["nm", OBJ], capture_output=True, text=True, check=True).stdout
    defined = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert decl <= defined

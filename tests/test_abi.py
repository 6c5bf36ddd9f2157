"""The C-ABI library loads and exports every entry point declared in
include/cvchip.h (CPU only; no compute calls)."""
import ctypes
import os
import re

from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "cvchip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(cvchip_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol(chip):
    lib = ctypes.CDLL(chip.LIB_PATH)
    decl = declared_symbols()
    assert len(decl) >= 20
    missing = [s for s in decl if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(chip.EXPORTS) == decl


def test_gauge_generator_is_su3_and_deterministic(chip):
    import numpy as np
    g1 = chip.gauge_random((4, 4, 4, 8), 2, 0.3)
    g2 = chip.gauge_random((4, 4, 4, 8), 2, 0.3)
    assert np.array_equal(g1, g2)
    U = (g1[..., 0::2] + 1j * g1[..., 1::2]).reshape(-1, 3, 3)
    assert np.abs(U @ U.conj().transpose(0, 2, 1) - np.eye(3)).max() < 1e-14
    assert np.abs(np.linalg.det(U) - 1).max() < 1e-14
    g3 = chip.gauge_random((4, 4, 4, 8), 3, 0.3)
    assert not np.array_equal(g1, g3)


def test_errors_without_gpu_work(chip):
    import pytest
    with pytest.raises(chip.CvchipError):
        chip.gauge_random((0, 4, 4, 4), 1, 0.3)


def test_gamma_tables_equal_the_oracle(orc):
    """The library builds its own gamma tables (from the chiral gamma_mu); they equal
    the oracle's, which are pinned to the reference's documented matrices."""
    import numpy as np
    import cvc_amd
    perm, sign = cvc_amd.gamma_tables()
    po, so = orc.gamma_tables()
    assert np.array_equal(perm, np.asarray(po)) and np.array_equal(sign, np.asarray(so))

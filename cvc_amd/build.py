"""Build recipe of ``libcvchip.so`` (HIP for gfx950 + host C++), in-tree.

``python -m cvc_amd.build`` or ``cvc_amd.build.build()``.  Every ``.hip`` file
under ``csrc/`` is compiled by hipcc for gfx950 only; host-only ``.cc`` files
by g++ with OpenMP and ``-ffp-contract=off`` (the gauge generator must give
the same bits on every host); the objects are linked into
``cvc_amd/libcvchip.so``.
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcvchip.so")
DROPIN_SRC = os.path.join(HERE, "dropin", "cvc_dropin.cc")
DROPIN_OBJ = os.path.join(HERE, "dropin", "cvc_dropin.o")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off: the compiler otherwise forms them from float2 code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    f"-I{INCLUDE}",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
             "-Wall", f"-I{INCLUDE}"]


def verbose_warnings(out: bytes) -> bool:
    # hipcc's host pass does not know the packed-fp32 feature flag; that note is harmless
    lines = [l for l in out.decode(errors="replace").splitlines() if "packed-fp32-ops" not in l]
    return bool("".join(lines).strip())


def _run_all(cmds):
    procs = [(c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)) for c in cmds]
    failed = []
    for c, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((c, out.decode(errors="replace")))
        elif out.strip() and verbose_warnings(out):
            # keep warnings visible but short
            text = "\n".join(l for l in out.decode(errors="replace").splitlines() if "packed-fp32-ops" not in l)
            sys.stderr.write(text[-4000:] + "\n")
    if failed:
        msg = "\n".join(" ".join(c) + "\n" + o for c, o in failed)
        raise RuntimeError("libcvchip build failed:\n" + msg)


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    hip_src = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    cc_src = sorted(glob.glob(os.path.join(CSRC, "*.cc")))
    cmds, objs = [], []
    for s in hip_src:
        o = os.path.join(OBJ, os.path.basename(s) + ".o")
        cmds.append([HIPCC, *HIP_FLAGS, "-c", s, "-o", o])
        objs.append(o)
    for s in cc_src:
        o = os.path.join(OBJ, os.path.basename(s) + ".o")
        cmds.append(["g++", *CXX_FLAGS, "-c", s, "-o", o])
        objs.append(o)
    if verbose:
        for c in cmds:
            print(" ".join(c))
    _run_all(cmds)
    tmp = LIB + ".tmp"
    link = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lgomp"]
    _run_all([link])
    shutil.move(tmp, LIB)
    # the drop-in shim is linked by a reference program, not into libcvchip.so
    _run_all([["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{INCLUDE}", "-c", DROPIN_SRC, "-o", DROPIN_OBJ]])
    # the drop-in contraction program (replaces the reference's cvc_exact2_xspace)
    prog = os.path.join(HERE, "dropin", "cvc_exact2_xspace")
    _run_all([["g++", "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}", prog + ".cc", f"-L{HERE}", "-lcvchip",
               "-Wl,-rpath,$ORIGIN/..", "-o", prog]])
    # a stand-in reference program that links the shim (used by the drop-in test)
    drv = os.path.join(os.path.dirname(HERE), "tests", "dropin")
    if os.path.exists(os.path.join(drv, "driver.cc")):
        _run_all([["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", os.path.join(drv, "driver.cc"), DROPIN_OBJ, f"-L{HERE}", "-lcvchip",
                   "-Wl,-rpath,$ORIGIN/../../cvc_amd", "-o", os.path.join(drv, "driver")]])
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

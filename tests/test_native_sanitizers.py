"""The library's host-only code (LIME readers and writers, the q-hat^2 orbits,
the gauge generator) built with AddressSanitizer and UBSan and fed every
truncation, single-byte corruptions and random files of each format
(tests/native/lime_fuzz.cc): every reader returns, with or without an error,
and no sanitizer reports anything.  CPU only."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT


def test_host_readers_under_sanitizers(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not available")
    exe = tmp_path / "lime_fuzz"
    srcs = [os.path.join(ROOT, "tests", "native", "lime_fuzz.cc")] + [
        os.path.join(ROOT, "cvc_amd", "csrc", f) for f in ("lime.cc", "analysis.cc", "gauge_gen.cc")]
    cmd = [gxx, "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", f"-I{ROOT}/include", f"-I{ROOT}/cvc_amd/csrc", *srcs, "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    files = tmp_path / "files"
    files.mkdir()
    r = subprocess.run([str(exe), str(files)], capture_output=True, text=True, timeout=300)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-3000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-3000:]
    assert "host readers under sanitizers: done" in out

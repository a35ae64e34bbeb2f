"""The LDS four-step FFTs of the shift search (matcha_amd/csrc/fft_lds.hpp) on
the CPU: a host build of the same step functions, run thread by thread,
against a direct DFT for every supported line length."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_lds_fft_lengths_match_direct_dft(tmp_path):
    if not Path(HIPCC).exists() and shutil.which("hipcc") is None:
        pytest.skip("hipcc not available")
    exe = tmp_path / "fft_check"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-I", str(ROOT / "matcha_amd" / "csrc"),
                    str(ROOT / "tests" / "fft_lds_check.cpp"), "-o", str(exe)], check=True,
                   stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    out = subprocess.run([str(exe)], check=True, stdout=subprocess.PIPE, text=True).stdout.split("\n")
    rows = [ln.split() for ln in out if ln.strip()]
    assert [int(r[0]) for r in rows] == [16, 32, 48, 64, 80, 96, 112, 128, 144, 160, 192]
    for n, ef, er in rows:
        assert float(ef) < 2e-6, (n, ef)
        assert float(er) < 2e-6, (n, er)

"""The FP64 special functions of the GPU basis build (matcha_amd/csrc/special.hpp)
on the CPU: spherical Bessel functions against SciPy, orthonormal Legendre
functions against the host recurrence of the projection tables."""
import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest
from scipy import special

from matcha_amd.core.FLEBasis3D import _legendre_orthonormal

ROOT = Path(__file__).resolve().parents[1]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def values(tmp_path_factory):
    if not Path(HIPCC).exists() and shutil.which("hipcc") is None:
        pytest.skip("hipcc not available")
    exe = tmp_path_factory.mktemp("special") / "special_check"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-I", str(ROOT / "matcha_amd" / "csrc"),
                    str(ROOT / "tests" / "special_check.cpp"), "-o", str(exe)], check=True,
                   stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    out = subprocess.run([str(exe)], check=True, stdout=subprocess.PIPE, text=True).stdout
    return [ln.split() for ln in out.splitlines() if ln.strip()]


def test_spherical_bessel_matches_scipy(values):
    rows = [r for r in values if r[0] == "j"]
    assert len(rows) > 400
    worst = 0.0
    for _, l, x, v in rows:
        l, x, v = int(l), float(x), float(v)
        ref = special.spherical_jn(l, x)
        # relative where the value is representable, else absolute against the order's scale
        err = abs(v - ref) / max(abs(ref), 1e-300)
        if abs(ref) > 1e-250:
            worst = max(worst, err)
            assert err < 1e-11, (l, x, v, ref)
        else:
            assert abs(v) < 1e-240, (l, x, v, ref)
    assert worst < 1e-11


def test_legendre_matches_the_host_recurrence(values):
    rows = [r for r in values if r[0] == "p"]
    assert len(rows) > 300
    for _, l, m, x, v in rows:
        l, m, x, v = int(l), int(m), float(x), float(v)
        ref = float(_legendre_orthonormal(l, np.array([x]))[l][m][0])
        assert abs(v - ref) <= 1e-13 * max(1.0, abs(ref)), (l, m, x, v, ref)

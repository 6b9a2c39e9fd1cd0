"""Host proof of the refinement steps of the short correctly rounded float32 division and
square root (geoswe_amd/csrc/gs_crmath.h): tests/native/cr_check.cpp feeds them every value
a 1-ulp hardware approximation can take, over every significand, and compares with the
host's IEEE results (a few seconds on 8 threads)."""
import shutil
import subprocess
from pathlib import Path

import pytest

NATIVE = Path(__file__).resolve().parent / "native"


def test_cr_refinements_exhaustive(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "cr_check"
    subprocess.run([gxx, "-O2", "-fopenmp", "-ffp-contract=off", "-std=c++17", "-o", str(exe),
                    str(NATIVE / "cr_check.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "failures 0" in r.stdout

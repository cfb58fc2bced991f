"""The device sine of RFCavityPass (at_amd/csrc/atx_libm.h) compiled for the
host and compared bit for bit with the host C library's sin, which the
reference calls (atintegrators/attrackfunc.c:24,36), on seeded random
arguments in every branch of the routine and on the branch edges. CPU only:
the same source runs on the device, where IEEE-754 add, multiply and fma
round identically."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "libm", "sin_check.hip")

# (lo, hi): the Taylor branch, the table branch, the cos(pi/2 - x) branch,
# the reduced branch at RF phases of a few turns of path length, and far out
RANGES = [(1e-9, 0.126), (0.126, 0.85546875), (0.85546875, 2.4262647628784180),
          (2.4262647628784180, 60.0), (60.0, 1.0e8)]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("libm") / "sin_check")
    subprocess.run([hipcc, "-O2", "--offload-arch=gfx950", SRC, "-o", exe], check=True,
                   capture_output=True)
    return exe


@pytest.mark.parametrize("lo,hi", RANGES)
def test_restated_sin_equals_host_libm(checker, lo, hi):
    out = subprocess.run([checker, repr(lo), repr(hi), "400000", "20251203"], check=True,
                         capture_output=True, text=True).stdout
    last = out.strip().splitlines()[-1].split()
    assert last[0] == "checked" and int(last[1]) > 400000
    assert int(last[3]) == 0, out

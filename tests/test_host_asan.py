"""CPU: the host side of the C ABI under AddressSanitizer and
LeakSanitizer. tests/asan/host_driver.cpp builds lattices with every pass
method and attribute kind, compiles them for all integrator types, changes
attributes between compilations and walks the error paths; the host runtime
(at_amd/csrc/atx_host.cpp) is rebuilt with the sanitizer for it, the
device code is the library's own build. No GPU is used."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "at_amd", "_build")
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC) or shutil.which("g++") is None,
                    reason="hipcc and g++ are needed")
def test_host_abi_under_address_sanitizer(tmp_path):
    from at_amd.build import SHAPES

    objs = [os.path.join(BUILD, f"atx_kernel_{b}_p{p}_u{u}.o")
            for b in ("strict", "fast") for p, u in SHAPES]
    objs.append(os.path.join(BUILD, "atx_adjoint.o"))
    if not all(os.path.exists(o) for o in objs):
        pytest.skip("device objects not built (python -m at_amd.build)")
    host = str(tmp_path / "host.o")
    drv = str(tmp_path / "driver.o")
    exe = str(tmp_path / "driver")
    san = ["-O1", "-g", "-fno-omit-frame-pointer"]
    subprocess.run([HIPCC, "-c", *san, "-std=c++17", "--offload-arch=gfx950", "-Xclang",
                    "-target-feature", "-Xclang", "-packed-fp32-ops", "-ffp-contract=off",
                    "-Xarch_host", "-fsanitize=address",
                    os.path.join(ROOT, "at_amd", "csrc", "atx_host.cpp"), "-o", host],
                   check=True, capture_output=True)
    subprocess.run(["g++", "-c", *san, "-std=c++17", "-fsanitize=address",
                    os.path.join(ROOT, "tests", "asan", "host_driver.cpp"), "-o", drv],
                   check=True, capture_output=True)
    subprocess.run([HIPCC, "-Xarch_host", "-fsanitize=address", drv, host, *objs, "-o", exe],
                   check=True, capture_output=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0")
    out = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert out.returncode == 0, out.stderr[-4000:]
    assert "host self-check passed" in out.stdout
    assert "AddressSanitizer" not in out.stderr and "LeakSanitizer" not in out.stderr

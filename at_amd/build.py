"""Build the engine's shared library in-tree with hipcc for gfx950.

    python -m at_amd.build          # builds at_amd/libatx.so

The kernel source is compiled twice (strict: reference operation order and
no FP contraction; fast: FMA contraction) and linked with the host runtime
into one library that exports the C ABI of include/atx.h.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libatx.so")
ARCH = os.environ.get("ATX_OFFLOAD_ARCH", "gfx950")

# No packed-FP32 instructions anywhere (they are not used by this FP64 code,
# and the flag keeps the compiler from forming them).
COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose_isa: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    objs = []
    for name, strict, contract in (("strict", 1, "off"), ("fast", 0, "fast")):
        obj = os.path.join(BUILD, f"atx_kernel_{name}.o")
        cmd = [cc, "-c", *COMMON, f"-DATX_STRICT={strict}", f"-ffp-contract={contract}",
               os.path.join(CSRC, "atx_kernel.hip"), "-o", obj]
        if verbose_isa:
            cmd += ["-Rpass-analysis=kernel-resource-usage"]
        _run(cmd)
        objs.append(obj)
    adj_obj = os.path.join(BUILD, "atx_adjoint.o")
    _run([cc, "-c", *COMMON, "-ffp-contract=fast", os.path.join(CSRC, "atx_adjoint.hip"), "-o",
          adj_obj] + (["-Rpass-analysis=kernel-resource-usage"] if verbose_isa else []))
    objs.append(adj_obj)
    host_obj = os.path.join(BUILD, "atx_host.o")
    _run([cc, "-c", *COMMON, "-ffp-contract=off", os.path.join(CSRC, "atx_host.cpp"), "-o", host_obj])
    objs.append(host_obj)
    tmp = LIB + ".tmp"
    _run([cc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose_isa="--isa" in sys.argv)

"""Build the engine's shared library in-tree with hipcc for gfx950.

    python -m at_amd.build          # builds at_amd/libatx.so

The kernel source is compiled once per build (strict: reference operation
order and no FP contraction; fast: FMA contraction) and launch shape, the
objects in parallel, and linked with the host runtime into one library that
exports the C ABI of include/atx.h.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

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


# Launch shapes with a kernel of their own: particles per lane x slices per
# loop iteration (include/atx.h atx_launch_shape). Two particles with four
# slices per iteration would need more than a lane's 256 registers.
SHAPES = [(1, 1), (1, 2), (1, 4), (2, 1), (2, 2)]


def _jobs() -> int:
    env = os.environ.get("MAX_JOBS", "")
    n = int(env) if env.isdigit() and int(env) > 0 else len(os.sched_getaffinity(0))
    return max(1, min(16, n))


def build(verbose_isa: bool = False) -> str:
    """Compile every object in parallel and link libatx.so. The strict build
    keeps the reference operation order (no contraction); the fast build
    contracts a*b+c within an expression only (-ffp-contract=on), so that the
    kernels of all launch shapes round identically."""
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    isa = ["-Rpass-analysis=kernel-resource-usage"] if verbose_isa else []
    cmds = []
    for name, strict, contract in (("strict", 1, "off"), ("fast", 0, "on")):
        for ppt, unroll in SHAPES:
            obj = os.path.join(BUILD, f"atx_kernel_{name}_p{ppt}_u{unroll}.o")
            cmds.append([cc, "-c", *COMMON, f"-DATX_STRICT={strict}", f"-DATX_PPT={ppt}",
                         f"-DATX_UNROLL={unroll}", f"-ffp-contract={contract}",
                         os.path.join(CSRC, "atx_kernel.hip"), "-o", obj] + isa)
    cmds.append([cc, "-c", *COMMON, "-ffp-contract=fast", os.path.join(CSRC, "atx_adjoint.hip"),
                 "-o", os.path.join(BUILD, "atx_adjoint.o")] + isa)
    cmds.append([cc, "-c", *COMMON, "-ffp-contract=off", os.path.join(CSRC, "atx_host.cpp"), "-o",
                 os.path.join(BUILD, "atx_host.o")])
    objs = [c[c.index("-o") + 1] for c in cmds]
    # the longest compilations first
    order = sorted(range(len(cmds)), key=lambda i: ("_p2_" not in objs[i], objs[i]))
    with ThreadPoolExecutor(max_workers=_jobs()) as pool:
        for f in [pool.submit(_run, cmds[i]) for i in order]:
            f.result()
    tmp = LIB + ".tmp"
    _run([cc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose_isa="--isa" in sys.argv)

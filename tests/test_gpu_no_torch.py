"""GPU: batched launches and the serving loop need no torch.

A child process in which ``import torch`` fails (and which loads the system
ROCm runtime, ``ATX_HIP_RUNTIME=system``) runs the batch and serving test
modules; it also checks that no torch module or library was loaded. Both
modules allocate, copy, stream and time through the engine's own C ABI
(include/atx.h: atx_device_alloc ... atx_event_elapsed)."""

import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

TESTS = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(TESTS)

CHILD = r"""
import sys
sys.modules["torch"] = None          # any import of torch now raises ImportError
import pytest
rc = pytest.main(["-x", "-q", "-p", "no:cacheprovider", "-m", "gpu",
                  "tests/test_gpu_batch.py", "tests/test_gpu_serving.py"])
maps = open("/proc/self/maps").read()
loaded = sorted({line.split()[-1] for line in maps.splitlines()
                 if "torch" in line.split()[-1]})
print("torch libraries loaded:", loaded, flush=True)
sys.exit(rc if rc else (3 if loaded else 0))
"""


def test_batch_and_serving_without_torch():
    env = dict(os.environ, ATX_HIP_RUNTIME="system")
    print("child: batch and serving tests with torch blocked", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD], cwd=ROOT, env=env, timeout=600,
                       capture_output=True, text=True)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0

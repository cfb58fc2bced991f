import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def _have_gpu() -> bool:
    # Ask the engine itself (HIP runtime), not torch: the GPU tests call
    # libatx.so and must fail loudly if it is missing on a GPU machine.
    try:
        from at_amd import engine

        return engine.device_count() > 0
    except ImportError:
        import glob

        return bool(glob.glob("/dev/dri/renderD*")) and os.path.exists("/dev/kfd")
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _have_gpu():
        return
    skip = pytest.mark.skip(reason="no HIP device")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)

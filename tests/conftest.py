import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch  # noqa: F401  (initialises the HIP runtime used for device arrays)
    import sparkle_amd as spk

    ctx = spk.Context(0)
    yield ctx
    ctx.close()


def print_case(*args):
    """Names the case before a GPU step, flushed, so a crash report identifies it."""
    print("CASE:", *args, flush=True)

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs on the MI355X box)")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as O

    O.build_lib()
    O.lib()
    return O


@pytest.fixture(scope="session")
def cuda():
    import torch

    assert torch.cuda.is_available(), "GPU test needs a device"
    from mnrt_amd import native

    native.load()
    native.check(native.load().mnrt_device_init(0), "mnrt_device_init")
    return torch

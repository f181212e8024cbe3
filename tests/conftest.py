import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    import oracle as orc

    orc._lib()
    return orc


@pytest.fixture(scope="session")
def gpu():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a GPU (run CPU tests with -m 'not gpu')")
    import feelmri_amd

    feelmri_amd._lib.load()
    return torch.device("cuda", 0)

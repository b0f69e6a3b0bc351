"""Shared pytest configuration: the ``gpu`` marker and repository imports."""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
if str(ROOT / 'tests') not in sys.path:
    sys.path.insert(0, str(ROOT / 'tests'))


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (ROCm device) and the built HIP library')


@pytest.fixture(scope='session')
def gpu_device():
    """One torch stream + the HIP library for every GPU test in the process."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no ROCm device visible')
    from nh_flood_2d_amd.engine import TorchDevice
    return TorchDevice()

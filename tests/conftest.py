"""pytest configuration: the gpu marker, the repository on sys.path, the GPU probe"""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (HIP device); the parity tests of the HIP path')


def gpu_available():
    from slipcam_amd import framecor
    return framecor.has_gpu()


@pytest.fixture(scope='session')
def gpu():
    if not gpu_available():
        pytest.skip('no GPU visible to HIP')
    return True

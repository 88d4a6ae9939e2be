import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (run with -m gpu)')
    config.addinivalue_line('markers', 'slow: CPU tests taking more than a few seconds')


@pytest.fixture(scope='session')
def gpu_available():
    import torch
    return torch.cuda.is_available()

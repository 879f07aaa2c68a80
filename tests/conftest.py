import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X GPU and the built HIP library')
    config.addinivalue_line('markers', 'slow: longer CPU test')


@pytest.fixture(scope='session')
def hip():
    """The HIP backend module; fails (not skips) when the GPU or library is missing."""
    import transiflow_amd
    from transiflow_amd import _lib
    _lib.load()
    from transiflow_amd.vector import torch
    torch()
    return transiflow_amd

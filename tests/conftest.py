import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, 'oracle')):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line(
        'markers', 'gpu: needs an MI355X (HIP device) and the built library')


def pytest_runtest_setup(item):
    # name the case before any GPU work, so a fault can be attributed
    if item.get_closest_marker('gpu') is not None:
        print('\n[gpu-case] %s' % item.nodeid, flush=True)


@pytest.fixture(scope='session')
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    import gpawdft_amd
    gpawdft_amd.library_loaded()
    return torch.device('cuda')

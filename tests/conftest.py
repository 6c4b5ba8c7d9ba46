import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: test needs an MI355X (runs the HIP kernels)')


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope='session')
def mech_c1_data():
    from spitfire_amd import synthetic
    return synthetic.mechanism_c1()


@pytest.fixture(scope='session')
def mech_53_data():
    from spitfire_amd import synthetic
    return synthetic.mechanism_53()


@pytest.fixture(scope='session')
def oracle_lib():
    from oracle import oracle
    if not os.path.exists(oracle.ORACLE_SO):
        oracle.build(ref=False)
    return oracle


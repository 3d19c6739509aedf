import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, 'tests', 'golden')


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a HIP device (MI355X)')


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope='session')
def golden():
    """The reference's golden regression vectors (tests/simulation/simpli_rep.h5)."""
    d = os.path.join(GOLDEN, 'simpli_rep')
    out = {
        'tissue': np.load(os.path.join(d, 'tissue.npy')).astype(np.int32),
        'optical_axis': np.load(os.path.join(d, 'optical_axis.npy')),
        'tissue_properties': np.load(os.path.join(d, 'tissue_properties.npy')),
        'data': [np.load(os.path.join(d, f'data_{t}.npy')) for t in range(5)],
    }
    return out


GOLDEN_LAYERS = [[(0.333, -0.004, 10, 'p'), (0.666, 0, 5, 'b'), (1.0, 0.004, 1, 'r')]]
GOLDEN_ROT = np.deg2rad([0, 30, 60, 90, 120, 150])
GOLDEN_TILTS = np.deg2rad([(0, 0), (5.5, 0), (5.5, 90), (5.5, 180), (5.5, 270)])


@pytest.fixture(scope='session')
def cube_bundles():
    """tests/cube.dat of the reference with the golden run's layers."""
    from fastpli_amd.io import fiber_bundles
    fbs = fiber_bundles.load(os.path.join(GOLDEN, 'cube.dat'))
    fbs.layers = GOLDEN_LAYERS
    return fbs


def rel_err(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)

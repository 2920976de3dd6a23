import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, 'tests', 'golden', 'cuau')


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs through the HIP library)')
    config.addinivalue_line('markers', 'slow: longer CPU test')


@pytest.fixture(scope='session')
def cuau():
    """The reference's CuAu example: snapshot state, reconstructions and measured data."""
    snap = dict(np.load(os.path.join(GOLDEN, 'snapshot.npz')))
    recon = dict(np.load(os.path.join(GOLDEN, 'recon.npz')))
    meas = np.load(os.path.join(GOLDEN, 'measured.npz'))['measured']
    with open(os.path.join(GOLDEN, 'manifest.json')) as fh:
        man = json.load(fh)
    return dict(snap=snap, recon=recon, meas=meas, man=man)


def rel(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP library)')


def load_golden(name):
    with np.load(os.path.join(TESTS, 'golden', name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope='session')
def wilson_golden():
    return load_golden('wilson_4x4x4x4.npz')


@pytest.fixture(scope='session')
def mobius_golden():
    return load_golden('mobius_4x2x2x4_ls6.npz')


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def max_site_rel_err(a, b, site_axes=None):
    """max over sites of |a-b|/|b| with a site = all axes except the trailing (spin, colour)."""
    a = np.asarray(a).reshape(-1, 12)
    b = np.asarray(b).reshape(-1, 12)
    num = np.linalg.norm(a - b, axis=1)
    den = np.linalg.norm(b, axis=1)
    return float(np.max(num / np.maximum(den, 1e-300)))

"""CPU checks of the C ABI: libgx.so loads, exports every symbol include/gx.h declares, and the
host-side synthetic-input utilities behave (no device work is done here)."""
import ctypes
import os
import re

import numpy as np
import pytest

import grid_amd as gx
from conftest import ROOT

HEADER = os.path.join(ROOT, 'include', 'gx.h')


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(gx_[a-z0-9_A-Z]+)\s*\(', text)))


def test_header_symbols_exported():
    L = ctypes.CDLL(gx.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) > 40
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_symbols()) == set(gx.exported_symbols())


def test_random_gauge_is_su3():
    dims = (4, 4, 2, 2)
    U = gx.random_gauge(dims, 1234, 0.5).reshape(-1, 3, 3)
    eye = np.eye(3)
    assert np.max(np.abs(U @ np.conj(np.swapaxes(U, -1, -2)) - eye)) < 1e-14
    assert np.max(np.abs(np.linalg.det(U) - 1)) < 1e-14
    # deterministic and seed dependent
    assert np.array_equal(U, gx.random_gauge(dims, 1234, 0.5).reshape(-1, 3, 3))
    assert not np.array_equal(U, gx.random_gauge(dims, 1235, 0.5).reshape(-1, 3, 3))


def test_random_gauge_generator_is_exp_i_eps_h():
    # U = exp(i eps H) with H Hermitian traceless: log(U)/(i eps) is Hermitian and traceless, and
    # for eps -> small the entries of H are O(1) Gaussians (variance 1/2 off-diagonal parts).
    dims = (8, 4, 4, 4)
    eps = 0.05
    U = gx.random_gauge(dims, 99, eps).reshape(-1, 3, 3)
    w, v = np.linalg.eig(U)
    theta = np.angle(w)
    H = np.einsum('nab,nb,nbc->nac', v, theta / eps, np.linalg.inv(v))
    assert np.max(np.abs(H - np.conj(np.swapaxes(H, -1, -2)))) < 1e-8
    assert np.max(np.abs(np.trace(H, axis1=1, axis2=2))) < 1e-8
    off = H[:, 0, 1]
    assert abs(np.var(off.real) - 0.5) < 0.05 and abs(np.var(off.imag) - 0.5) < 0.05


def test_gaussian_fermion_statistics():
    psi = gx.gaussian_fermion((8, 8, 4, 4), 2, 7)
    assert abs(np.mean(psi.real)) < 0.02 and abs(np.var(psi.real) - 1) < 0.03
    assert abs(np.var(psi.imag) - 1) < 0.03


def test_point_source():
    dims = (4, 4, 4, 4)
    psi = gx.point_source(dims, 1, (1, 2, 3, 0), 2, 1)
    assert psi.sum() == 1.0 and psi[0, 3, 2, 1, 0, 2, 1] == 1.0


def test_context_without_gpu_raises():
    import torch
    if torch.cuda.is_available():
        pytest.skip('GPU present')
    with pytest.raises(gx.GxError):
        gx.Context(0)

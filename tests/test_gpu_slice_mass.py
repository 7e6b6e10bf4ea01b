"""Multi-RHS Wilson operators with a mass per right-hand side (gx_op_set_slice_masses), through the
C ABI: slice s of every operator and of the red-black solve equals the oracle's single-RHS Wilson
operator of mass m_s on slice s of the input (WilsonFermionImplementation.h:71,113-176), so
requests of different masses can share one multi-RHS launch.

Tolerances: per-site relative 1e-13 (FP64 fields) and 1e-6 (FP32) against the FP64 oracle;
solves: |b - M x| / |b| <= 1e-10 per slice with the oracle's M of that slice's mass, and the slice
within 1e-8 of the oracle's red-black solve.
"""
import numpy as np
import pytest

import grid_amd as gx
import npref as R
from conftest import max_site_rel_err
from oracle.pyoracle import Oracle, EVEN, ODD

pytestmark = pytest.mark.gpu

TOL = {gx.F64: 1e-13, gx.F32: 1e-6}
DIMS = (8, 4, 4, 8)
MASSES = np.array([0.05, 0.12, -0.2, 0.3, 0.07])


@pytest.fixture(scope='module')
def ctx():
    c = gx.Context(0)
    yield c
    c.close()


@pytest.fixture(scope='module')
def setup():
    U = gx.random_gauge(DIMS, 61, 0.4)
    psi = gx.gaussian_fermion(DIMS, len(MASSES), 62)
    oracles = [Oracle(U, DIMS, 1, 'wilson', mass=m) for m in MASSES]
    return U, psi, oracles


def sl(a, k):
    return np.ascontiguousarray(a[..., k:k + 1, :, :])


@pytest.mark.parametrize('prec', [gx.F64, gx.F32])
@pytest.mark.parametrize('fused', [1, 0])
def test_slice_mass_operators(ctx, setup, prec, fused):
    U, psi, oracles = setup
    D = gx.FermionOp(ctx, U, DIMS, len(MASSES), 'wilson', mass=0.5)
    D.set_slice_masses(MASSES)
    assert np.array_equal(D.slice_masses(), MASSES)
    old = gx.tune_get(gx.TUNE_FUSED)
    gx.tune_set(gx.TUNE_FUSED, fused)
    try:
        fi, fo = D.field(prec).upload(psi), D.field(prec)
        for name in ('M', 'Mdag'):
            getattr(D, name)(fi, fo)
            out = fo.download()
            for k, O in enumerate(oracles):
                assert max_site_rel_err(sl(out, k), O.apply(name, sl(psi, k))) < TOL[prec], (name, k)
        po = R.pick(psi, DIMS, ODD)
        hi, ho = D.field(prec, ODD).upload(po), D.field(prec, ODD)
        for name in ('Mooee', 'MooeeInv', 'MooeeDag', 'MooeeInvDag'):
            getattr(D, name)(hi, ho)
            out = ho.download()
            for k, O in enumerate(oracles):
                assert max_site_rel_err(sl(out, k), O.apply(name, sl(po, k), ODD)) < TOL[prec], (name, k)
        for name in ('Mpc', 'MpcDag', 'MpcDagMpc'):
            getattr(D, name)(hi, ho)
            out = ho.download()
            for k, O in enumerate(oracles):
                assert max_site_rel_err(sl(out, k), O.schur(name, sl(po, k))) < TOL[prec] * 10, (name, k)
    finally:
        gx.tune_set(gx.TUNE_FUSED, old)
    D.close()


@pytest.mark.parametrize('mixed', [0, 1, 2])
def test_slice_mass_schur_solve(ctx, setup, mixed):
    U, psi, oracles = setup
    D = gx.FermionOp(ctx, U, DIMS, len(MASSES), 'wilson', mass=0.5)
    D.set_slice_masses(MASSES)
    fs, fx = D.field().upload(psi), D.field()
    st = D.schur_solve(fs, fx, tol=1e-11, maxit=5000, mixed=mixed)
    assert st['converged'], st
    x = fx.download()
    for k, O in enumerate(oracles):
        b = sl(psi, k)
        r = O.apply('M', sl(x, k)) - b
        assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-10, (k, mixed)
        ref, _ = O.schur_solve(b, tol=1e-11)
        assert np.linalg.norm(sl(x, k) - ref) / np.linalg.norm(ref) < 1e-8, (k, mixed)
    D.close()


def test_slice_mass_clone_and_reset(ctx, setup):
    """A clone with per-slice masses matches a fresh operator; set_mass and None return to one mass."""
    U, psi, oracles = setup
    base = gx.FermionOp(ctx, U, DIMS, 1, 'wilson', mass=0.1)
    cl = base.clone(ctx, len(MASSES), 0.1)
    cl.set_slice_masses(MASSES)
    po = R.pick(psi, DIMS, ODD)
    hi, ho = cl.field(gx.F64, ODD).upload(po), cl.field(gx.F64, ODD)
    cl.Mpc(hi, ho)
    out = ho.download()
    for k, O in enumerate(oracles):
        assert max_site_rel_err(sl(out, k), O.schur('Mpc', sl(po, k))) < 1e-12
    cl.set_slice_masses(None)
    assert np.all(cl.slice_masses() == 0.1)
    cl.set_slice_masses(MASSES)
    cl.set_mass(0.3)
    assert np.all(cl.slice_masses() == 0.3)
    cl.Mpc(hi, ho)
    O3 = Oracle(U, DIMS, 1, 'wilson', mass=0.3)
    out = ho.download()
    for k in range(len(MASSES)):
        assert max_site_rel_err(sl(out, k), O3.schur('Mpc', sl(po, k))) < 1e-12
    with pytest.raises(gx.GxError) as e:
        cl.set_slice_masses(MASSES[:3])
    assert e.value.code == 1
    cl.close()
    base.close()
    M = gx.FermionOp(ctx, U, DIMS, 4, 'mobius', mass=0.1, M5=1.8, b=1.5, c=0.5)
    with pytest.raises(gx.GxError) as e:
        M.set_slice_masses([0.1] * 4)
    assert e.value.code == 1
    M.close()

"""ConjugateGradientMultiShift: every shifted Schur system (MpcDagMpc + shift) solved from one Krylov
space, checked against the oracle's MpcDagMpc (residual of each shift) and the oracle's CG (the
unshifted system)."""
import numpy as np
import pytest

import grid_amd as gx
import npref as R
from conftest import rel_err
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu
ODD = gx.ODD


@pytest.fixture(scope='module')
def ctx():
    c = gx.Context(0)
    yield c
    c.close()


def norm(a):
    return np.sqrt(np.sum(np.abs(a) ** 2))


@pytest.mark.parametrize('kind,Ls', [('wilson', 1), ('mobius', 8)])
def test_multishift_every_shift(ctx, kind, Ls):
    """Each psi_s solves (MpcDagMpc + shift_s) psi_s = src to its tolerance (residual through the
    oracle's MpcDagMpc); the reported true residual is that residual; the unshifted solution is the
    oracle CG's; with equal tolerances heavier shifts stop no later than lighter ones."""
    dims, mass = (8, 8, 8, 4), 0.05
    U = gx.random_gauge(dims, 401, 0.3)
    D = gx.FermionOp(ctx, U, dims, Ls, kind, mass=mass)
    src = R.pick(gx.gaussian_fermion(dims, Ls, 402), dims, ODD)
    fs = D.field(gx.F64, ODD).upload(src)
    shifts = [0.0, 0.003, 0.05, 0.4, 2.0]
    tols = [1e-10] * len(shifts)
    psis = [D.field(gx.F64, ODD) for _ in shifts]
    sts = D.cg_multishift(fs, psis, shifts, tols)
    O = Oracle(U, dims, Ls, kind, mass=mass)
    for s, (sh, f) in enumerate(zip(shifts, psis)):
        x = f.download()
        res = norm(O.schur('MpcDagMpc', x) + sh * x - src) / norm(src)
        assert sts[s]['converged'] and res < 10 * tols[s], (s, res, sts[s])
        assert abs(sts[s]['true_residual'] - res) < 1e-3 * tols[s] + 1e-6 * res, (s, res, sts[s])
    ref, _ = O.cg(src, tol=1e-10)
    assert rel_err(psis[0].download(), ref) < 1e-8
    its = [st['iterations'] for st in sts]
    assert all(its[i + 1] <= its[i] for i in range(len(its) - 1)), its
    assert its[0] > its[-1], its


def test_multishift_tolerances_and_errors(ctx):
    """Per-shift tolerances stop each shift at its own residual; a zero source gives zero solutions;
    a lighter shift after the first, or a repeated solution field, is rejected."""
    dims, mass = (8, 4, 4, 8), 0.1
    U = gx.random_gauge(dims, 411, 0.3)
    D = gx.FermionOp(ctx, U, dims, 1, 'wilson', mass=mass)
    src = R.pick(gx.gaussian_fermion(dims, 1, 412), dims, ODD)
    fs = D.field(gx.F64, ODD).upload(src)
    shifts, tols = [0.01, 0.1, 1.0], [1e-10, 1e-6, 1e-8]
    psis = [D.field(gx.F64, ODD) for _ in shifts]
    sts = D.cg_multishift(fs, psis, shifts, tols)
    O = Oracle(U, dims, 1, 'wilson', mass=mass)
    for s, sh in enumerate(shifts):
        x = psis[s].download()
        res = norm(O.schur('MpcDagMpc', x) + sh * x - src) / norm(src)
        assert sts[s]['converged'] and res < 10 * tols[s], (s, res)
    assert max(sts[1]['iterations'], sts[2]['iterations']) < sts[0]['iterations'], sts
    zero = D.field(gx.F64, ODD)
    for f in psis:
        f.upload(src)
    zs = D.cg_multishift(zero, psis, shifts, tols)
    assert all(st['converged'] for st in zs)
    assert all(np.count_nonzero(f.download()) == 0 for f in psis)
    with pytest.raises(gx.GxError):
        D.cg_multishift(fs, psis, [0.1, 0.01, 1.0], tols)
    with pytest.raises(gx.GxError):
        D.cg_multishift(fs, [psis[0], psis[0], psis[2]], shifts, tols)


def test_multishift_maximum_shifts(ctx):
    """32 shifts (the kernel's maximum) all converge; 33 are refused."""
    dims, mass = (4, 4, 4, 8), 0.1
    U = gx.random_gauge(dims, 421, 0.3)
    D = gx.FermionOp(ctx, U, dims, 1, 'wilson', mass=mass)
    src = R.pick(gx.gaussian_fermion(dims, 1, 422), dims, ODD)
    fs = D.field(gx.F64, ODD).upload(src)
    shifts = list(np.geomspace(1e-3, 10.0, 32))
    psis = [D.field(gx.F64, ODD) for _ in range(33)]
    sts = D.cg_multishift(fs, psis[:32], shifts, [1e-10] * 32)
    O = Oracle(U, dims, 1, 'wilson', mass=mass)
    for s in (0, 15, 31):
        x = psis[s].download()
        res = norm(O.schur('MpcDagMpc', x) + shifts[s] * x - src) / norm(src)
        assert sts[s]['converged'] and res < 1e-9, (s, res)
    with pytest.raises(gx.GxError):
        D.cg_multishift(fs, psis, shifts + [20.0], [1e-10] * 33)

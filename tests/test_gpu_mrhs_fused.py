"""CG with its reductions in the hop epilogues, through the C ABI.

With FP32 / FP16-storage inner solves, CGmultiRHS takes d_s = |Mpc p_s|^2 from per-tile sums the
Mpc hop leaves, and the last MpcDag hop updates r_s -= a_s (MpcDag Mpc p)_s and sums |r_s|^2 without
storing MpcDag Mpc p (BlockConjugateGradient.h:412-545 computes the same scalars from separate
passes).  The default 4-wave hop kernel and the general tile shape have these epilogues; under the
persistent general grid, the row-tile kernel or (FP32) the matrix-core kernel the iteration runs
the separate passes.  This test
solves the same 12-RHS Wilson system (one mass per right-hand side, 8x4x4x8) both ways:

  - every slice of every solve reaches |b - M x| / |b| <= 1e-10 with the oracle's Wilson M of that
    slice's mass (WilsonFermionImplementation.h:71,113-176);
  - the fused and the separate-pass solutions agree to 1e-9 relative per slice (both are FP64
    defect-correction results at the 1e-11 red-black tolerance).

The single-RHS CG of FP32 / FP16-storage inner solves does the same on Mobius operators (the Mpc
hop's |t|^2 over all slices; r -= a MpcDag t in the matrix epilogue of the last MpcDag hop):
Mobius 8^4 x Ls8 defect correction, fused and separate-pass, each against the oracle's red-black CG
solution (ConjugateGradient.h:68-257, CayleyFermion5DImplementation.h:155-329) to 1e-7.
"""
import numpy as np
import pytest

import grid_amd as gx
from oracle.pyoracle import Oracle, ODD
import npref as R

pytestmark = pytest.mark.gpu

DIMS = (8, 4, 4, 8)
NRHS = 12
MASSES = np.linspace(0.04, 0.4, NRHS)
KNOBS = {
    'fused': {},
    'general': {gx.TUNE_DHOP_SHAPE: 6},
    'rows': {gx.TUNE_DHOP_TILE: 1},
    'mfma': {gx.TUNE_DHOP_MFMA: 1},
}


@pytest.fixture(scope='module')
def ctx():
    c = gx.Context(0)
    yield c
    c.close()


@pytest.fixture(scope='module')
def system():
    U = gx.random_gauge(DIMS, 71, 0.4)
    src = gx.gaussian_fermion(DIMS, NRHS, 72)
    return U, src, [Oracle(U, DIMS, 1, 'wilson', mass=m) for m in MASSES]


def sl(a, k):
    return np.ascontiguousarray(a[..., k:k + 1, :, :])


def solve(ctx, U, src, mixed, knobs):
    D = gx.FermionOp(ctx, U, DIMS, NRHS, 'wilson', mass=0.1)
    D.set_slice_masses(MASSES)
    prec = gx.F32 if mixed == 1 else gx.F16
    saved = {k: gx.tune_get(k, prec) for k in knobs}
    for k, v in knobs.items():
        gx.tune_set(k, v, prec)
    try:
        fs, fx = D.field().upload(src), D.field()
        st = D.schur_solve(fs, fx, tol=1e-11, maxit=5000, mixed=mixed)
        x = fx.download()
    finally:
        for k, v in saved.items():
            gx.tune_set(k, v, prec)
        D.close()
    return st, x


@pytest.mark.parametrize('mixed', [1, 2])
def test_fused_reductions_match_separate_passes(ctx, system, mixed):
    U, src, oracles = system
    names = [n for n in KNOBS if not (n == 'mfma' and mixed == 2)]
    sols = {}
    for name in names:
        st, x = solve(ctx, U, src, mixed, KNOBS[name])
        assert st['converged'], (name, st)
        for k, O in enumerate(oracles):
            b = sl(src, k)
            r = O.apply('M', sl(x, k)) - b
            assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-10, (name, k, mixed)
        sols[name] = x
    ref = sols['fused']
    for name in names[1:]:
        for k in range(NRHS):
            d = np.linalg.norm(sl(sols[name], k) - sl(ref, k)) / np.linalg.norm(sl(ref, k))
            assert d <= 1e-9, (name, k, mixed, d)


@pytest.mark.parametrize('inner', [gx.F32, gx.F16])
def test_single_rhs_fused_reductions(ctx, inner):
    dims, Ls, mass = (8, 8, 8, 8), 8, 0.1
    U = gx.random_gauge(dims, 73, 0.4)
    src = R.pick(gx.gaussian_fermion(dims, Ls, 74), dims, ODD)
    kw = dict(mass=mass, M5=1.8, b=1.5, c=0.5)
    ref, _ = Oracle(U, dims, Ls, 'mobius', **kw).cg(src, tol=1e-10)
    D = gx.FermionOp(ctx, U, dims, Ls, 'mobius', **kw)
    fs = D.field(gx.F64, ODD).upload(src)
    for name in ('fused', 'general', 'rows'):
        knobs = KNOBS[name]
        saved = {k: gx.tune_get(k, inner) for k in knobs}
        for k, v in knobs.items():
            gx.tune_set(k, v, inner)
        try:
            fx = D.field(gx.F64, ODD)
            st = D.mixed_cg(fs, fx, tol=1e-10, inner_tol=1e-5, inner_prec=inner)
            x = fx.download()
        finally:
            for k, v in saved.items():
                gx.tune_set(k, v, inner)
        assert st['converged'] and st['true_residual'] < 1e-9, (name, st)
        assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-7, name
    D.close()

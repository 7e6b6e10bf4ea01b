"""Block solvers over the fifth-dimension slices: sliceInnerProductMatrix / sliceMaddMatrix
against numpy, and BlockCGrQ (one block Krylov space for all right-hand sides) against the
oracle's single-RHS FP64 CG."""
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


def fields(ctx, dims, Ls, subset, seeds):
    out = []
    for sd in seeds:
        a = gx.gaussian_fermion(dims, Ls, sd)
        if subset != gx.FULL:
            a = R.pick(a, dims, subset)
        out.append((gx.Field(ctx, dims, Ls, gx.F64, subset).upload(a), a))
    return out


@pytest.mark.parametrize('subset', [ODD, gx.FULL])
@pytest.mark.parametrize('dims,Ls', [((4, 4, 4, 4), 5), ((8, 4, 6, 4), 12), ((4, 2, 2, 6), 16)])
def test_slice_inner_product_matrix(ctx, dims, Ls, subset):
    """m[s, t] = <x_s, y_t> (conjugate-linear in the left slice) equals numpy's contraction; m(x, x)
    is Hermitian with the slice norms on its diagonal."""
    (fx, x), (fy, y) = fields(ctx, dims, Ls, subset, (301, 302))
    m = gx.slice_inner_product_matrix(fx, fy)
    ref = np.einsum('nsk,ntk->st', x.reshape(-1, Ls, 12).conj(), y.reshape(-1, Ls, 12))
    assert np.max(np.abs(m - ref)) < 1e-12 * np.sqrt(np.sum(np.abs(x) ** 2) * np.sum(np.abs(y) ** 2))
    mxx = gx.slice_inner_product_matrix(fx, fx)
    assert np.max(np.abs(mxx - mxx.conj().T)) < 1e-12 * np.max(np.abs(mxx))
    norms = np.sum(np.abs(x) ** 2, axis=tuple(i for i in range(x.ndim) if i != x.ndim - 3))
    assert np.allclose(mxx.diagonal().real, norms, rtol=1e-12)


def mul(x, a):
    """sum_s x_s a(s, t) over the slice axis of a (..., Ls, 4, 3) array"""
    Ls = a.shape[0]
    return np.einsum('nsk,st->ntk', x.reshape(-1, Ls, 12), a).reshape(x.shape)


@pytest.mark.parametrize('subset', [ODD, gx.FULL])
@pytest.mark.parametrize('Ls', [3, 12])
def test_slice_madd_matrix(ctx, Ls, subset):
    """r_t = y_t + scale sum_s x_s a(s, t), also with r = x, r = y and without y (sliceMulMatrix)."""
    dims = (4, 4, 2, 6)
    rng = np.random.default_rng(7)
    a = rng.normal(size=(Ls, Ls)) + 1j * rng.normal(size=(Ls, Ls))
    scale = -0.75
    (fx, x), (fy, y), (fr, _) = fields(ctx, dims, Ls, subset, (311, 312, 313))
    ref = y + scale * mul(x, a)
    gx.slice_madd_matrix(fr, a, fx, fy, scale)
    assert rel_err(fr.download(), ref) < 1e-14
    gx.slice_madd_matrix(fr, a, fx, None, scale)
    assert rel_err(fr.download(), scale * mul(x, a)) < 1e-14
    gx.slice_madd_matrix(fx, a, fx, fy, scale)          # r aliases x
    assert rel_err(fx.download(), ref) < 1e-14
    fx.upload(x)
    gx.slice_madd_matrix(fy, a, fx, fy, scale)          # r aliases y
    assert rel_err(fy.download(), ref) < 1e-14
    # identity coefficients: r = y + x exactly
    fy.upload(y)
    gx.slice_madd_matrix(fr, np.eye(Ls), fx, fy, 1.0)
    assert np.array_equal(fr.download(), x + y)


@pytest.mark.parametrize('nrhs', [1, 4, 8, 16])
def test_block_cgrq_matches_single_rhs(ctx, nrhs):
    """BlockCGrQ reaches the oracle's single-RHS FP64 CG solution for every right-hand side, with
    true residuals at the tolerance, in no more iterations than CGmultiRHS needs for its slowest
    right-hand side (the block Krylov space contains every single-RHS one)."""
    dims, mass = (8, 8, 4, 8), 0.1
    U = gx.random_gauge(dims, 321, 0.5)
    src = R.pick(gx.gaussian_fermion(dims, nrhs, 322), dims, ODD)
    D = gx.FermionOp(ctx, U, dims, nrhs, 'wilson', mass=mass)
    fs, fp, fm = D.field(gx.F64, ODD).upload(src), D.field(gx.F64, ODD), D.field(gx.F64, ODD)
    sts = D.block_cgrq(fs, fp, tol=1e-10)
    smr = D.block_cg(fs, fm, tol=1e-10)
    x = fp.download()
    O1 = Oracle(U, dims, 1, 'wilson', mass=mass)
    for k in range(nrhs):
        ref, _ = O1.cg(np.ascontiguousarray(src[..., k:k + 1, :, :]), tol=1e-10)
        assert sts[k]['converged'] and sts[k]['true_residual'] < 1e-9, sts[k]
        assert sts[k]['computed_residual'] < 1e-10, sts[k]
        assert rel_err(x[..., k:k + 1, :, :], ref) < 1e-8
    assert sts[0]['iterations'] <= max(s['iterations'] for s in smr) + 1, (sts[0], smr)  # +1: rounding


def test_block_cgrq_guess_and_errors(ctx):
    """A nonzero guess enters through its residual; a zero right-hand side (rank-deficient block)
    and a Mobius operator are rejected."""
    dims, nrhs, mass = (8, 8, 4, 8), 3, 0.1
    U = gx.random_gauge(dims, 331, 0.5)
    src = R.pick(gx.gaussian_fermion(dims, nrhs, 332), dims, ODD)
    guess = 0.3 * R.pick(gx.gaussian_fermion(dims, nrhs, 333), dims, ODD)
    D = gx.FermionOp(ctx, U, dims, nrhs, 'wilson', mass=mass)
    fs, fp = D.field(gx.F64, ODD).upload(src), D.field(gx.F64, ODD).upload(guess)
    sts = D.block_cgrq(fs, fp, tol=1e-10)
    x = fp.download()
    O1 = Oracle(U, dims, 1, 'wilson', mass=mass)
    for k in range(nrhs):
        ref, _ = O1.cg(np.ascontiguousarray(src[..., k:k + 1, :, :]), tol=1e-10)
        assert sts[k]['converged'] and sts[k]['true_residual'] < 1e-9, sts[k]
        assert rel_err(x[..., k:k + 1, :, :], ref) < 1e-8
    z = src.copy()
    z[..., 1, :, :] = 0
    with pytest.raises(gx.GxError):
        D.block_cgrq(D.field(gx.F64, ODD).upload(z), D.field(gx.F64, ODD), tol=1e-10)
    M = gx.FermionOp(ctx, U, dims, 4, 'mobius', mass=0.1)
    with pytest.raises(gx.GxError):
        M.block_cgrq(M.field(gx.F64, ODD), M.field(gx.F64, ODD))

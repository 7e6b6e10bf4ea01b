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


@pytest.mark.parametrize('nrhs', [1, 3, 6])
def test_block_cgrq_vec_mobius_schur(ctx, nrhs):
    """BlockCGrQVec over a list of Mobius half fields (A = MpcDagMpc, the operator applied field by
    field): every solution equals the oracle's single-RHS FP64 CG solution, and the block's
    iteration count is at most the slowest single solve's."""
    dims, Ls = (8, 4, 4, 8), 8
    kw = dict(mass=0.05, M5=1.8, b=1.5, c=0.5)
    U = gx.random_gauge(dims, 341, 0.3)
    D = gx.FermionOp(ctx, U, dims, Ls, 'mobius', **kw)
    O = Oracle(U, dims, Ls, 'mobius', **kw)
    srcs = [R.pick(gx.gaussian_fermion(dims, Ls, 342 + k), dims, ODD) for k in range(nrhs)]
    fs = [D.field(gx.F64, ODD).upload(s) for s in srcs]
    fx = [D.field(gx.F64, ODD) for _ in range(nrhs)]
    sts = D.block_cgrq_vec(fs, fx, tol=1e-10, maxit=5000)
    worst = 0
    for k in range(nrhs):
        ref, rst = O.cg(srcs[k], tol=1e-10, maxit=5000)
        worst = max(worst, rst['iterations'])
        assert sts[k]['converged'] and sts[k]['true_residual'] < 1e-9, sts[k]
        assert rel_err(fx[k].download(), ref) < 1e-8, k
    assert sts[0]['iterations'] <= worst + 1, (sts[0], worst)


def test_block_cgrq_vec_mobius_mdagm_full(ctx):
    """BlockCGrQVec with Mdag M on full Mobius fields (MdagMLinearOperator, as in Grid's
    tests/solver/Test_mobius_bcg_nosplit.cc): each solution's residual |Mdag M x - b| / |b|,
    recomputed with the oracle's M and Mdag, is at the tolerance; guesses are honoured."""
    dims, Ls, nrhs = (4, 4, 4, 4), 6, 2
    kw = dict(mass=0.1, M5=1.8, b=1.5, c=0.5)
    U = gx.random_gauge(dims, 351, 0.3)
    D = gx.FermionOp(ctx, U, dims, Ls, 'mobius', **kw)
    O = Oracle(U, dims, Ls, 'mobius', **kw)
    srcs = [gx.gaussian_fermion(dims, Ls, 352 + k) for k in range(nrhs)]
    fs = [D.field().upload(s) for s in srcs]
    fx = [D.field().upload(0.2 * gx.gaussian_fermion(dims, Ls, 360 + k)) for k in range(nrhs)]
    sts = D.block_cgrq_vec(fs, fx, tol=1e-10, maxit=20000)
    for k in range(nrhs):
        x = fx[k].download()
        r = O.apply('Mdag', O.apply('M', x)) - srcs[k]
        res = np.linalg.norm(r) / np.linalg.norm(srcs[k])
        assert sts[k]['converged'] and res < 2e-10 and abs(res - sts[k]['true_residual']) < 1e-3 * res, (k, res, sts[k])


def test_block_cgrq_vec_matches_slice_form(ctx):
    """On a multi-RHS Wilson operator the list form (single-RHS operator, one field per source)
    and the slice form (one field, right-hand side = fifth-dimension slice) run the same algorithm:
    same iteration count and solutions within rounding."""
    dims, nrhs, mass = (8, 8, 4, 8), 4, 0.1
    U = gx.random_gauge(dims, 371, 0.5)
    src = R.pick(gx.gaussian_fermion(dims, nrhs, 372), dims, ODD)
    B = gx.FermionOp(ctx, U, dims, nrhs, 'wilson', mass=mass)
    fs, fp = B.field(gx.F64, ODD).upload(src), B.field(gx.F64, ODD)
    sl = B.block_cgrq(fs, fp, tol=1e-10)
    D = gx.FermionOp(ctx, U, dims, 1, 'wilson', mass=mass)
    vs = [D.field(gx.F64, ODD).upload(np.ascontiguousarray(src[..., k:k + 1, :, :])) for k in range(nrhs)]
    vx = [D.field(gx.F64, ODD) for _ in range(nrhs)]
    ve = D.block_cgrq_vec(vs, vx, tol=1e-10)
    x = fp.download()
    assert abs(sl[0]['iterations'] - ve[0]['iterations']) <= 1
    for k in range(nrhs):
        assert rel_err(vx[k].download(), x[..., k:k + 1, :, :]) < 1e-9


def test_block_cgrq_vec_errors(ctx):
    dims = (4, 4, 4, 4)
    D = gx.FermionOp(ctx, gx.random_gauge(dims, 381, 0.5), dims, 1, 'wilson', mass=0.1)
    half = [D.field(gx.F64, ODD) for _ in range(2)]
    full = [D.field() for _ in range(2)]
    with pytest.raises(gx.GxError):   # half fields with Mdag M
        D.block_cgrq_vec(half, [D.field(gx.F64, ODD) for _ in range(2)], herm=gx.HERM_MDAGM)
    with pytest.raises(gx.GxError):   # full fields with MpcDagMpc
        D.block_cgrq_vec(full, [D.field() for _ in range(2)], herm=gx.HERM_SCHUR)
    with pytest.raises(gx.GxError):   # a solution that is also a source
        D.block_cgrq_vec(half, [half[1], D.field(gx.F64, ODD)])
    with pytest.raises(gx.GxError):   # zero sources: rank-deficient block
        D.block_cgrq_vec(half, [D.field(gx.F64, ODD) for _ in range(2)])


@pytest.mark.parametrize('inner', [gx.F32, gx.F16])
def test_mixed_cg_batched_wilson(ctx, inner):
    """MixedPrecisionConjugateGradientBatched on a single-RHS Wilson operator (the batch packed into
    one multi-RHS solve): every source reaches the oracle's FP64 CG solution; a nonzero guess is
    honoured; the sources stay unchanged."""
    dims, n, mass = (8, 8, 4, 8), 5, 0.08
    U = gx.random_gauge(dims, 391, 0.5)
    D = gx.FermionOp(ctx, U, dims, 1, 'wilson', mass=mass)
    O = Oracle(U, dims, 1, 'wilson', mass=mass)
    srcs = [R.pick(gx.gaussian_fermion(dims, 1, 392 + k), dims, ODD) for k in range(n)]
    fs = [D.field(gx.F64, ODD).upload(s) for s in srcs]
    fx = [D.field(gx.F64, ODD) for _ in range(n)]
    fx[2].upload(0.1 * srcs[2])
    sts = D.mixed_cg_batched(fs, fx, tol=1e-10, inner_tol=1e-5, inner_prec=inner)
    for k in range(n):
        ref, _ = O.cg(srcs[k], tol=1e-10)
        assert sts[k]['converged'] and sts[k]['true_residual'] < 2e-10, sts[k]
        assert rel_err(fx[k].download(), ref) < 1e-8, k
        assert np.array_equal(fs[k].download(), srcs[k])


def test_mixed_cg_batched_mobius(ctx):
    """On a Mobius operator the batch runs source by source and gives what gx_mixed_cg gives each."""
    dims, Ls, n = (8, 4, 4, 8), 6, 2
    kw = dict(mass=0.05, M5=1.8, b=1.5, c=0.5)
    U = gx.random_gauge(dims, 401, 0.3)
    D = gx.FermionOp(ctx, U, dims, Ls, 'mobius', **kw)
    srcs = [R.pick(gx.gaussian_fermion(dims, Ls, 402 + k), dims, ODD) for k in range(n)]
    fs = [D.field(gx.F64, ODD).upload(s) for s in srcs]
    fx = [D.field(gx.F64, ODD) for _ in range(n)]
    sts = D.mixed_cg_batched(fs, fx, tol=1e-10)
    for k in range(n):
        one = D.field(gx.F64, ODD)
        D.mixed_cg(fs[k], one, tol=1e-10, inner_tol=1e-5)
        assert sts[k]['converged'] and np.array_equal(fx[k].download(), one.download()), k


@pytest.mark.parametrize('subset', [ODD, gx.FULL])
def test_inner_product_and_madd_matrix_lists(ctx, subset):
    """InnerProductMatrix and MaddMatrix over lists of fields against numpy, with R aliasing X or Y."""
    dims, Ls, n = (4, 4, 2, 6), 3, 4
    xs = fields(ctx, dims, Ls, subset, range(410, 410 + n))
    ys = fields(ctx, dims, Ls, subset, range(420, 420 + n))
    X = np.stack([a for _, a in xs]).reshape(n, -1)
    Y = np.stack([a for _, a in ys]).reshape(n, -1)
    m = gx.inner_product_matrix([f for f, _ in xs], [f for f, _ in ys])
    assert np.max(np.abs(m - X.conj() @ Y.T)) < 1e-12 * np.linalg.norm(X) * np.linalg.norm(Y)
    rng = np.random.default_rng(9)
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    ref = Y + 0.5 * (a.T @ X)
    rs = [gx.Field(ctx, dims, Ls, gx.F64, subset) for _ in range(n)]
    gx.madd_matrix(rs, a, [f for f, _ in xs], [f for f, _ in ys], 0.5)
    got = np.stack([r.download() for r in rs]).reshape(n, -1)
    assert np.max(np.abs(got - ref)) < 1e-13 * np.max(np.abs(ref))
    gx.madd_matrix([f for f, _ in xs], a, [f for f, _ in xs], [f for f, _ in ys], 0.5)   # R aliases X
    got = np.stack([f.download() for f, _ in xs]).reshape(n, -1)
    assert np.max(np.abs(got - ref)) < 1e-13 * np.max(np.abs(ref))

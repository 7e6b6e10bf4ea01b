"""DhopDir (the single legs of the hop) and the force DhopDeriv / DhopDerivEO / DhopDerivOE.

Pins: the eight legs sum to the hop (which the oracle pins); one leg on a point source equals the
numpy restatement -1/2 U (1 -+ gamma_mu) e (chiral gamma basis of TwoSpinor.h:67-139); the force
is numpy's colour outer product of the leg with A, and the dagger force the same on gamma5-rotated
fields (gamma5 (1 - gamma_mu) gamma5 = 1 + gamma_mu)."""
import numpy as np
import pytest

import grid_amd as gx

pytestmark = pytest.mark.gpu

GAMMA = [np.array(g, dtype=np.complex128) for g in (
    [[0, 0, 0, 1j], [0, 0, 1j, 0], [0, -1j, 0, 0], [-1j, 0, 0, 0]],
    [[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]],
    [[0, 0, 1j, 0], [0, 0, 0, -1j], [-1j, 0, 0, 0], [0, 1j, 0, 0]],
    [[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]])]
G5 = np.diag([1, 1, -1, -1]).astype(np.complex128)


@pytest.fixture(scope='module')
def ctx():
    c = gx.Context(0)
    yield c
    c.close()


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def op(ctx, kind, Ls, dims, seed):
    U = gx.random_gauge(dims, seed, 0.5)
    return U, gx.FermionOp(ctx, U, dims, Ls, kind, mass=0.1)


@pytest.mark.parametrize('phases', [None, (1, 1, 1, -1), (1j, 1, np.exp(0.3j), -1)])
@pytest.mark.parametrize('kind,Ls', [('wilson', 1), ('mobius', 4)])
def test_dhopdir_legs_sum_to_dhop(ctx, kind, Ls, phases):
    """The eight legs sum to Dhop, also with antiperiodic and twisted boundary phases (the legs read
    the same phase-carrying double-stored links)."""
    dims = (4, 6, 4, 8)
    U = gx.random_gauge(dims, 601, 0.5)
    D = gx.FermionOp(ctx, U, dims, Ls, kind, mass=0.1, phases=phases)
    psi = D.field().upload(gx.gaussian_fermion(dims, Ls, 602))
    total = np.zeros(psi.shape, dtype=np.complex128)
    out = D.field()
    for mu in range(4):
        for disp in (1, -1):
            D.DhopDir(psi, out, mu, disp)
            total += out.download()
    ref = D.field()
    D.Dhop(psi, ref, 0)
    assert rel(total, ref.download()) < 1e-14


@pytest.mark.parametrize('mu', [0, 1, 2, 3])
def test_dhopdir_point_source(ctx, mu):
    """A point source at y in (spin sp, colour c): the +mu leg is -1/2 U_mu(y - mu)[:, c] (1 - g_mu)[:, sp]
    at y - mu only; the -mu leg is -1/2 conj(U_mu(y)[c, :]) (1 + g_mu)[:, sp] at y + mu only."""
    dims = (4, 4, 6, 4)
    U, D = op(ctx, 'wilson', 1, dims, 611 + mu)
    y, sp, c = (1, 2, 3, 2), 1, 2
    src = D.field()
    src.point_source(y, sp, c)
    out = D.field()
    step = np.zeros(4, dtype=int)
    step[mu] = 1
    for disp in (1, -1):
        D.DhopDir(src, out, mu, disp)
        got = out.download()[..., 0, :, :]
        x = tuple((np.array(y) - disp * step) % np.array(dims))
        idx = (x[3], x[2], x[1], x[0])
        one = np.eye(4) - disp * GAMMA[mu]          # (1 - g) forward, (1 + g) backward
        if disp == 1:
            col = U[idx][mu][:, c]
        else:
            col = np.conj(U[(y[3], y[2], y[1], y[0])][mu][c, :])
        expect = -0.5 * np.outer(one[:, sp], col)
        assert np.max(np.abs(got[idx] - expect)) < 1e-15
        mask = np.ones(got.shape[:4], dtype=bool)
        mask[idx] = False
        assert np.count_nonzero(got[mask]) == 0


def outer(chi, A):
    """mat[..., a, b] = sum over s and spin of chi_a conj(A_b)"""
    return np.einsum('...sia,...sib->...ab', chi, np.conj(A))


@pytest.mark.parametrize('kind,Ls', [('wilson', 1), ('mobius', 4)])
def test_dhop_deriv(ctx, kind, Ls):
    dims = (4, 4, 4, 6)
    _, D = op(ctx, kind, Ls, dims, 621)
    a = gx.gaussian_fermion(dims, Ls, 622)
    b = gx.gaussian_fermion(dims, Ls, 623)
    A, B = D.field().upload(a), D.field().upload(b)
    chi = D.field()
    ref = {0: np.zeros((*a.shape[:4], 4, 3, 3), dtype=np.complex128)}
    ref[1] = ref[0].copy()
    g5b = D.field().upload(np.einsum('ij,...jc->...ic', G5, b))
    for mu in range(4):
        D.DhopDir(B, chi, mu, 1)
        ref[0][..., mu, :, :] = outer(chi.download(), a)
        D.DhopDir(g5b, chi, mu, 1)
        ref[1][..., mu, :, :] = outer(np.einsum('ij,...jc->...ic', G5, chi.download()), a)
    t, z, y, x = np.indices(a.shape[:4])
    even = (x + y + z + t) % 2 == 0
    for dag in (0, 1):
        mat = D.DhopDeriv(A, B, dag)
        assert rel(mat, ref[dag]) < 1e-14, dag
        ae, ao = D.field(gx.F64, gx.EVEN), D.field(gx.F64, gx.ODD)
        be, bo = D.field(gx.F64, gx.EVEN), D.field(gx.F64, gx.ODD)
        gx.pick_checkerboard(0, ae, A)
        gx.pick_checkerboard(1, ao, A)
        gx.pick_checkerboard(0, be, B)
        gx.pick_checkerboard(1, bo, B)
        meo = D.DhopDeriv(ae, bo, dag, 'EO')
        moe = D.DhopDeriv(ao, be, dag, 'OE')
        assert rel(meo[even], ref[dag][even]) < 1e-14 and np.count_nonzero(meo[~even]) == 0
        assert rel(moe[~even], ref[dag][~even]) < 1e-14 and np.count_nonzero(moe[even]) == 0
        with pytest.raises(gx.GxError):
            D.DhopDeriv(ao, be, dag, 'EO')          # wrong checkerboards

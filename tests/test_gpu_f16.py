"""FP16 storage (binary16 values with a per-site float scale, FP32 arithmetic): round trips, the
hop and fifth-dimension kernels against the oracle, and FP16-inner defect-correction CG."""
import numpy as np
import pytest

import grid_amd as gx
import npref as R
from conftest import max_site_rel_err, rel_err
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

# binary16 keeps 11 significant bits: a component's error is at most 2^-11 of its site's largest
# component, so a site's relative error is below ~2^-11 * sqrt(24) / sqrt(24) ~ 5e-4 per store.
EPS16 = 2.0 ** -11


@pytest.fixture(scope='module')
def ctx():
    c = gx.Context(0)
    yield c
    c.close()


def half(a, dims, cb):
    return R.pick(a, dims, cb)


@pytest.mark.parametrize('subset', [gx.FULL, gx.ODD])
def test_f16_round_trip(ctx, subset):
    dims, Ls = (8, 4, 4, 6), 3
    a = gx.gaussian_fermion(dims, Ls, 11)
    if subset != gx.FULL:
        a = half(a, dims, subset)
    a[0, 0, 0, 0] *= 1e-6   # a tiny site keeps its relative precision (per-site scale)
    # sites whose largest component is subnormal in FP32 (1/max overflows), one with exact zeros
    a[0, 0, 0, 1] *= 1e-41
    a[0, 0, 1, 0] *= 1e-41
    a[0, 0, 1, 0, :, 1:, :] = 0
    f = gx.Field(ctx, dims, Ls, gx.F16, subset).upload(a)
    b = f.download()
    assert np.all(np.isfinite(b))
    site_max = np.max(np.maximum(np.abs(a.real), np.abs(a.imag)), axis=(-2, -1), keepdims=True)
    # subnormal FP32 values carry fewer mantissa bits: a few 2^-11 of their site's maximum
    tol = np.where(site_max < 1e-37, 4 * EPS16, EPS16) * site_max
    assert np.all(np.abs(b.real - a.real) <= tol + 1e-300)
    assert np.all(np.abs(b.imag - a.imag) <= tol + 1e-300)
    # precision change on the device agrees with the host packing
    g64 = gx.Field(ctx, dims, Ls, gx.F64, subset).upload(a)
    g16 = gx.Field(ctx, dims, Ls, gx.F16, subset)
    gx.precision_change(g16, g64)
    assert np.array_equal(g16.download(), b)
    back = gx.Field(ctx, dims, Ls, gx.F64, subset)
    gx.precision_change(back, g16)
    assert np.array_equal(back.download(), b)
    assert f.norm2() == pytest.approx(np.vdot(b, b).real, rel=1e-6)


@pytest.mark.parametrize('kind,Ls', [('wilson', 1), ('mobius', 8)])
def test_f16_hop_and_mpc(ctx, kind, Ls):
    """DhopEO and the Schur operator in FP16 storage agree with the FP64 oracle to the storage
    precision (per-site relative error of a few 2^-11)."""
    dims = (8, 8, 4, 4)
    U = gx.random_gauge(dims, 71, 0.4)
    D = gx.FermionOp(ctx, U, dims, Ls, kind, mass=0.1)
    O = Oracle(U, dims, Ls, kind, mass=0.1)
    src = half(gx.gaussian_fermion(dims, Ls, 72), dims, gx.ODD)
    fi, fo = D.field(gx.F16, gx.ODD).upload(src), D.field(gx.F16, gx.EVEN)
    D.DhopEO(fi, fo, 0)
    ref = O.dhop(src, 0, 0)
    assert max_site_rel_err(fo.download(), ref) < 8 * EPS16
    for fused in (0, 1):
        gx.tune_set(gx.TUNE_FUSED, fused)
        try:
            mi, mo = D.field(gx.F16, gx.ODD).upload(src), D.field(gx.F16, gx.ODD)
            D.Mpc(mi, mo)
            want = O.schur('Mpc', src)
            assert rel_err(mo.download(), want) < 8 * EPS16
        finally:
            gx.tune_set(gx.TUNE_FUSED, 1)


@pytest.mark.parametrize('kind,Ls', [('wilson', 1), ('mobius', 8)])
def test_f16_inner_mixed_cg(ctx, kind, Ls):
    """Defect correction with FP16-storage inner solves reaches the FP64 tolerance and the FP64 CG
    solution."""
    dims = (8, 8, 8, 4)
    U = gx.random_gauge(dims, 81, 0.3)
    D = gx.FermionOp(ctx, U, dims, Ls, kind, mass=0.05)
    src = half(gx.gaussian_fermion(dims, Ls, 82), dims, gx.ODD)
    fs = D.field(gx.F64, gx.ODD).upload(src)
    x64, x16 = D.field(gx.F64, gx.ODD), D.field(gx.F64, gx.ODD)
    s64 = D.cg(fs, x64, tol=1e-10)
    s16 = D.mixed_cg(fs, x16, tol=1e-10, inner_tol=1e-5, inner_prec=gx.F16)
    assert s16['converged'] and s16['true_residual'] < 1e-9, s16
    assert s16['outer_iterations'] >= 1 and s16['inner_iterations'] > 0
    assert rel_err(x16.download(), x64.download()) < 1e-8


@pytest.mark.parametrize('sloppy', [gx.F32, gx.F16])
@pytest.mark.parametrize('kind,Ls', [('wilson', 1), ('mobius', 8)])
def test_reliable_update_cg(ctx, kind, Ls, sloppy):
    """ConjugateGradientReliableUpdate: sloppy-precision Krylov vectors with FP64 reliable updates
    reach the FP64 tolerance and the FP64 CG solution; with delta = 0.1 several updates happen."""
    dims = (8, 8, 8, 4)
    U = gx.random_gauge(dims, 85, 0.3)
    D = gx.FermionOp(ctx, U, dims, Ls, kind, mass=0.05)
    src = half(gx.gaussian_fermion(dims, Ls, 86), dims, gx.ODD)
    fs = D.field(gx.F64, gx.ODD).upload(src)
    x64, xru = D.field(gx.F64, gx.ODD), D.field(gx.F64, gx.ODD)
    s64 = D.cg(fs, x64, tol=1e-10)
    sru = D.reliable_update_cg(fs, xru, tol=1e-10, delta=0.1, sloppy_prec=sloppy)
    assert sru['converged'] and sru['true_residual'] < 1e-9, sru
    assert sru['outer_iterations'] >= 2, sru          # reliable updates performed
    assert sru['final_iterations'] < 0.2 * s64['iterations'] + 5, (sru, s64)
    assert rel_err(xru.download(), x64.download()) < 1e-8


@pytest.mark.parametrize('kind,Ls', [('wilson', 1), ('mobius', 8)])
def test_reliable_update_fallback(ctx, kind, Ls):
    """setFallbackLinop: FP16 sloppy vectors switch to FP32 below the transition |r|^2.  Transition 0
    never switches (the same digest as plain FP16 reliable updates); a transition above |src|^2
    switches after the first iteration and then matches the FP32 iteration count within a few; a
    middle one switches part way.  Every variant reaches the FP64 tolerance and solution."""
    dims = (8, 8, 8, 4)
    U = gx.random_gauge(dims, 87, 0.3)
    D = gx.FermionOp(ctx, U, dims, Ls, kind, mass=0.05)
    src = half(gx.gaussian_fermion(dims, Ls, 88), dims, gx.ODD)
    fs = D.field(gx.F64, gx.ODD).upload(src)
    ssq = fs.norm2()
    x64 = D.field(gx.F64, gx.ODD)
    s64 = D.cg(fs, x64, tol=1e-10)
    ref = x64.download()
    x16, x32 = D.field(gx.F64, gx.ODD), D.field(gx.F64, gx.ODD)
    s16 = D.reliable_update_cg(fs, x16, tol=1e-10, delta=0.1, sloppy_prec=gx.F16)
    s32 = D.reliable_update_cg(fs, x32, tol=1e-10, delta=0.1, sloppy_prec=gx.F32)
    for ftol in (0.0, 1e-5 * ssq, 10.0 * ssq):
        x = D.field(gx.F64, gx.ODD)
        st = D.reliable_update_cg_fallback(fs, x, tol=1e-10, delta=0.1, fallback_transition_tol=ftol)
        assert st['converged'] and st['true_residual'] < 1e-9, (ftol, st)
        assert rel_err(x.download(), ref) < 1e-8, ftol
        if ftol == 0.0:
            assert x.digest() == x16.digest() and st['inner_iterations'] == s16['inner_iterations'], (st, s16)
        if ftol > ssq:
            assert abs(st['inner_iterations'] - s32['inner_iterations']) <= 3, (st, s32)
    bad = D.field(gx.F64, gx.ODD)
    for args in ((gx.F32, gx.F32, 0.0), (gx.F16, gx.F64, 0.0), (gx.F16, gx.F32, -1.0)):
        with pytest.raises(gx.GxError):
            D.reliable_update_cg_fallback(fs, bad, sloppy_prec=args[0], fallback_prec=args[1],
                                          fallback_transition_tol=args[2])

"""GPU parity at the workloads of the five benchmark configurations, through the C ABI.

Every test builds its inputs with the seeds the benchmark uses, so the kernels, tile maps and
solver paths checked here are the ones that are timed.

  C1  Wilson 8^4, eps 0.5, m 0.1, FP64: DhopEO per site against the oracle; red-black CG to 1e-10
      with the iteration count within 2 of the oracle's CG.
  C2  Mobius 16^4 x Ls16 (b 1.5, c 0.5, M5 1.8, mf 0.01, eps 0.3): DhopEO (both daggers), Meooe,
      MeooeDag, Mpc and MpcDag at the full volume, FP64 (per-site relative error <= 1e-13) and
      FP32 (<= 1e-6), under every hop tile map; the Mobius red-black solve against the oracle's on
      8^4 x Ls16 with the same parameters; the full-size FP64 and mixed solves' residual
      |b - M x| / |b| recomputed by the oracle's M (<= 1e-10).
  C3  Mobius 24^4 x Ls12, mf 0.005: red-black solves in FP64 CG and in defect correction with FP32
      and FP16-storage inner solves; each full-system residual recomputed by the oracle's M
      (<= 1e-10) and each mixed solution within 1e-7 of the FP64 one.
  C4  Wilson 32^3 x 64, eps 0.3, m 0.05, the 12 spin-colour point sources of one seeded site: one
      mixed-precision multi-RHS red-black solve (FP32 inner); every right-hand side's residual
      recomputed by the oracle's Wilson M (<= 1e-10); the FP16-inner solve on a sample of three.
  C5  the serving mix over 8^4, 12^4, 16^4, 16^3 x 32 and 24^4 with 1-12 sources: every request's
      solution bitwise equal to the one it gets served alone, and the oracle's residual of a
      sample of served solutions (<= 1e-10).
"""
import numpy as np
import pytest

import grid_amd as gx
from conftest import max_site_rel_err
from grid_amd import serving
from oracle.pyoracle import Oracle, EVEN, ODD

pytestmark = pytest.mark.gpu

MOBIUS = dict(M5=1.8, b=1.5, c=0.5)
TARGET = 1e-10   # full-system residual every solve must reach
TOL = {gx.F64: 1e-13, gx.F32: 1e-6}


@pytest.fixture(scope='module')
def ctx():
    c = gx.Context(0)
    yield c
    c.close()


def pick(a, dims, cb):
    """Full lexicographic field -> half field of checkerboard cb (Grid red-black order)."""
    lx, ly, lz, lt = dims
    t, z, y = np.meshgrid(np.arange(lt), np.arange(lz), np.arange(ly), indexing='ij')
    off = (cb + t + z + y) % 2
    xs = off[..., None] + 2 * np.arange(lx // 2)[None, None, None, :]
    idx = xs.reshape(lt, lz, ly, lx // 2, *([1] * (a.ndim - 4)))
    return np.ascontiguousarray(np.take_along_axis(a, idx, axis=3))


def full_residual(O, x, b):
    """|b - M x| / |b| with the oracle's M (x, b full host fields of one right-hand side)."""
    r = O.apply('M', x) - b
    return float(np.linalg.norm(r.ravel()) / np.linalg.norm(b.ravel()))


class Knobs:
    def __init__(self, **kw):
        self.kw = {getattr(gx, 'TUNE_' + k.upper()): v for k, v in kw.items()}

    def __enter__(self):
        self.old = {k: gx.tune_get(k) for k in self.kw}
        for k, v in self.kw.items():
            gx.tune_set(k, v)

    def __exit__(self, *a):
        for k, v in self.old.items():
            gx.tune_set(k, v)


# ---------------- C1 ----------------
def test_c1_wilson_8x8x8x8(ctx):
    dims = (8, 8, 8, 8)
    U = gx.random_gauge(dims, 1001, 0.5)
    src = gx.gaussian_fermion(dims, 1, 2001)
    O = Oracle(U, dims, 1, 'wilson', mass=0.1)
    D = gx.FermionOp(ctx, U, dims, 1, 'wilson', mass=0.1)
    po = pick(src, dims, ODD)
    fi, fo = D.field(gx.F64, ODD).upload(po), D.field(gx.F64, EVEN)
    D.DhopEO(fi, fo, 0)
    assert max_site_rel_err(fo.download(), O.dhop(po, EVEN, 0)) < 1e-13
    x = D.field(gx.F64, ODD)
    st = D.cg(fi, x, tol=1e-10, maxit=1000)
    _, ref = O.cg(po, tol=1e-10, maxit=1000)
    assert st['converged'] and st['true_residual'] < 2e-10, st
    assert abs(st['iterations'] - ref['iterations']) <= 2, (st, ref)
    D.close()


# ---------------- C2 ----------------
C2_DIMS, C2_LS = (16, 16, 16, 16), 16


@pytest.fixture(scope='module')
def c2():
    U = gx.random_gauge(C2_DIMS, 1002, 0.3)
    psi = gx.gaussian_fermion(C2_DIMS, C2_LS, 2002)
    O = Oracle(U, C2_DIMS, C2_LS, 'mobius', mass=0.01, **MOBIUS)
    po = pick(psi, C2_DIMS, ODD)
    ref = {('DhopEO', 0): O.dhop(po, EVEN, 0), ('DhopEO', 1): O.dhop(po, EVEN, 1),
           ('Meooe', 0): O.apply('Meooe', po, ODD), ('MeooeDag', 0): O.apply('MeooeDag', po, ODD),
           ('Mpc', 0): O.schur('Mpc', po), ('MpcDag', 0): O.schur('MpcDag', po)}
    return U, psi, po, O, ref


@pytest.mark.parametrize('prec', [gx.F64, gx.F32])
@pytest.mark.parametrize('tile_map,rows', [(0, 0), (1, 0), (2, 0), (3, 0), (4, 0), (1, 1), (4, 1)])
def test_c2_operators_full_volume(ctx, c2, prec, tile_map, rows):
    U, _, po, _, ref = c2
    print(f'C2 16^4xLs16 prec {prec} tile map {tile_map} row tiles {rows}', flush=True)
    D = gx.FermionOp(ctx, U, C2_DIMS, C2_LS, 'mobius', mass=0.01, **MOBIUS)
    tol = TOL[prec]
    fi = D.field(prec, ODD).upload(po)
    with Knobs(dhop_map=tile_map, dhop_pf=1, dhop_nt=1, dhop_tile=rows):
        for order in (0, 1):
            with Knobs(dhop_order=order):
                for dag in (0, 1):
                    fo = D.field(prec, EVEN)
                    D.DhopEO(fi, fo, dag)
                    assert max_site_rel_err(fo.download(), ref[('DhopEO', dag)]) < tol, (order, dag)
        fo = D.field(prec, EVEN)
        D.Meooe(fi, fo)
        assert max_site_rel_err(fo.download(), ref[('Meooe', 0)]) < tol
        D.MeooeDag(fi, fo)
        assert max_site_rel_err(fo.download(), ref[('MeooeDag', 0)]) < tol
        for fused in (1, 0):
            with Knobs(fused=fused):
                for name in ('Mpc', 'MpcDag'):
                    fo = D.field(prec, ODD)
                    getattr(D, name)(fi, fo)
                    assert max_site_rel_err(fo.download(), ref[(name, 0)]) < 10 * tol, (name, fused)
    D.close()


def test_c2_schur_solve_matches_oracle(ctx):
    """Mobius red-black solve with C2's parameters (mf 0.01, Ls 16) on 8^4, FP64 CG and FP32
    defect correction, against the oracle's red-black solve."""
    dims = (8, 8, 8, 8)
    U = gx.random_gauge(dims, 1102, 0.3)
    src = gx.gaussian_fermion(dims, C2_LS, 2102)
    O = Oracle(U, dims, C2_LS, 'mobius', mass=0.01, **MOBIUS)
    ref, rst = O.schur_solve(src, tol=1e-12, maxit=20000)
    assert rst['converged'] and rst['unprec_residual'] < TARGET
    D = gx.FermionOp(ctx, U, dims, C2_LS, 'mobius', mass=0.01, **MOBIUS)
    fs = D.field().upload(src)
    for mixed in (0, 1, 2):
        fx = D.field()
        st = D.schur_solve(fs, fx, tol=1e-12, maxit=20000, mixed=mixed)
        x = fx.download()
        assert st['converged'] and st['unprec_residual'] < TARGET, (mixed, st)
        assert full_residual(O, x, src) < TARGET, mixed
        # FP64 CG follows the oracle's iteration closely; a defect-correction solution differs from
        # the FP64 one by up to the condition number (~1e4 at mf 0.01) times the 1e-12 stopping
        # residuals, whatever the inner arithmetic's summation order
        err = float(np.linalg.norm((x - ref).ravel()) / np.linalg.norm(ref.ravel()))
        assert err < (1e-8 if mixed == 0 else 3e-8), (mixed, err)
        if mixed == 0:
            assert abs(st['iterations'] - rst['iterations']) <= 2, (st, rst)
    D.close()


@pytest.mark.parametrize('mixed', [0, 1])
def test_c2_full_size_solve(ctx, c2, mixed):
    """The benchmark's C2 solves (16^4 x Ls16, Gaussian source): full-system residual recomputed
    by the oracle's M."""
    U, _, _, O, _ = c2
    src = gx.gaussian_fermion(C2_DIMS, C2_LS, 3003)
    D = gx.FermionOp(ctx, U, C2_DIMS, C2_LS, 'mobius', mass=0.01, **MOBIUS)
    fs, fx = D.field().upload(src), D.field()
    st = D.schur_solve(fs, fx, tol=1e-12, maxit=20000, mixed=mixed)
    assert st['converged'], st
    r = full_residual(O, fx.download(), src)
    assert r < TARGET and abs(r - st['unprec_residual']) <= 1e-2 * r, (r, st)
    D.close()


# ---------------- C3 ----------------
def test_c3_mixed_precision_solves(ctx):
    dims, Ls = (24, 24, 24, 24), 12
    kw = dict(mass=0.005, **MOBIUS)
    U = gx.random_gauge(dims, 1003, 0.3)
    D = gx.FermionOp(ctx, U, dims, Ls, 'mobius', **kw)
    src_f = D.field().gaussian(2003)
    src = src_f.download()
    O = Oracle(U, dims, Ls, 'mobius', **kw)
    x64 = None
    for mixed in (0, 1, 2):
        print(f'C3 24^4xLs12 mixed={mixed}', flush=True)
        fx = D.field()
        st = D.schur_solve(src_f, fx, tol=1e-12, maxit=50000, mixed=mixed)
        x = fx.download()
        fx.close()
        r = full_residual(O, x, src)
        assert st['converged'] and r < TARGET, (mixed, r, st)
        if x64 is None:
            x64 = x
        else:
            err = float(np.linalg.norm((x - x64).ravel()) / np.linalg.norm(x64.ravel()))
            assert err < 1e-7, (mixed, err)
    D.close()


# ---------------- C4 ----------------
def test_c4_multi_rhs_point_sources(ctx):
    dims, nrhs, mass = (32, 32, 32, 64), 12, 0.05
    U = gx.random_gauge(dims, 1004, 0.3)
    rng = np.random.default_rng(4004)
    x0 = tuple(int(rng.integers(d)) for d in dims)
    base = gx.FermionOp(ctx, U, dims, 1, 'wilson', mass=mass)
    B = base.clone(ctx, nrhs, mass)
    O = Oracle(U, dims, 1, 'wilson', mass=mass)
    sb = B.field().point_source(x0)
    t4 = base.field()
    for mixed, sample in ((1, range(nrhs)), (2, (0, 5, 11))):
        print(f'C4 32^3x64 12 point sources mixed={mixed}', flush=True)
        xb = B.field()
        st = B.schur_solve(sb, xb, tol=1e-11, maxit=20000, mixed=mixed)
        assert st['converged'] and st['unprec_residual'] < TARGET, st
        for k in sample:
            gx.extract_slice(t4, xb, k)
            b = gx.point_source(dims, 1, x0, k // 3, k % 3)
            r = full_residual(O, t4.download(), b)
            assert r < TARGET, (mixed, k, r)
        xb.close()
    B.close()
    base.close()


# ---------------- C5 ----------------
def test_c5_serving_mix_volumes(ctx):
    vols = serving.VOLUMES
    reqs, t = [], 0.0
    nsrcs = (12, 1, 7, 3, 12, 5, 2, 12, 9, 4)
    masses = (0.05, 0.3, 0.12, 0.2, 0.07, 0.25, 0.05, 0.16, 0.09, 0.28)
    for i in range(10):
        reqs.append(dict(dims=vols[i % 5], nsrc=nsrcs[i], mass=masses[i], gauge=i % 2, seed=7000 + i, arrival=t))
        t += 0.01
    kw = dict(n_gauge=2, gauge_seed=1500, gauge_eps=0.3, tol=1e-11, mixed=2)
    mix, info = serving.serve(reqs, slots=4, budget=64e9, want_solutions=True, **kw)
    assert all(r['status'] == 0 for r in mix), [r['status'] for r in mix]
    assert any(b['t_admit'] < a['t_done'] for a, b in zip(mix, mix[1:]))
    for i, r in enumerate(reqs):
        alone, _ = serving.serve([dict(r, arrival=0.0)], slots=1, budget=64e9, realtime=False, **kw)
        assert alone[0]['status'] == 0 and alone[0]['digest'] == mix[i]['digest'], i
        assert mix[i]['residual'] < TARGET, mix[i]
    # the oracle's residual of the 24^4 request with 12 sources and of the 16^3 x 32 request
    for i in (4, 3):
        r = reqs[i]
        U = gx.random_gauge(r['dims'], 1500 + r['gauge'], 0.3)
        src = gx.gaussian_fermion(r['dims'], r['nsrc'], r['seed'])
        O = Oracle(U, r['dims'], 1, 'wilson', mass=r['mass'])
        for k in range(r['nsrc']):
            x = np.ascontiguousarray(mix[i]['solution'][..., k:k + 1, :, :])
            assert full_residual(O, x, np.ascontiguousarray(src[..., k:k + 1, :, :])) < TARGET, (i, k)

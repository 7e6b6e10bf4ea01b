"""GPU parity: the HIP library (through the C ABI) against the CPU oracle and the golden fixtures.

Tolerances (per site, max over sites of |gpu - ref| / |ref|): 1e-13 for FP64 fields,
1e-6 for FP32 fields (compared with the FP64 oracle).  Solvers: true FP64 residual <= tol and
agreement with the oracle's FP64 CG solution.
"""
import numpy as np
import pytest

import grid_amd as gx
import npref as R
from conftest import max_site_rel_err, rel_err
from oracle.pyoracle import Oracle, EVEN, ODD

pytestmark = pytest.mark.gpu

TOL = {gx.F64: 1e-13, gx.F32: 1e-6}


@pytest.fixture(scope='module')
def ctx():
    c = gx.Context(0)
    yield c
    c.close()


def inputs(dims, Ls, seed, eps=0.5):
    return gx.random_gauge(dims, seed, eps), gx.gaussian_fermion(dims, Ls, seed + 17)


def half(a, dims, cb):
    return R.pick(a, dims, cb)


def run_half(D, prec, fn, psi_h, cb_in):
    i = D.field(prec, cb_in).upload(psi_h)
    o = D.field(prec, 1 - cb_in)
    fn(i, o)
    return o.download()


@pytest.mark.parametrize('dims', [(8, 8, 8, 8), (4, 6, 8, 6), (16, 8, 4, 12)])
@pytest.mark.parametrize('prec', [gx.F64, gx.F32])
def test_wilson_dhop(ctx, dims, prec):
    U, psi = inputs(dims, 1, 11)
    O = Oracle(U, dims, 1, 'wilson', mass=0.1)
    D = gx.FermionOp(ctx, U, dims, 1, 'wilson', mass=0.1)
    for dag in (0, 1):
        fi = D.field(prec).upload(psi)
        fo = D.field(prec)
        D.Dhop(fi, fo, dag)
        assert max_site_rel_err(fo.download(), O.dhop(psi, -1, dag)) < TOL[prec]
        po, pe = half(psi, dims, ODD), half(psi, dims, EVEN)
        out = run_half(D, prec, lambda i, o: D.DhopEO(i, o, dag), po, ODD)
        assert max_site_rel_err(out, O.dhop(po, EVEN, dag)) < TOL[prec]
        out = run_half(D, prec, lambda i, o: D.DhopOE(i, o, dag), pe, EVEN)
        assert max_site_rel_err(out, O.dhop(pe, ODD, dag)) < TOL[prec]
    D.close()


def test_wilson_golden(ctx, wilson_golden):
    G = wilson_golden
    dims = tuple(int(x) for x in G['dims'])
    D = gx.FermionOp(ctx, G['U'], dims, 1, 'wilson', mass=float(G['mass']))
    fi, fo = D.field().upload(G['psi']), D.field()
    D.Dhop(fi, fo, 0)
    assert max_site_rel_err(fo.download(), G['dhop']) < 1e-13
    D.Dhop(fi, fo, 1)
    assert max_site_rel_err(fo.download(), G['dhop_dag']) < 1e-13
    D.M(fi, fo)
    assert max_site_rel_err(fo.download(), G['M']) < 1e-13
    D.Mdag(fi, fo)
    assert max_site_rel_err(fo.download(), G['Mdag']) < 1e-13
    out = run_half(D, gx.F64, lambda i, o: D.DhopEO(i, o, 0), half(G['psi'], dims, ODD), ODD)
    assert max_site_rel_err(out, G['dhop_eo']) < 1e-13
    Dap = gx.FermionOp(ctx, G['U'], dims, 1, 'wilson', mass=0.1, phases=(1, 1, 1, -1))
    Dap.Dhop(fi, fo, 0)
    assert max_site_rel_err(fo.download(), G['dhop_antiperiodic_t']) < 1e-13
    sol = D.field()
    st = D.schur_solve(fi, sol, tol=1e-12)
    assert st['converged'] and st['unprec_residual'] < 1e-10
    assert rel_err(sol.download(), G['x']) < 1e-9


def test_mobius_golden(ctx, mobius_golden):
    G = mobius_golden
    dims = tuple(int(x) for x in G['dims'])
    Ls = int(G['Ls'])
    D = gx.FermionOp(ctx, G['U'], dims, Ls, 'mobius', mass=float(G['mass']), M5=float(G['M5']), b=float(G['b']),
                     c=float(G['c']))
    fi, fo = D.field().upload(G['psi']), D.field()
    D.M(fi, fo)
    assert max_site_rel_err(fo.download(), G['M']) < 1e-13
    D.Mdag(fi, fo)
    assert max_site_rel_err(fo.download(), G['Mdag']) < 1e-13
    po, pe = half(G['psi'], dims, ODD), half(G['psi'], dims, EVEN)
    assert max_site_rel_err(run_half(D, gx.F64, D.Meooe, po, ODD), G['meooe_o']) < 1e-13
    assert max_site_rel_err(run_half(D, gx.F64, D.MeooeDag, po, ODD), G['meooedag_o']) < 1e-13
    for name, key in [('Mooee', 'mooee_e'), ('MooeeDag', 'mooeedag_e'), ('MooeeInv', 'mooeeinv_e'),
                      ('MooeeInvDag', 'mooeeinvdag_e')]:
        i = D.field(gx.F64, EVEN).upload(pe)
        o = D.field(gx.F64, EVEN)
        getattr(D, name)(i, o)
        assert max_site_rel_err(o.download(), G[key]) < 1e-13, name
    for name, key in [('Mpc', 'mpc_o'), ('MpcDag', 'mpcdag_o')]:
        i = D.field(gx.F64, ODD).upload(po)
        o = D.field(gx.F64, ODD)
        getattr(D, name)(i, o)
        assert max_site_rel_err(o.download(), G[key]) < 1e-13, name


@pytest.mark.parametrize('Ls,dims', [(8, (8, 8, 8, 8)), (16, (8, 4, 4, 8)), (12, (4, 6, 4, 8))])
@pytest.mark.parametrize('prec', [gx.F64, gx.F32])
def test_mobius_ops(ctx, Ls, dims, prec):
    U, psi = inputs(dims, Ls, 23, eps=0.3)
    kw = dict(mass=0.01, M5=1.8, b=1.5, c=0.5)
    O = Oracle(U, dims, Ls, 'mobius', **kw)
    D = gx.FermionOp(ctx, U, dims, Ls, 'mobius', **kw)
    tol = TOL[prec]
    fi, fo = D.field(prec).upload(psi), D.field(prec)
    D.M(fi, fo)
    assert max_site_rel_err(fo.download(), O.apply('M', psi)) < tol
    D.Mdag(fi, fo)
    assert max_site_rel_err(fo.download(), O.apply('Mdag', psi)) < tol
    po, pe = half(psi, dims, ODD), half(psi, dims, EVEN)
    for dag in (0, 1):
        out = run_half(D, prec, lambda i, o: D.DhopEO(i, o, dag), po, ODD)
        assert max_site_rel_err(out, O.dhop(po, EVEN, dag)) < tol
    assert max_site_rel_err(run_half(D, prec, D.Meooe, po, ODD), O.apply('Meooe', po, ODD)) < tol
    assert max_site_rel_err(run_half(D, prec, D.MeooeDag, po, ODD), O.apply('MeooeDag', po, ODD)) < tol
    for name in ['Mooee', 'MooeeDag', 'MooeeInv', 'MooeeInvDag', 'Meooe5D', 'MeooeDag5D']:
        i = D.field(prec, EVEN).upload(pe)
        o = D.field(prec, EVEN)
        getattr(D, name)(i, o)
        assert max_site_rel_err(o.download(), O.apply(name, pe, EVEN)) < tol, name
    for name in ['M5D', 'M5Ddag']:   # Grid's M5D(psi, chi) with chi = psi on entry
        i = D.field(prec, EVEN).upload(pe)
        o = D.field(prec, EVEN).upload(pe)
        getattr(D, name)(i, o)
        assert max_site_rel_err(o.download(), O.apply(name, pe, EVEN)) < tol, name
    for name in ['Mpc', 'MpcDag', 'MpcDagMpc']:
        i = D.field(prec, ODD).upload(po)
        o = D.field(prec, ODD)
        getattr(D, name)(i, o)
        assert max_site_rel_err(o.download(), O.schur(name, po)) < tol * 10, name
    D.close()


@pytest.mark.parametrize('fused', [0, 1])
@pytest.mark.parametrize('kind,Ls,dims', [('wilson', 1, (8, 8, 8, 8)), ('wilson', 12, (4, 6, 4, 8)),
                                          ('mobius', 16, (8, 4, 4, 8)), ('mobius', 12, (4, 6, 4, 8)),
                                          ('mobius', 5, (8, 4, 4, 4))])
@pytest.mark.parametrize('prec', [gx.F64, gx.F32])
def test_schur_fused_and_unfused(ctx, fused, kind, Ls, dims, prec):
    U, psi = inputs(dims, Ls, 91, eps=0.3)
    kw = dict(mass=0.05, M5=1.8, b=1.5, c=0.5)
    O = Oracle(U, dims, Ls, kind, **kw)
    D = gx.FermionOp(ctx, U, dims, Ls, kind, **kw)
    po = half(psi, dims, ODD)
    gx.tune_set(gx.TUNE_FUSED, fused)
    try:
        for name in ['Mpc', 'MpcDag', 'MpcDagMpc']:
            i = D.field(prec, ODD).upload(po)
            o = D.field(prec, ODD)
            getattr(D, name)(i, o)
            assert max_site_rel_err(o.download(), O.schur(name, po)) < TOL[prec] * 10, (name, fused)
    finally:
        gx.tune_set(gx.TUNE_FUSED, 1)
    D.close()


@pytest.mark.parametrize('prec', [gx.F64, gx.F32])
def test_physical_source_maps(ctx, prec):
    dims, Ls = (8, 4, 4, 8), 8
    kw = dict(mass=0.05, M5=1.8, b=1.5, c=0.5)
    U = gx.random_gauge(dims, 71, 0.3)
    b4 = gx.gaussian_fermion(dims, 1, 72)
    sol = gx.gaussian_fermion(dims, Ls, 73)
    O = Oracle(U, dims, Ls, 'mobius', **kw)
    D = gx.FermionOp(ctx, U, dims, Ls, 'mobius', **kw)
    f4 = gx.Field(ctx, dims, 1, prec, gx.FULL).upload(b4)
    f5 = D.field(prec)
    D.ImportPhysicalFermionSource(f4, f5)
    assert max_site_rel_err(f5.download(), O.import_physical(b4)) < TOL[prec] * 10
    f5.upload(sol)
    D.ExportPhysicalFermionSolution(f5, f4)
    assert max_site_rel_err(f4.download(), O.export_physical(sol)) < TOL[prec]


@pytest.mark.parametrize('nrhs', [4, 12])
def test_multi_rhs_wilson_matches_single(ctx, nrhs):
    dims = (8, 4, 4, 8)
    U, psi = inputs(dims, nrhs, 31)
    O1 = Oracle(U, dims, 1, 'wilson', mass=0.05)
    D = gx.FermionOp(ctx, U, dims, nrhs, 'wilson', mass=0.05)
    po = half(psi, dims, ODD)
    out = run_half(D, gx.F64, lambda i, o: D.DhopEO(i, o, 0), po, ODD)
    for k in range(nrhs):
        ref = O1.dhop(np.ascontiguousarray(po[..., k:k + 1, :, :]), EVEN, 0)
        assert max_site_rel_err(out[..., k:k + 1, :, :], ref) < 1e-13


def test_reductions(ctx):
    dims = (8, 8, 4, 4)
    x = gx.gaussian_fermion(dims, 4, 1)
    y = gx.gaussian_fermion(dims, 4, 2)
    D = gx.FermionOp(ctx, gx.random_gauge(dims, 5, 0.5), dims, 4, 'wilson')
    fx, fy, fz = D.field().upload(x), D.field().upload(y), D.field()
    assert abs(fx.norm2() - np.vdot(x, x).real) / np.vdot(x, x).real < 1e-14
    ip = gx.inner_product(fx, fy)
    assert abs(ip - np.vdot(x, y)) / abs(np.vdot(x, y)) < 1e-12
    n = gx.axpy_norm(fz, -0.7, fx, fy)
    z = -0.7 * x + y
    assert abs(n - np.vdot(z, z).real) / np.vdot(z, z).real < 1e-14
    assert rel_err(fz.download(), z) < 1e-15
    # bitwise reproducible
    assert all(fx.norm2() == fx.norm2() for _ in range(3))
    assert gx.inner_product(fx, fy) == ip


def test_checkerboard_errors(ctx):
    dims = (4, 4, 4, 4)
    D = gx.FermionOp(ctx, gx.random_gauge(dims, 5, 0.5), dims, 1, 'wilson')
    i, o = D.field(gx.F64, EVEN), D.field(gx.F64, ODD)
    with pytest.raises(gx.GxError) as e:
        D.DhopEO(i, o)
    assert e.value.code == 5
    with pytest.raises(gx.GxError) as e:
        D.Dhop(i, o)
    assert e.value.code == 1


@pytest.mark.parametrize('kind,dims,Ls,mass', [('wilson', (8, 8, 8, 8), 1, 0.1), ('mobius', (8, 8, 4, 4), 8, 0.05)])
def test_cg_matches_oracle(ctx, kind, dims, Ls, mass):
    U = gx.random_gauge(dims, 41, 0.5 if kind == 'wilson' else 0.3)
    src = half(gx.gaussian_fermion(dims, Ls, 42), dims, ODD)
    kw = dict(mass=mass, M5=1.8, b=1.5, c=0.5)
    O = Oracle(U, dims, Ls, kind, **kw)
    ref, rst = O.cg(src, tol=1e-10)
    D = gx.FermionOp(ctx, U, dims, Ls, kind, **kw)
    fs, fp = D.field(gx.F64, ODD).upload(src), D.field(gx.F64, ODD)
    st = D.cg(fs, fp, tol=1e-10)
    assert st['converged'] and st['true_residual'] < 1e-9
    assert abs(st['iterations'] - rst['iterations']) <= 2
    assert rel_err(fp.download(), ref) < 1e-8
    # mixed precision: FP32 inner, FP64 outer; reaches the same FP64 tolerance
    fm = D.field(gx.F64, ODD)
    mst = D.mixed_cg(fs, fm, tol=1e-10, inner_tol=1e-5)
    assert mst['converged'] and mst['true_residual'] < 1e-9
    assert rel_err(fm.download(), ref) < 1e-7


@pytest.mark.parametrize('mixed', [0, 1, 2, 3])
def test_schur_solve_matches_oracle(ctx, mixed):
    """SchurRedBlackDiagMooeeSolve with FP64 CG, defect correction (FP32, FP16 inner) and
    reliable-update CG (mixed = 3) all reach the oracle's red-black solution."""
    dims = (8, 8, 8, 8)
    U, src = inputs(dims, 1, 51)
    O = Oracle(U, dims, 1, 'wilson', mass=0.1)
    ref, rst = O.schur_solve(src, tol=1e-10)
    D = gx.FermionOp(ctx, U, dims, 1, 'wilson', mass=0.1)
    fs, fsol = D.field().upload(src), D.field()
    st = D.schur_solve(fs, fsol, tol=1e-10, mixed=mixed)
    assert st['converged'] and st['unprec_residual'] < 1e-8
    assert rel_err(fsol.download(), ref) < 1e-8
    if mixed == 3:
        assert st['outer_iterations'] >= 1 and st['inner_iterations'] > 0, st   # reliable updates ran
        M = gx.FermionOp(ctx, U, dims, 4, 'wilson', mass=0.1)
        with pytest.raises(gx.GxError):        # multi-RHS operators take CGmultiRHS forms only
            M.schur_solve(M.field(), M.field(), mixed=3)


@pytest.mark.parametrize('kind,Ls', [('wilson', 1), ('mobius', 8), ('wilson', 4)])
def test_graph_replay_matches_eager(ctx, kind, Ls):
    """CG iteration batches replayed as hipGraphs give bitwise the eager result, on every replay,
    with eager work and synchronisations between the replays."""
    dims = (8, 8, 4, 4)
    U = gx.random_gauge(dims, 81, 0.4)
    D = gx.FermionOp(ctx, U, dims, Ls, kind, mass=0.1)
    srcs = [half(gx.gaussian_fermion(dims, Ls, 82 + i), dims, ODD) for i in range(3)]
    multi = kind == 'wilson' and Ls > 1
    solve = (lambda s, x: D.block_cg(s, x, tol=1e-9)) if multi else (lambda s, x: D.cg(s, x, tol=1e-9))
    eager = []
    gx.tune_set(gx.TUNE_CG_GRAPH, 0)
    try:
        for s in srcs:
            fs, fx = D.field(gx.F64, ODD).upload(s), D.field(gx.F64, ODD)
            solve(fs, fx)
            eager.append(fx.download())
        gx.tune_set(gx.TUNE_CG_GRAPH, 1)
        fs, fx = D.field(gx.F64, ODD), D.field(gx.F64, ODD)
        for rep in range(2):
            for s, ref in zip(srcs, eager):
                fs.upload(s)
                fx.zero()
                solve(fs, fx)
                ctx.sync()
                assert np.array_equal(fx.download(), ref), (rep, kind, Ls)
                fs.norm2()  # eager work between replays
    finally:
        gx.tune_set(gx.TUNE_CG_GRAPH, 1)


@pytest.mark.parametrize('kind,Ls', [('wilson', 1), ('mobius', 4)])
def test_graph_follows_mass_and_gauge(ctx, kind, Ls):
    """A captured CG batch bakes in the links and the mass: after set_mass or import_gauge the
    graph-mode solve must agree bitwise with an eager solve of the changed operator."""
    dims = (8, 4, 4, 4)
    U1, U2 = gx.random_gauge(dims, 91, 0.4), gx.random_gauge(dims, 92, 0.4)
    src = half(gx.gaussian_fermion(dims, Ls, 93), dims, ODD)
    D = gx.FermionOp(ctx, U1, dims, Ls, kind, mass=0.1)
    fs, fx = D.field(gx.F64, ODD).upload(src), D.field(gx.F64, ODD)

    def solve():
        fx.zero()
        D.cg(fs, fx, tol=1e-9)
        ctx.sync()
        return fx.download()

    try:
        gx.tune_set(gx.TUNE_CG_GRAPH, 1)
        solve()                      # capture with (U1, m=0.1)
        D.set_mass(0.2)
        got_m = solve()
        D.import_gauge(U2)
        got_u = solve()
        gx.tune_set(gx.TUNE_CG_GRAPH, 0)
        D.import_gauge(U1)
        want_m = solve()             # eager (U1, m=0.2)
        D.import_gauge(U2)
        want_u = solve()             # eager (U2, m=0.2)
    finally:
        gx.tune_set(gx.TUNE_CG_GRAPH, 1)
    assert np.array_equal(got_m, want_m)
    assert np.array_equal(got_u, want_u)
    assert not np.array_equal(got_m, got_u)


@pytest.mark.parametrize('inner', [gx.F64, gx.F32, gx.F16])
def test_block_cg_matches_single_rhs(ctx, inner):
    """CGmultiRHS: every right-hand side agrees with the oracle's single-RHS FP64 CG (FP64, or FP32
    and FP16-storage inner solves under FP64 defect correction)."""
    dims, nrhs, mass = (8, 8, 4, 8), 4, 0.1
    U = gx.random_gauge(dims, 61, 0.5)
    src = half(gx.gaussian_fermion(dims, nrhs, 62), dims, ODD)
    D = gx.FermionOp(ctx, U, dims, nrhs, 'wilson', mass=mass)
    fs, fp = D.field(gx.F64, ODD).upload(src), D.field(gx.F64, ODD)
    sts = D.block_cg(fs, fp, tol=1e-10, inner_prec=inner)
    x = fp.download()
    O1 = Oracle(U, dims, 1, 'wilson', mass=mass)
    for k in range(nrhs):
        ref, rst = O1.cg(np.ascontiguousarray(src[..., k:k + 1, :, :]), tol=1e-10)
        assert sts[k]['converged'] and sts[k]['true_residual'] < 1e-9, sts[k]
        assert rel_err(x[..., k:k + 1, :, :], ref) < 1e-8


@pytest.mark.parametrize('inner', [gx.F64, gx.F32])
def test_block_cg_from_guess(ctx, inner):
    """A nonzero starting guess enters through its residual src - A psi: the solves reach the
    oracle's solutions (which start from zero)."""
    dims, nrhs, mass = (8, 8, 4, 8), 3, 0.1
    U = gx.random_gauge(dims, 63, 0.5)
    src = half(gx.gaussian_fermion(dims, nrhs, 64), dims, ODD)
    guess = 0.3 * half(gx.gaussian_fermion(dims, nrhs, 65), dims, ODD)
    D = gx.FermionOp(ctx, U, dims, nrhs, 'wilson', mass=mass)
    fs, fp = D.field(gx.F64, ODD).upload(src), D.field(gx.F64, ODD).upload(guess)
    sts = D.block_cg(fs, fp, tol=1e-10, inner_prec=inner)
    x = fp.download()
    O1 = Oracle(U, dims, 1, 'wilson', mass=mass)
    for k in range(nrhs):
        ref, _ = O1.cg(np.ascontiguousarray(src[..., k:k + 1, :, :]), tol=1e-10)
        assert sts[k]['converged'] and sts[k]['true_residual'] < 1e-9, sts[k]
        assert rel_err(x[..., k:k + 1, :, :], ref) < 1e-8


@pytest.mark.parametrize('inner', [gx.F32, gx.F16])
def test_mixed_cg_from_guess(ctx, inner):
    """Defect correction from a nonzero guess, and from the zero guess, reaches the oracle's CG
    solution; the final FP64 solve starts from the last outer residual."""
    dims, mass = (8, 8, 8, 8), 0.1
    U = gx.random_gauge(dims, 66, 0.5)
    src = half(gx.gaussian_fermion(dims, 1, 67), dims, ODD)
    ref, _ = Oracle(U, dims, 1, 'wilson', mass=mass).cg(src, tol=1e-10)
    D = gx.FermionOp(ctx, U, dims, 1, 'wilson', mass=mass)
    fs = D.field(gx.F64, ODD).upload(src)
    for guess in (np.zeros_like(src), 0.3 * half(gx.gaussian_fermion(dims, 1, 68), dims, ODD)):
        fx = D.field(gx.F64, ODD).upload(guess)
        st = D.mixed_cg(fs, fx, tol=1e-10, inner_tol=1e-5, inner_prec=inner)
        assert st['converged'] and st['true_residual'] < 1e-9, st
        assert st['outer_iterations'] >= 1 and st['inner_iterations'] > 0
        assert rel_err(fx.download(), ref) < 1e-7


@pytest.mark.parametrize('kind,dims,Ls,prec', [('mobius', (8, 8, 8, 8), 16, gx.F64), ('wilson', (8, 4, 8, 8), 12, gx.F32),
                                               ('wilson', (8, 8, 4, 4), 1, gx.F64), ('mobius', (4, 6, 4, 8), 5, gx.F16)])
def test_tile_maps_bitwise(ctx, kind, dims, Ls, prec):
    """Every workgroup->tile map (linear, XCD slabs, XCD runs, blocked sweep) only reorders the
    work: the hop's output is bitwise the same."""
    U, psi = inputs(dims, Ls, 121)
    D = gx.FermionOp(ctx, U, dims, Ls, kind, mass=0.1)
    fi, fo = D.field(prec, ODD).upload(half(psi, dims, ODD)), D.field(prec, EVEN)
    saved = gx.tune_get(gx.TUNE_DHOP_MAP)
    outs = []
    try:
        for m in range(5):
            gx.tune_set(gx.TUNE_DHOP_MAP, m)
            fo.zero()
            D.DhopEO(fi, fo, 0)
            outs.append(fo.download())
    finally:
        gx.tune_set(gx.TUNE_DHOP_MAP, saved)
    for m in range(1, 5):
        assert np.array_equal(outs[m], outs[0]), m
    O = Oracle(U, dims, Ls, kind, mass=0.1)
    assert max_site_rel_err(outs[0], O.dhop(half(psi, dims, ODD), EVEN, 0)) < {gx.F64: 1e-13, gx.F32: 1e-6, gx.F16: 5e-3}[prec]


class _Knobs:
    def __init__(self, **kw):
        self.kw = {getattr(gx, 'TUNE_' + k.upper()): v for k, v in kw.items()}

    def __enter__(self):
        self.old = {k: gx.tune_get(k) for k in self.kw}
        for k, v in self.kw.items():
            gx.tune_set(k, v)

    def __exit__(self, *a):
        for k, v in self.old.items():
            gx.tune_set(k, v)


@pytest.mark.parametrize('Ls,dims', [(24, (8, 4, 4, 8)), (20, (4, 4, 4, 8)), (32, (4, 4, 2, 8)), (48, (4, 2, 2, 4))])
@pytest.mark.parametrize('prec', [gx.F64, gx.F32])
def test_mobius_ops_large_ls(ctx, Ls, dims, prec):
    """Mobius operators with Ls > 16 (Grid runs physical-point DWF at Ls = 24,
    tests/Test_dwf_mixedcg_prec_halfcomms.cc:43) through the general-extent kernels, against the oracle."""
    U, psi = inputs(dims, Ls, 27, eps=0.3)
    kw = dict(mass=0.01, M5=1.8, b=1.5, c=0.5)
    O = Oracle(U, dims, Ls, 'mobius', **kw)
    D = gx.FermionOp(ctx, U, dims, Ls, 'mobius', **kw)
    tol = TOL[prec]
    fi, fo = D.field(prec).upload(psi), D.field(prec)
    D.M(fi, fo)
    assert max_site_rel_err(fo.download(), O.apply('M', psi)) < tol
    D.Mdag(fi, fo)
    assert max_site_rel_err(fo.download(), O.apply('Mdag', psi)) < tol
    po, pe = half(psi, dims, ODD), half(psi, dims, EVEN)
    for dag in (0, 1):
        out = run_half(D, prec, lambda i, o: D.DhopEO(i, o, dag), po, ODD)
        assert max_site_rel_err(out, O.dhop(po, EVEN, dag)) < tol
    assert max_site_rel_err(run_half(D, prec, D.Meooe, po, ODD), O.apply('Meooe', po, ODD)) < tol
    assert max_site_rel_err(run_half(D, prec, D.MeooeDag, po, ODD), O.apply('MeooeDag', po, ODD)) < tol
    for name in ['Mooee', 'MooeeDag', 'MooeeInv', 'MooeeInvDag', 'Meooe5D', 'MeooeDag5D']:
        i = D.field(prec, EVEN).upload(pe)
        o = D.field(prec, EVEN)
        getattr(D, name)(i, o)
        assert max_site_rel_err(o.download(), O.apply(name, pe, EVEN)) < tol, name
    for fused in (1, 0):
        with _Knobs(fused=fused):
            for name in ['Mpc', 'MpcDag', 'MpcDagMpc']:
                i = D.field(prec, ODD).upload(po)
                o = D.field(prec, ODD)
                getattr(D, name)(i, o)
                assert max_site_rel_err(o.download(), O.schur(name, po)) < tol * 10, (name, fused)
    D.close()


def test_mobius_ls24_red_black_solves(ctx):
    """Red-black solves at Ls = 24 (FP64 CG, FP32 and FP16-storage defect correction) reach the
    oracle's solution."""
    dims, Ls = (4, 4, 4, 8), 24
    kw = dict(mass=0.01, M5=1.8, b=1.5, c=0.5)
    U = gx.random_gauge(dims, 61, 0.3)
    src = gx.gaussian_fermion(dims, Ls, 62)
    O = Oracle(U, dims, Ls, 'mobius', **kw)
    ref, rst = O.schur_solve(src, tol=1e-12, maxit=20000)
    D = gx.FermionOp(ctx, U, dims, Ls, 'mobius', **kw)
    fs = D.field().upload(src)
    for mixed in (0, 1, 2):
        fx = D.field()
        st = D.schur_solve(fs, fx, tol=1e-12, maxit=20000, mixed=mixed)
        assert st['converged'] and st['unprec_residual'] < 1e-10, (mixed, st)
        assert rel_err(fx.download(), ref) < 1e-8, mixed
        if mixed == 0:
            assert abs(st['iterations'] - rst['iterations']) <= 2, (st, rst)
    D.close()


@pytest.mark.parametrize('kind,Ls,dims', [('mobius', 8, (8, 4, 4, 8)), ('mobius', 12, (4, 6, 4, 8)),
                                          ('mobius', 16, (8, 4, 4, 8)), ('wilson', 12, (8, 4, 6, 8)),
                                          ('wilson', 16, (8, 4, 4, 8)), ('wilson', 8, (8, 8, 4, 4))])
@pytest.mark.parametrize('prec', [gx.F64, gx.F32, gx.F16])
def test_tile_shapes_bitwise(ctx, kind, Ls, dims, prec):
    """Every tile shape (whole sites per wave, or W waves of 64 consecutive entries, one workgroup
    per tile or a persistent grid, or 2 x 2 row tiles staged in LDS) and tile map computes the same
    legs in the same order: DhopEO, Mpc and MpcDag are bitwise equal across shapes."""
    from grid_amd.autotune import shape_candidates
    U, psi = inputs(dims, Ls, 29, eps=0.3)
    kw = dict(mass=0.05, M5=1.8, b=1.5, c=0.5)
    D = gx.FermionOp(ctx, U, dims, Ls, kind, **kw)
    po = half(psi, dims, ODD)
    fi = D.field(prec, ODD).upload(po)
    ref = None
    shapes = [(w, p, 0) for w in shape_candidates(Ls, prec) for p in ((0, 1) if w else (0,))] + [(0, 0, 1)]
    for w, persist, rows in shapes:
        for tile_map in (0, 4):
            with _Knobs(dhop_shape=w, dhop_persist=persist, dhop_map=tile_map, dhop_order=0, fused=1, dhop_tile=rows):
                outs = []
                for fn in (lambda o: D.DhopEO(fi, o, 0), lambda o: D.DhopEO(fi, o, 1), lambda o: D.Mpc(fi, o),
                           lambda o: D.MpcDag(fi, o)):
                    fo = D.field(prec, EVEN if len(outs) < 2 else ODD)
                    fn(fo)
                    outs.append(fo.download())
            if ref is None:
                ref = outs
            else:
                for a, b in zip(outs, ref):
                    assert np.array_equal(a, b), (w, persist, tile_map, rows)
    D.close()


@pytest.mark.parametrize('nrhs', [24, 48])
def test_multi_rhs_wilson_large(ctx, nrhs):
    """A multi-RHS Wilson operator with more than 16 right-hand sides: the hop of every right-hand
    side equals the oracle's single-RHS hop, and the mixed-precision multi-RHS red-black solve
    brings every right-hand side to the tolerance."""
    dims, mass = (8, 4, 4, 8), 0.1
    U, psi = inputs(dims, nrhs, 33)
    O1 = Oracle(U, dims, 1, 'wilson', mass=mass)
    D = gx.FermionOp(ctx, U, dims, nrhs, 'wilson', mass=mass)
    po = half(psi, dims, ODD)
    out = run_half(D, gx.F64, lambda i, o: D.DhopEO(i, o, 0), po, ODD)
    for k in range(nrhs):
        ref = O1.dhop(np.ascontiguousarray(po[..., k:k + 1, :, :]), EVEN, 0)
        assert max_site_rel_err(out[..., k:k + 1, :, :], ref) < 1e-13
    fs, fx = D.field().upload(psi), D.field()
    st = D.schur_solve(fs, fx, tol=1e-11, maxit=5000, mixed=1)
    assert st['converged'] and st['unprec_residual'] < 1e-10, st
    x = fx.download()
    for k in (0, nrhs // 2, nrhs - 1):
        Mx = O1.apply('M', np.ascontiguousarray(x[..., k:k + 1, :, :]))
        assert rel_err(Mx, psi[..., k:k + 1, :, :]) < 1e-10, k
    D.close()


@pytest.mark.parametrize('nrhs,dims', [(4, (8, 4, 4, 8)), (12, (8, 4, 6, 8)), (16, (4, 4, 4, 8))])
def test_mfma_multi_rhs_hop(ctx, nrhs, dims):
    """The FP32 multi-RHS Wilson hop with the link products on the matrix cores: DhopEO (both
    daggers) and the fused Mpc agree with the oracle's per-right-hand-side results to the FP32
    tolerance, and a mixed-precision multi-RHS solve with it brings every right-hand side to the
    FP64 tolerance with the solution of the VALU kernel."""
    mass = 0.1
    U, psi = inputs(dims, nrhs, 37)
    O1 = Oracle(U, dims, 1, 'wilson', mass=mass)
    D = gx.FermionOp(ctx, U, dims, nrhs, 'wilson', mass=mass)
    po = half(psi, dims, ODD)
    with _Knobs(dhop_mfma=1):
        for dag in (0, 1):
            out = run_half(D, gx.F32, lambda i, o: D.DhopEO(i, o, dag), po, ODD)
            for k in range(nrhs):
                ref = O1.dhop(np.ascontiguousarray(po[..., k:k + 1, :, :]), EVEN, dag)
                assert max_site_rel_err(out[..., k:k + 1, :, :], ref) < 1e-6, (dag, k)
        i, o = D.field(gx.F32, ODD).upload(po), D.field(gx.F32, ODD)
        D.Mpc(i, o)
        mpc = o.download()
        for k in range(nrhs):
            ref = O1.schur('Mpc', np.ascontiguousarray(po[..., k:k + 1, :, :]))
            assert max_site_rel_err(mpc[..., k:k + 1, :, :], ref) < 1e-5, k
        fs, fx = D.field().upload(psi), D.field()
        st = D.schur_solve(fs, fx, tol=1e-11, maxit=5000, mixed=1)
        x_mfma = fx.download()
    fx2 = D.field()
    D.schur_solve(fs, fx2, tol=1e-11, maxit=5000, mixed=1)
    assert st['converged'] and st['unprec_residual'] < 1e-10, st
    assert rel_err(x_mfma, fx2.download()) < 1e-9
    D.close()

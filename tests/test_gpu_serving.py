"""Serving path on the GPU: device sources, digests, link-sharing clones, the multi-RHS red-black
solve, and the gx_serve runtime (continuous admission, overlap, alone-vs-mix identity)."""
import numpy as np
import pytest

import grid_amd as gx
from grid_amd import serving
from conftest import rel_err
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def ctx():
    c = gx.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize('prec,tol', [(gx.F64, 1e-13), (gx.F32, 1e-6)])
def test_device_gaussian_matches_host(ctx, prec, tol):
    dims, Ls = (8, 4, 6, 4), 3
    host = gx.gaussian_fermion(dims, Ls, 4242)
    f = gx.Field(ctx, dims, Ls, prec).gaussian(4242)
    got = f.download()
    assert np.max(np.abs(got - host)) <= tol * np.max(np.abs(host))
    # the numbering is the host's: a different seed gives a different field
    assert rel_err(gx.Field(ctx, dims, Ls, prec).gaussian(4243).download(), host) > 0.5


def test_digest_tracks_bits(ctx):
    dims = (4, 4, 4, 4)
    a = gx.gaussian_fermion(dims, 2, 7)
    f, g = gx.Field(ctx, dims, 2).upload(a), gx.Field(ctx, dims, 2).upload(a)
    assert f.digest() == g.digest() == f.digest()
    b = a.copy()
    b.flat[123] = np.nextafter(b.flat[123].real, np.inf) + 1j * b.flat[123].imag  # one ulp
    assert g.upload(b).digest() != f.digest()
    c = a.copy()
    c.flat[[5, 9]] = c.flat[[9, 5]]  # same multiset of values, other positions
    assert g.upload(c).digest() != f.digest()


def test_clone_matches_fresh_operator(ctx):
    """A clone reading a resident operator's links computes bitwise what a freshly built operator
    on the same links computes, holds no links of its own, and follows share_links."""
    dims, nrhs, mass = (8, 4, 4, 8), 3, 0.17
    U1, U2 = gx.random_gauge(dims, 31, 0.4), gx.random_gauge(dims, 32, 0.4)
    base1 = gx.FermionOp(ctx, U1, dims, 1, 'wilson', mass=0.1)
    base2 = gx.FermionOp(ctx, U2, dims, 1, 'wilson', mass=0.1)
    ctx2 = gx.Context(0)
    cl = base1.clone(ctx2, nrhs, mass)
    src = gx.gaussian_fermion(dims, nrhs, 33)
    for U, base in ((U1, None), (U2, base2)):
        if base is not None:
            cl.share_links(base)
        fresh = gx.FermionOp(ctx, U, dims, nrhs, 'wilson', mass=mass)
        a_in, a_out = cl.field().upload(src), cl.field()
        b_in, b_out = fresh.field().upload(src), fresh.field()
        cl.M(a_in, a_out)
        fresh.M(b_in, b_out)
        ctx2.sync()
        assert np.array_equal(a_out.download(), b_out.download())
        assert cl.footprint() < fresh.footprint()
    ctx2.close()


def test_base_with_clones_refuses_import_and_destroy(ctx):
    """While a clone reads a base operator's links, the base refuses ImportGauge and destroy with
    GX_EINVAL (its links stay valid); once the clone is gone both work.  Only error codes are
    checked: the refused calls do no device work."""
    dims = (4, 4, 4, 4)
    U1, U2 = gx.random_gauge(dims, 51, 0.4), gx.random_gauge(dims, 52, 0.4)
    base = gx.FermionOp(ctx, U1, dims, 1, 'wilson', mass=0.1)
    cl = base.clone(ctx, 2, 0.2)
    for call in (lambda: base.import_gauge(U2), base.close):
        with pytest.raises(gx.GxError) as e:
            call()
        assert e.value.code == 1  # GX_EINVAL
    assert base.h
    cl2 = base.clone(ctx, 3, 0.3)
    cl.close()
    with pytest.raises(gx.GxError):
        base.close()
    cl2.close()
    base.import_gauge(U2)
    base.close()


def test_multi_rhs_schur_solve(ctx):
    """Red-black solve of a multi-RHS Wilson operator (FP64, FP32 and FP16-storage inner solves):
    every right-hand side agrees with the oracle's single-RHS solve and the reported residual is the
    worst per-RHS one."""
    dims, nrhs, mass = (8, 8, 4, 4), 3, 0.12
    U = gx.random_gauge(dims, 41, 0.4)
    src = gx.gaussian_fermion(dims, nrhs, 42)
    D = gx.FermionOp(ctx, U, dims, nrhs, 'wilson', mass=mass)
    O = Oracle(U, dims, 1, 'wilson', mass=mass)
    for mixed in (0, 1, 2):
        fs, fx = D.field().upload(src), D.field()
        st = D.schur_solve(fs, fx, tol=1e-10, mixed=mixed)
        x = fx.download()
        assert st['converged'] and st['true_residual'] < 1e-9 and st['unprec_residual'] < 1e-8, st
        worst = 0.0
        for k in range(nrhs):
            ref, _ = O.schur_solve(np.ascontiguousarray(src[..., k:k + 1, :, :]), tol=1e-10)
            assert rel_err(x[..., k:k + 1, :, :], ref) < 1e-8
            Mx = O.apply('M', np.ascontiguousarray(x[..., k:k + 1, :, :]))
            worst = max(worst, rel_err(Mx, src[..., k:k + 1, :, :]))
        assert abs(worst - st['unprec_residual']) <= 1e-3 * worst + 1e-15


@pytest.mark.parametrize('mixed', [1, 2])
def test_serve_footprint_matches_allocation(ctx, mixed):
    """The admission footprint equals what a served request really holds on the device (FP32 and
    FP16-storage inner solves)."""
    dims, nrhs = (8, 4, 4, 8), 5
    base = gx.FermionOp(ctx, gx.random_gauge(dims, 51, 0.3), dims, 1, 'wilson', mass=0.1)
    cl = base.clone(ctx, nrhs, 0.2)
    fs, fx = cl.field().gaussian(52), cl.field()
    cl.schur_solve(fs, fx, tol=1e-10, mixed=mixed)
    held = cl.footprint() + 2 * (2 * 12 * 16 * np.prod(dims) // 2 * nrhs)
    assert held == serving.footprint(dims, nrhs, mixed=mixed)


SMALL = ((4, 4, 4, 4), (8, 4, 4, 4), (4, 4, 4, 8), (8, 8, 4, 4))


@pytest.mark.parametrize('mixed', [1, 2])
def test_serve_alone_equals_mix(ctx, mixed):
    """Continuous batching: requests admitted while others run give bitwise the result they give
    alone, every one reaches its tolerance, and the served solutions agree with the oracle (FP32
    and FP16-storage inner solves)."""
    reqs = serving.make_trace(14, rate=400.0, seed=77, volumes=SMALL, nsrc_range=(1, 4), n_gauge=2)
    kw = dict(n_gauge=2, gauge_seed=900, gauge_eps=0.4, tol=1e-10, mixed=mixed)
    mix, info = serving.serve(reqs, slots=3, budget=2e9, want_solutions=True, **kw)
    assert all(r['status'] == 0 for r in mix), [r['status'] for r in mix]
    assert len({r['slot'] for r in mix}) > 1
    # some request was admitted before an earlier one finished: solves overlapped
    assert any(b['t_admit'] < a['t_done'] for a, b in zip(mix, mix[1:]))
    for i, r in enumerate(reqs):
        alone, _ = serving.serve([dict(r, arrival=0.0)], slots=1, budget=2e9, realtime=False, **kw)
        assert alone[0]['status'] == 0
        assert alone[0]['digest'] == mix[i]['digest'], i
        assert mix[i]['solver_residual'] < 1e-9 and mix[i]['residual'] < 1e-8, mix[i]
    # two served solutions against the oracle (host gauge and sources from the same recipes)
    for i in (0, len(reqs) - 1):
        r = reqs[i]
        U = gx.random_gauge(r['dims'], 900 + r['gauge'], 0.4)
        src = gx.gaussian_fermion(r['dims'], r['nsrc'], r['seed'])
        O = Oracle(U, r['dims'], 1, 'wilson', mass=r['mass'])
        for k in range(r['nsrc']):
            ref, _ = O.schur_solve(np.ascontiguousarray(src[..., k:k + 1, :, :]), tol=1e-10)
            assert rel_err(mix[i]['solution'][..., k:k + 1, :, :], ref) < 1e-7
    s = serving.summarize(mix, info)
    assert s['completed'] == len(reqs) and s['p99_latency_s'] >= s['p50_latency_s'] > 0


def test_serve_admission_respects_budget(ctx):
    """With a budget that fits one request at a time, requests run one by one (no two solves
    overlap) and a request larger than the whole budget is refused, not run."""
    dims = (8, 4, 4, 4)
    one = serving.footprint(dims, 2, mixed=True)
    reqs = [dict(dims=dims, nsrc=2, mass=0.2, gauge=0, seed=10 + i, arrival=0.0) for i in range(4)]
    reqs.append(dict(dims=(8, 8, 8, 8), nsrc=4, mass=0.2, gauge=0, seed=99, arrival=0.0))
    res, _ = serving.serve(reqs, slots=3, budget=1.5 * one, n_gauge=1, realtime=False)
    ok = sorted((r for r in res if r['status'] == 0), key=lambda r: r['t_admit'])
    assert len(ok) == 4 and res[-1]['status'] == 2  # GX_ENOMEM
    for a, b in zip(ok, ok[1:]):
        assert b['t_admit'] >= a['t_done'] - 1e-6 or b['slot'] == a['slot']


def test_device_point_sources_and_slices(ctx):
    """Device point sources equal the host ones; the spin-colour batch puts component s in slice
    s; ExtractSlice/InsertSlice move one fifth-dimension slice bitwise."""
    dims, x = (4, 6, 4, 8), (3, 1, 2, 5)
    for prec in (gx.F64, gx.F32):
        f = gx.Field(ctx, dims, 3, prec).point_source(x, 2, 1)
        assert np.array_equal(f.download(), gx.point_source(dims, 3, x, 2, 1))
        b = gx.Field(ctx, dims, 12, prec).point_source(x).download()
        for s in range(12):
            want = gx.point_source(dims, 1, x, s // 3, s % 3)
            assert np.array_equal(b[..., s:s + 1, :, :], want)
    a = gx.gaussian_fermion(dims, 5, 8)
    f5, f4 = gx.Field(ctx, dims, 5).upload(a), gx.Field(ctx, dims, 1)
    gx.extract_slice(f4, f5, 3)
    assert np.array_equal(f4.download(), a[..., 3:4, :, :])
    g5 = gx.Field(ctx, dims, 5)
    g5.zero()
    gx.insert_slice(f4, g5, 1)
    got = g5.download()
    assert np.array_equal(got[..., 1:2, :, :], a[..., 3:4, :, :]) and not np.any(got[..., [0, 2, 3, 4], :, :])


def test_serve_reads_nersc_gauge(ctx, tmp_path):
    """Resident gauge fields from NERSC files (gauge_dir): files holding the synthetic fields give
    bitwise the synthetic run's solutions; another configuration in one file changes exactly the
    requests on that field; a missing file is refused before any solve."""
    dims = (8, 8, 8, 8)
    reqs = [dict(dims=dims, nsrc=2, mass=0.1, gauge=i % 2, seed=300 + i, arrival=0.0) for i in range(4)]
    kw = dict(slots=2, budget=2e9, n_gauge=2, gauge_seed=910, gauge_eps=0.3, tol=1e-10, mixed=2, realtime=False)
    for g in range(2):
        gx.nersc_write(ctx, str(tmp_path / f'ckpoint_lat.8x8x8x8.{g}'), gx.random_gauge(dims, 910 + g, 0.3), dims)
    syn, _ = serving.serve(reqs, **kw)
    real, _ = serving.serve(reqs, gauge_dir=str(tmp_path), **kw)
    assert all(r['status'] == 0 for r in syn + real)
    assert [r['digest'] for r in real] == [r['digest'] for r in syn]
    gx.nersc_write(ctx, str(tmp_path / 'ckpoint_lat.8x8x8x8.0'), gx.random_gauge(dims, 2000, 0.3), dims, two_row=True)
    other, _ = serving.serve(reqs, gauge_dir=str(tmp_path), **kw)
    assert all(r['status'] == 0 and r['residual'] < 1e-8 for r in other)
    for r, o in zip(reqs, other):
        same = o['digest'] == syn[reqs.index(r)]['digest']
        assert same == (r['gauge'] == 1), r
    with pytest.raises(gx.GxError):
        serving.serve(reqs, gauge_dir=str(tmp_path / 'missing'), **kw)


@pytest.mark.parametrize('mixed', [1, 2])
def test_serve_fuses_masses(ctx, mixed):
    """fuse_rhs: requests waiting for the same lattice and gauge field join one multi-RHS solve
    with a mass per slice.  Every request reaches the tolerance; a fused request's solution agrees
    with the one it gets alone within the solve tolerance's reach (the fused solve stops when its
    hardest right-hand side converges, so the two differ in the last iterations); a request served
    in a solve of its own gives bitwise its alone result; a fused solution satisfies the oracle's
    M of its own mass."""
    dims = (8, 4, 4, 8)
    rng = np.random.default_rng(7)
    reqs = [dict(dims=dims, nsrc=int(rng.integers(1, 5)), mass=float(rng.uniform(0.05, 0.3)),
                 gauge=int(i % 3 == 2), seed=700 + i, arrival=0.0) for i in range(9)]
    kw = dict(n_gauge=2, gauge_seed=930, gauge_eps=0.4, tol=1e-10, mixed=mixed, realtime=False, want_solutions=True)
    mix, _ = serving.serve(reqs, slots=2, budget=4e9, fuse_rhs=12, **kw)
    assert all(r['status'] == 0 and r['residual'] < 1e-9 for r in mix), [(r['status'], r['residual']) for r in mix]
    assert any(r['batch'] > 1 for r in mix)
    for i, r in enumerate(reqs):
        alone, _ = serving.serve([r], slots=1, budget=4e9, **kw)
        assert alone[0]['status'] == 0 and alone[0]['batch'] == 1
        if mix[i]['batch'] == 1:
            assert alone[0]['digest'] == mix[i]['digest'], i
        else:
            assert rel_err(mix[i]['solution'], alone[0]['solution']) < 1e-7, i
    i = next(i for i, m in enumerate(mix) if m['batch'] > 1)
    r = reqs[i]
    O = Oracle(gx.random_gauge(dims, 930 + r['gauge'], 0.4), dims, 1, 'wilson', mass=r['mass'])
    src = gx.gaussian_fermion(dims, r['nsrc'], r['seed'])
    for k in range(r['nsrc']):
        b = np.ascontiguousarray(src[..., k:k + 1, :, :])
        x = np.ascontiguousarray(mix[i]['solution'][..., k:k + 1, :, :])
        assert np.linalg.norm(O.apply('M', x) - b) / np.linalg.norm(b) < 1e-9


def test_slice_sources_and_digests(ctx):
    """gaussian_slices puts the n-slice field of a seed into slices s0.. of a wider field (bitwise
    the narrow field's values); digest_slices of those slices equals the narrow field's slice
    digest and changes when one value in them changes."""
    dims = (4, 6, 4, 8)
    narrow = gx.Field(ctx, dims, 3).gaussian(91)
    wide = gx.Field(ctx, dims, 7)
    wide.zero()
    wide.gaussian_slices(91, 2, 3)
    w = wide.download()
    assert np.array_equal(w[..., 2:5, :, :], narrow.download())
    assert not np.any(w[..., [0, 1, 5, 6], :, :])
    assert wide.digest_slices(2, 3) == narrow.digest_slices(0, 3)
    assert wide.digest_slices(1, 3) != narrow.digest_slices(0, 3)
    w[0, 0, 0, 0, 3, 1, 2] += 1e-9
    assert gx.Field(ctx, dims, 7).upload(w).digest_slices(2, 3) != narrow.digest_slices(0, 3)
    with pytest.raises(gx.GxError):
        wide.gaussian_slices(91, 5, 3)

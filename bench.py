"""Benchmark: Mobius domain-wall even-odd Dslash (DhopEO) on one MI355X.

Workload (the reference's Mobius configuration): 16^4 x Ls=16, b=1.5, c=0.5, M5=1.8, mf=0.01,
random SU(3) links exp(i 0.3 H), Gaussian 5D source, FP64 (the reported value) and FP32.
A step is one DhopEO over the whole odd->even half lattice, inputs resident in HBM.

value = algorithmic (compulsory) bytes per call / average kernel time, with
bytes = H (24 Ls + 72) c   (H = V4/2 half sites, c = bytes per complex): each input spinor and
each link read once, each output spinor written once.  Kernel time comes from HIP events on the
stream the kernels are launched on.  Also reported: GF/s at 1320 flop per 5D site, the FP32
figure, Schur CG time-to-solution in FP64 and in defect correction with FP32 and with FP16-storage
inner solves, and the CPU oracle's DhopEO on the
same workload (cpu_baseline).  Every solve runs its red-black CG to 1e-11 (Wilson) or 1e-12
(Mobius) on the normal equations so that the full-system residual |b - M x| / |b| (reported as
unprec_residual, with meets_target) is <= 1e-10.

Further legs, reported beside the headline value (each can be switched off):
  c3     Mobius 24^4 x Ls12 (mf 0.005): red-black solve in FP64 CG, in defect correction with
         FP32 and FP16-storage inner solves, and in reliable-update CG;
  c4     Wilson 32^3x64 (m 0.05), the 12 spin-colour point sources of one site: one mixed-precision
         multi-RHS solve against 12 sequential FP64 solves;
  serve  a Poisson trace of Wilson solve requests (volumes 8^4 .. 24^4, 1-12 sources, masses
         0.05-0.3, 8 resident gauge fields per volume) through the serving runtime: throughput and
         p50/p99 latency (--serve-requests 4096 for the full trace);
  sweep  the serving trace at several arrival rates (20-60/s, 4 slots), with fused solves
         (--sweep-fused-rates) and with 8 slots (--sweep-wide-rates): capacity against p50/p99;
  graphs hipGraph-captured CG batches against eager ones (C1, C2 FP32, C4 12-RHS FP32);
  large  DhopEO at >= 1 GB working sets (Mobius 24^4 x Ls16 FP64, 32^4 x Ls16 FP32);
  ms     multishift CG, Mobius 16^4 x Ls16, 12 shifts, against one plain CG solve;
  block  BlockCGrQ against CGmultiRHS on the c4 red-black system.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--legs dhop,large,cg,c3,c4,serve,sweep,...]
                       [--no-tune | --knobs map,pf,nt,order[,shape,persist] | --knobs f64=.../f32=.../f16=...]
                       [--serve-requests N] [--serve-rate R] [--serve-slots S]
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
DIMS = (16, 16, 16, 16)
LS = 16
PARAMS = dict(mass=0.01, M5=1.8, b=1.5, c=0.5)
EPS = 0.3
METRIC = 'Dslash GB/s (% HBM peak); CG time-to-1e-10 mixed vs FP64; solve p50/p99'


def dhop_kernel_name(gx):
    """Name of the kernel the FP64 DhopEO launches under the current knobs (as rocprofv3 lists it)."""
    if gx.tune_get(gx.TUNE_DHOP_TILE):
        return 'gx::k_dhop_rows<double, false, 0>'
    if gx.tune_get(gx.TUNE_DHOP_SHAPE):
        return 'gx::k_dhop_g<double, false, 0>'
    pf = 'true' if gx.tune_get(gx.TUNE_DHOP_PF) else 'false'
    return f'gx::k_dhop<double, {LS}, false, {pf}, {gx.tune_get(gx.TUNE_DHOP_ORDER)}, 0>'
# Every solve must reach |b - M x| / |b| <= 1e-10 on the full (unpreconditioned) system.  The
# red-black CG's tolerance applies to the normal equations of the Schur system, whose residual
# runs below the full-system one by a factor the operator sets: 1.5-2x for these Wilson systems,
# 10-50x for the light-quark Mobius ones.  The solves use 1e-11 (Wilson) and 1e-12 (Mobius), the
# same for the FP64 baseline and the mixed solves, and report the full residual.
CG_TOL = 1e-11
CG_TOL_MOBIUS = 1e-12
TARGET_RESIDUAL = 1e-10


def pick(a, dims, cb):
    """Full lexicographic field -> half field of checkerboard cb (Grid red-black order)."""
    lx, ly, lz, lt = dims
    t, z, y = np.meshgrid(np.arange(lt), np.arange(lz), np.arange(ly), indexing='ij')
    off = (cb + t + z + y) % 2
    xs = off[..., None] + 2 * np.arange(lx // 2)[None, None, None, :]
    idx = xs.reshape(lt, lz, ly, lx // 2, *([1] * (a.ndim - 4)))
    return np.ascontiguousarray(np.take_along_axis(a, idx, axis=3))


def traffic_from_profile(kernel, shape=0):
    """Measured HBM bytes per launch of the FP64 DhopEO kernel `kernel` at tile shape `shape` from the
    committed rocprofv3 counter summaries (profiles/dhop_mobius_16x16x16x16_ls16_f64*_traffic.json,
    one per kernel and shape); (bytes, file) or (None, None) when no summary covers them."""
    import glob
    for p in sorted(glob.glob(os.path.join(ROOT, 'profiles', 'dhop_mobius_16x16x16x16_ls16_f64*_traffic.json'))):
        try:
            with open(p) as f:
                d = json.load(f)
        except Exception:
            continue
        if d.get('kernel') == kernel and int(d.get('shape', 0)) == int(shape):
            return d.get('traffic_bytes_per_launch'), os.path.relpath(p, ROOT)
    return None, None


def time_dhop(ctx, D, fi, fo, steps, warmup):
    for _ in range(warmup):
        D.DhopEO(fi, fo, 0)
    ctx.sync()
    ctx.record(0)
    for _ in range(steps):
        D.DhopEO(fi, fo, 0)
    ctx.record(1)
    ctx.sync()
    return ctx.elapsed_ms(0, 1) / steps


def physical_cores():
    """Physical cores of the host (lscpu: cores per socket x sockets), or os.cpu_count()."""
    try:
        out = subprocess.run(['lscpu'], capture_output=True, text=True, timeout=10).stdout
        f = {k.strip(): v.strip() for k, v in (ln.split(':', 1) for ln in out.splitlines() if ':' in ln)}
        return int(f['Core(s) per socket']) * int(f['Socket(s)'])
    except Exception:
        return os.cpu_count() or 1


def cpu_threads():
    """Threads for the all-cores CPU run: the host's physical cores, capped at the CPU share this
    job is given (OMP_NUM_THREADS when set), since more threads than the share only time-slice."""
    cores = physical_cores()
    share = int(os.environ.get('OMP_NUM_THREADS', '0') or 0)
    return (min(cores, share) if share > 0 else cores), cores


def median_time(fn, calls, warm=2):
    for _ in range(warm):
        fn()
    ts = []
    for _ in range(calls):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


C2_CPU_ITERS = 40


def cpu_baseline(gx, U, psi_o, c2_iterations=463):
    """The CPU restatement of Grid's algorithm (oracle/cpu_ref.cpp, OpenMP, -O3 -march) timed on this
    host: C2 DhopEO median over 20 calls with all physical cores (capped at the job's CPU share) and
    with one thread; red-black solve time-to-solution of C1 and C4 (one point source) with all
    cores, at the tolerances of the GPU legs, and of C2 projected from its first iterations.  About
    30-40 s of CPU work."""
    from oracle import pyoracle
    nthr, cores = cpu_threads()
    H = np.prod(DIMS) // 2
    byts = H * (24 * LS + 72) * 16
    O = pyoracle.Oracle(U, DIMS, LS, 'mobius', **PARAMS)
    out = {'threads_all': nthr, 'host_physical_cores': cores}
    for name, n in (('all_cores', nthr), ('one_thread', 1)):
        pyoracle.set_threads(n)
        sec = median_time(lambda: O.dhop(psi_o, 0, 0), 20, warm=1)
        out[name] = {'threads': n, 'dhop_eo_median_ms': round(sec * 1e3, 3), 'gbs': round(byts / sec / 1e9, 3),
                     'gflops_1320': round(1320.0 * H * LS / sec / 1e9, 3),
                     'gflops_1344': round(1344.0 * H * LS / sec / 1e9, 3)}
    pyoracle.set_threads(nthr)
    solves = {}
    # C2: the GPU cg leg's source and tolerance.  A full CPU solve takes about 80 s on 16 cores, so
    # the sample is the first C2_CPU_ITERS iterations of the same solve, projected to the iteration
    # count of the converged solve (463 on this system: the GPU FP64 solve, whose count the oracle's
    # CG matches within 2 iterations).
    src = gx.gaussian_fermion(DIMS, LS, 3003)
    t0 = time.perf_counter()
    _, st = O.schur_solve(src, tol=CG_TOL_MOBIUS, maxit=C2_CPU_ITERS)
    sec = time.perf_counter() - t0
    per_it = sec / max(st['iterations'], 1)
    solves['c2_mobius_16x16x16x16_ls16'] = {'sample_seconds': round(sec, 3), 'sample_iterations': st['iterations'],
                                            'seconds_per_iteration': round(per_it, 5),
                                            'iterations_to_solution': c2_iterations,
                                            'projected_seconds': round(per_it * c2_iterations, 2),
                                            'cg_tol': CG_TOL_MOBIUS}
    del O
    # C1: Wilson 8^4, m 0.1, eps 0.5
    d1 = (8, 8, 8, 8)
    O1 = pyoracle.Oracle(gx.random_gauge(d1, 1001, 0.5), d1, 1, 'wilson', mass=0.1)
    src1 = gx.gaussian_fermion(d1, 1, 2001)
    t0 = time.perf_counter()
    _, st = O1.schur_solve(src1, tol=CG_TOL, maxit=20000)
    solves['c1_wilson_8x8x8x8'] = {'seconds': round(time.perf_counter() - t0, 4), 'iterations': st['iterations'],
                                   'unprec_residual': st['unprec_residual'], 'cg_tol': CG_TOL}
    # C4: Wilson 32^3x64, m 0.05, one of the 12 point sources (spin 0, colour 0)
    d4 = (32, 32, 32, 64)
    rng = np.random.default_rng(4004)
    x0 = tuple(int(rng.integers(d)) for d in d4)
    O4 = pyoracle.Oracle(gx.random_gauge(d4, 1004, 0.3), d4, 1, 'wilson', mass=0.05)
    src4 = gx.point_source(d4, 1, x0, 0, 0)
    t0 = time.perf_counter()
    _, st = O4.schur_solve(src4, tol=CG_TOL, maxit=20000)
    solves['c4_wilson_32x32x32x64_1rhs'] = {'seconds': round(time.perf_counter() - t0, 3), 'iterations': st['iterations'],
                                            'unprec_residual': st['unprec_residual'], 'cg_tol': CG_TOL}
    out['solves'] = solves
    return out, byts


def bench_dhop_large(gx, ctx, tune):
    """DhopEO at working sets well above the 256 MB Infinity Cache, so HBM sets the rate: Mobius
    24^4 x Ls16 FP64 (1.21 GB per call) and 32^4 x Ls16 FP32 (1.91 GB), C2's parameters."""
    out = {}
    for dims, prec, name, c in (((24, 24, 24, 24), gx.F64, 'mobius_24x24x24x24_ls16_f64', 16),
                                ((32, 32, 32, 32), gx.F32, 'mobius_32x32x32x32_ls16_f32', 8)):
        U = gx.random_gauge(dims, 1012, EPS)
        D = gx.FermionOp(ctx, U, dims, LS, 'mobius', **PARAMS)
        tune('dhop_' + name, D, prec)
        fi = D.field(prec, 1)
        full = D.field(prec)
        full.gaussian(2012)
        gx.pick_checkerboard(1, fi, full)
        full.close()
        fo = D.field(prec, 0)
        ms = time_dhop(ctx, D, fi, fo, 50, 5)
        H = np.prod(dims) // 2
        byts = H * (24 * LS + 72) * c
        out[name] = {'ms': round(ms, 5), 'bytes': int(byts), 'gbs': round(byts / (ms * 1e-3) / 1e9, 2),
                     'frac_of_8tbs': round(byts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     'gflops_1320': round(1320.0 * H * LS / (ms * 1e-3) / 1e9, 1)}
        fi.close()
        fo.close()
        D.close()
    return out


def bench_mixed_dwf(gx, ctx, dims=(24, 24, 24, 24), Ls=12, tol=CG_TOL_MOBIUS, tune=None):
    """Mobius 24^4 x Ls12 (b=1.5, c=0.5, M5=1.8, mf=0.005, eps 0.3), Gaussian 5D source: red-black
    solve in FP64 CG, in defect correction with FP32 and with FP16-storage inner solves (first
    inner tolerance 1e-5 and 1e-3, FP64 outer residual and patch-up), and in reliable-update CG (FP32 Krylov
    vectors, FP64 reliable updates with delta 0.01, FP64 clean-up), each to |b - M x| / |b| <= 1e-10.
    Wall clock around each solve."""
    kw = dict(mass=0.005, M5=1.8, b=1.5, c=0.5)
    U = gx.random_gauge(dims, 1003, 0.3)
    D = gx.FermionOp(ctx, U, dims, Ls, 'mobius', **kw)
    if tune:  # each precision's knobs on this operator: FP64 outer, FP32 and FP16 inner
        tune('c3_f64', D, gx.F64, rounds=2, cg_iters=32)
        tune('c3', D, gx.F32, rounds=2, cg_iters=32)
        tune('c3_f16', D, gx.F16, rounds=2, cg_iters=32)
    src = D.field().gaussian(2003)
    out, ref = {'workload': 'Mobius 24^4 x Ls12 mf=0.005, Gaussian source', 'cg_tol': tol,
                'target_unprec_residual': TARGET_RESIDUAL}, None
    for name, mixed in (('fp64', 0), ('mixed_f32', 1), ('mixed_f16', 2), ('reliable_update_f32', 3)):
        x = D.field()
        ctx.sync()
        t0 = time.perf_counter()
        st = D.schur_solve(src, x, tol=tol, maxit=50000, mixed=mixed)
        ctx.sync()
        sec = time.perf_counter() - t0
        d = dict(seconds=round(sec, 4), iterations=st['iterations'], inner_iterations=st['inner_iterations'],
                 outer=st['outer_iterations'], final_iterations=st['final_iterations'],
                 true_residual=st['true_residual'], unprec_residual=st['unprec_residual'],
                 meets_target=bool(st['unprec_residual'] <= TARGET_RESIDUAL))
        if ref is None:
            ref = x
        else:
            t = D.field()
            gx.axpby(t, 1.0, -1.0, x, ref)
            d['rel_err_vs_fp64'] = (t.norm2() / ref.norm2()) ** 0.5
            d['speedup_vs_fp64'] = round(out['fp64']['seconds'] / sec, 3)
        out[name] = d
    return out


def bench_multishift(gx, ctx, dims=(16, 16, 16, 16), Ls=16, nshift=12, tol=1e-10):
    """Leg 'ms': ConjugateGradientMultiShift on the Mobius 16^4 x Ls16 Schur system (m 0.01,
    eps 0.3, Gaussian odd source), nshift shifts geometrically spaced in [1e-4, 3], every shift to
    tol, against one plain CG solve of the lightest system.  Wall clock, best of two."""
    U = gx.random_gauge(dims, 1002, 0.3)
    D = gx.FermionOp(ctx, U, dims, Ls, 'mobius', mass=0.01, M5=1.8, b=1.5, c=0.5)
    full = D.field().gaussian(2002)
    src = D.field(gx.F64, gx.ODD)
    gx.pick_checkerboard(1, src, full)
    out = {'workload': f'Mobius {dims[0]}^4 x Ls{Ls} m=0.01, {nshift} shifts in [1e-4, 3]', 'tol': tol}
    best = None
    for _ in range(2):
        x = D.field(gx.F64, gx.ODD)
        ctx.sync()
        t0 = time.perf_counter()
        st = D.cg(src, x, tol=tol, maxit=20000)
        ctx.sync()
        sec = time.perf_counter() - t0
        best = sec if best is None else min(best, sec)
    out['cg'] = {'seconds': round(best, 4), 'iterations': st['iterations'],
                 'ms_per_iteration': round(1e3 * best / st['iterations'], 4)}
    shifts = list(np.geomspace(1e-4, 3.0, nshift))
    psis = [D.field(gx.F64, gx.ODD) for _ in range(nshift)]
    best = None
    for _ in range(2):
        ctx.sync()
        t0 = time.perf_counter()
        sts = D.cg_multishift(src, psis, shifts, [tol] * nshift, maxit=20000)
        ctx.sync()
        sec = time.perf_counter() - t0
        best = sec if best is None else min(best, sec)
    k = max(s['iterations'] for s in sts)
    worst = max(s['true_residual'] for s in sts)
    # the multishift vector update's (k_ms_update) algorithmic bytes over the solve: per iteration r read, p read and written, and
    # psi_s, p_s read and written for each shift still iterating (FP64 half fields)
    fbytes = (np.prod(dims) // 2) * Ls * 12 * 16
    its = [s['iterations'] for s in sts]
    vec_bytes = fbytes * (3 * k + 4 * sum(its))
    out['multishift'] = {'seconds': round(best, 4), 'iterations': k, 'ms_per_iteration': round(1e3 * best / k, 4),
                         'shift_iterations': its, 'update_bytes_per_solve': int(vec_bytes),
                         'update_bytes_per_iteration_avg': int(vec_bytes / k),
                         'max_true_residual': worst, 'converged': all(s['converged'] for s in sts),
                         'vs_cg_per_iteration': round((best / k) / (out['cg']['seconds'] / out['cg']['iterations']), 3)}
    return out


def bench_block_cgrq(gx, ctx, dims=(32, 32, 32, 64), mass=0.05, eps=0.3, tol=1e-11):
    """Leg 'block': the C4 Schur system (Wilson 32^3x64 m 0.05, the red-black sources of the
    12 point sources of one site) solved by CGmultiRHS and by BlockCGrQ in FP64, to tol; wall clock,
    iterations and the agreement of the two solutions."""
    nrhs = 12
    U = gx.random_gauge(dims, 1004, eps)
    rng = np.random.default_rng(4004)
    x0 = tuple(int(rng.integers(d)) for d in dims)
    B = gx.FermionOp(ctx, U, dims, nrhs, 'wilson', mass=mass)
    full = B.field().point_source(x0)
    se, so = B.field(gx.F64, gx.EVEN), B.field(gx.F64, gx.ODD)
    gx.pick_checkerboard(0, se, full)
    gx.pick_checkerboard(1, so, full)
    te, to, so2, b = B.field(gx.F64, gx.EVEN), B.field(gx.F64, gx.ODD), B.field(gx.F64, gx.ODD), B.field(gx.F64, gx.ODD)
    B.MooeeInv(se, te)
    B.Meooe(te, to)
    gx.axpby(so2, 1.0, -1.0, so, to)
    B.MpcDag(so2, b)          # RedBlackSource
    out = {'workload': 'Wilson 32^3x64 m=0.05, 12 point sources (red-black system)', 'tol': tol}
    sols = {}
    for name in ('cgmultirhs', 'blockcgrq'):
        best = None
        for _ in range(2):
            x = B.field(gx.F64, gx.ODD)
            ctx.sync()
            t0 = time.perf_counter()
            st = B.block_cg(b, x, tol=tol, maxit=20000) if name == 'cgmultirhs' else B.block_cgrq(b, x, tol=tol, maxit=20000)
            ctx.sync()
            sec = time.perf_counter() - t0
            best = sec if best is None else min(best, sec)
        sols[name] = x
        out[name] = {'seconds': round(best, 4), 'iterations': max(s['iterations'] for s in st),
                     'max_true_residual': max(s['true_residual'] for s in st)}
    t = B.field(gx.F64, gx.ODD)
    gx.axpby(t, 1.0, -1.0, sols['blockcgrq'], sols['cgmultirhs'])
    out['rel_diff'] = (t.norm2() / sols['cgmultirhs'].norm2()) ** 0.5
    return out


def bench_multi_rhs(gx, ctx, dims=(32, 32, 32, 64), mass=0.05, eps=0.3, tol=CG_TOL, tune=None):
    """Wilson 32^3x64 (eps 0.3, m 0.05), the 12 spin-colour point sources of one seeded site:
    mixed-precision multi-RHS red-black solves (CGmultiRHS inner solves in FP32, and in FP16
    storage, sharing every link load across the 12 sources, FP64 defect correction and patch-up)
    against 12 sequential FP64 red-black CG solves, all to |b - M x| / |b| <= 1e-10 per source.
    Wall clock around each solve;
    'speedup' is the faster mixed variant's (named in 'speedup_variant')."""
    nrhs = 12
    U = gx.random_gauge(dims, 1004, eps)
    rng = np.random.default_rng(4004)
    x0 = tuple(int(rng.integers(d)) for d in dims)
    base = gx.FermionOp(ctx, U, dims, 1, 'wilson', mass=mass)
    if tune:
        tune('c4_sequential', base, gx.F64, rounds=2, cg_iters=32)
    s1 = base.field()
    seq, seq_sec, seq_it, seq_res = [], 0.0, [], 0.0
    for k in range(nrhs):
        s1.point_source(x0, k // 3, k % 3)
        x = base.field()
        ctx.sync()
        t0 = time.perf_counter()
        st = base.schur_solve(s1, x, tol=tol, maxit=20000, mixed=False)
        ctx.sync()
        seq_sec += time.perf_counter() - t0
        seq_it.append(st['iterations'])
        seq_res = max(seq_res, st['unprec_residual'])
        seq.append(x)
    B = base.clone(ctx, nrhs, mass)
    if tune:  # the 12-RHS FP64 outer operator's own knobs (the inner precisions are tuned below)
        tune('c4_f64', B, gx.F64, dhop_iters=5, rounds=2, cg_iters=8)
    sb = B.field().point_source(x0)
    t4 = base.field()
    out = {'workload': 'Wilson 32^3x64 m=0.05, 12 point sources', 'origin': list(x0),
           'cg_tol': tol, 'target_unprec_residual': TARGET_RESIDUAL,
           'sequential_fp64': {'seconds': round(seq_sec, 4), 'per_rhs_seconds': round(seq_sec / nrhs, 4),
                               'iterations': seq_it, 'max_unprec_residual': seq_res,
                               'meets_target': bool(seq_res <= TARGET_RESIDUAL)}}
    best = None
    for name, mixed, prec, tkey in (('multi_rhs_mixed', 1, gx.F32, 'c4'), ('multi_rhs_mixed_f16', 2, gx.F16, 'c4_f16')):
        if tune:
            tune(tkey, B, prec, dhop_iters=5, rounds=2, cg_iters=16)
        xb = B.field()
        ctx.sync()
        t0 = time.perf_counter()
        stb = B.schur_solve(sb, xb, tol=tol, maxit=20000, mixed=mixed)
        ctx.sync()
        blk_sec = time.perf_counter() - t0
        worst = 0.0
        for k in range(nrhs):
            gx.extract_slice(t4, xb, k)
            gx.axpby(t4, 1.0, -1.0, t4, seq[k])
            worst = max(worst, (t4.norm2() / seq[k].norm2()) ** 0.5)
        xb.close()
        out[name] = {'inner_storage': {gx.F32: 'fp32', gx.F16: 'fp16'}[prec],
                     'seconds': round(blk_sec, 4), 'per_rhs_seconds': round(blk_sec / nrhs, 4),
                     'inner_iterations': stb['inner_iterations'], 'outer': stb['outer_iterations'],
                     'final_iterations': stb['final_iterations'], 'true_residual': stb['true_residual'],
                     'unprec_residual': stb['unprec_residual'],
                     'meets_target': bool(stb['unprec_residual'] <= TARGET_RESIDUAL),
                     'speedup': round(seq_sec / blk_sec, 3),
                     'max_rel_diff_vs_sequential': worst}
        if best is None or blk_sec < out[best]['seconds']:
            best = name
    out['speedup'] = out[best]['speedup']
    out['speedup_variant'] = best
    # the FP32 12-RHS DhopEO (3.02 GB per call) with the link products on the VALU and on the
    # matrix cores (v_mfma_f32_16x16x4f32), tuned knobs otherwise
    H = np.prod(dims) // 2
    byts = H * (24 * nrhs + 72) * 8
    fi, fo = B.field(gx.F32, gx.ODD), B.field(gx.F32, gx.EVEN)
    full = B.field(gx.F32)
    full.gaussian(2030)
    gx.pick_checkerboard(1, fi, full)
    full.close()
    hop = {}
    for name, mf in (('valu', 0), ('mfma', 1)):
        gx.tune_set(gx.TUNE_DHOP_MFMA, mf)
        ms = time_dhop(ctx, B, fi, fo, 20, 3)
        hop[name] = {'ms': round(ms, 4), 'gbs': round(byts / (ms * 1e-3) / 1e9, 1),
                     'frac_of_8tbs': round(byts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    gx.tune_set(gx.TUNE_DHOP_MFMA, 0)
    out['dhop_eo_12rhs_f32'] = dict(hop, bytes=int(byts))
    fi.close()
    fo.close()
    return out


def bench_graphs(gx, ctx):
    """Leg 'graphs': CG iterations replayed as captured hipGraphs against plain launches, on three
    workloads: C1 (Wilson 8^4 FP64, launch-bound), C2 (Mobius 16^4 x Ls16 FP32) and the C4 lattice
    (Wilson 32^3x64 with 12 right-hand sides as one FP32 CG).  A fixed number of iterations
    (tolerance 0), best of two runs, ms per iteration; the other knobs at the library defaults."""
    out = {}
    cases = (('c1_wilson_8x8x8x8_f64', (8, 8, 8, 8), 1, 'wilson', dict(mass=0.1), gx.F64, 0.5, 400),
             ('c2_mobius_16x16x16x16_ls16_f32', (16, 16, 16, 16), 16, 'mobius', dict(mass=0.01, **{
                 k: v for k, v in PARAMS.items() if k != 'mass'}), gx.F32, EPS, 100),
             ('c4_wilson_32x32x32x64_12rhs_f32', (32, 32, 32, 64), 12, 'wilson', dict(mass=0.05), gx.F32, EPS, 24))
    for name, dims, Ls, kind, kw, prec, eps, iters in cases:
        D = gx.FermionOp(ctx, gx.random_gauge(dims, 1020, eps), dims, Ls, kind, **kw)
        full = D.field(prec)
        full.gaussian(2020)
        src, x = D.field(prec, 1), D.field(prec, 1)
        gx.pick_checkerboard(1, src, full)
        full.close()
        res = {}
        for graph in (0, 1):
            gx.tune_set(gx.TUNE_CG_GRAPH, graph)
            best = None
            for _ in range(2):
                x.zero()
                ctx.sync()
                t0 = time.perf_counter()
                D.cg(src, x, tol=0.0, maxit=iters)
                ctx.sync()
                sec = time.perf_counter() - t0
                best = sec if best is None else min(best, sec)
            res['graph' if graph else 'eager'] = round(1e3 * best / iters, 4)
        res['graph_speedup'] = round(res['eager'] / res['graph'], 3)
        out[name] = res
        for f in (src, x):
            f.close()
        D.close()
    gx.tune_set(gx.TUNE_CG_GRAPH, 1)
    return out


def bench_serve_sweep(n, rates, slots=4, mixed=2, fused_rates=(), fuse_rhs=16, wide_rates=(), wide_slots=8):
    """Leg 'sweep': the same Poisson trace shape at several arrival rates, so capacity and latency
    are reported together (p50/p99 grow once the rate passes the GPU's service rate); at
    `fused_rates` also with fused solves (waiting requests of one lattice and gauge field join an
    admitted request's solve, up to fuse_rhs right-hand sides, a mass per slice); at `wide_rates`
    also with `wide_slots` slots (more solves in flight on the one GPU)."""
    out = {f'{r:g}_per_s': bench_serving(n, r, slots, mixed) for r in rates}
    for r in fused_rates:
        out[f'{r:g}_per_s_fused{fuse_rhs}'] = bench_serving(n, r, slots, mixed, fuse_rhs=fuse_rhs)
    for r in wide_rates:
        out[f'{r:g}_per_s_{wide_slots}slots'] = bench_serving(n, r, wide_slots, mixed)
    return out


def bench_serving(n, rate, slots=4, mixed=1, fuse_rhs=0):
    """Poisson trace of n Wilson solve requests at `rate` per second through the serving runtime
    (mixed = 1: FP32 inner solves, 2: FP16-storage inner solves; fuse_rhs > 0: fused solves)."""
    from grid_amd import serving
    trace = serving.make_trace(n, rate, seed=5000)
    res, info = serving.serve(trace, slots=slots, budget=128e9, n_gauge=8, gauge_eps=0.3, tol=CG_TOL, mixed=mixed,
                              progress=max(64, n // 16), fuse_rhs=fuse_rhs)
    s = serving.summarize(res, info)
    s['meets_target'] = bool(s['max_residual'] <= TARGET_RESIDUAL)
    s.update({'arrival_rate_per_s': rate, 'slots': slots, 'admission_budget_bytes': 128e9, 'fuse_rhs': fuse_rhs,
              'inner_storage': {1: 'fp32', 2: 'fp16'}[mixed],
              'mix': 'volumes 8^4,12^4,16^4,16^3x32,24^4; 1-12 Gaussian sources; m in [0.05,0.3]; mixed CG to 1e-11 (|b - M x|/|b| <= 1e-10)'})
    return {k: (float(f'{v:.6g}') if isinstance(v, float) else v) for k, v in s.items()}


T_START = time.perf_counter()


def leg(name, fn):
    """Run one leg, with a progress line on stderr before and after it."""
    print(f'bench: {name} ... ({time.perf_counter() - T_START:.0f} s)', file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    r = fn()
    print(f'bench: {name} done in {time.perf_counter() - t0:.1f} s', file=sys.stderr, flush=True)
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=100)
    ap.add_argument('--warmup', type=int, default=10)
    ap.add_argument('--no-cg', action='store_true')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-c3', action='store_true')
    ap.add_argument('--no-c4', action='store_true')
    ap.add_argument('--serve-requests', type=int, default=256)
    ap.add_argument('--serve-rate', type=float, default=20.0)
    ap.add_argument('--serve-slots', type=int, default=4)
    ap.add_argument('--serve-mixed', type=int, default=2, choices=(1, 2),
                    help='inner solves of the served requests: 1 FP32, 2 FP16 storage')
    ap.add_argument('--sweep-requests', type=int, default=384)
    ap.add_argument('--sweep-rates', default='20,30,40,50,60')
    ap.add_argument('--sweep-fused-rates', default='40,50',
                    help='rates at which the sweep also serves with fused solves (16 right-hand sides)')
    ap.add_argument('--sweep-wide-rates', default='60',
                    help='rates at which the sweep also serves with --sweep-wide-slots slots')
    ap.add_argument('--sweep-wide-slots', type=int, default=8)
    ap.add_argument('--legs', default='dhop,large,cg,c3,c4,serve,sweep,graphs,cpu,ms,block',
                    help='comma list of legs to run: dhop, large (DhopEO at >= 1 GB working sets), cg, c3, c4, '
                         'serve, sweep (serving at several rates), graphs (hipGraph against eager CG), cpu, '
                         'ms (multishift CG), block (BlockCGrQ); profiling runs pick one')
    ap.add_argument('--no-graphs', action='store_true',
                    help='solver iterations as plain launches (for profiler runs; same kernels and timings)')
    ap.add_argument('--no-tune', action='store_true',
                    help='keep the library default launch knobs instead of autotuning each workload')
    ap.add_argument('--knobs', default='',
                    help='explicit hop knobs map,pf,nt,order[,shape,persist] for every precision (implies '
                         '--no-tune), e.g. 0,1,1,0; or per storage precision, e.g. '
                         'f64=4,1,0,0,6,1/f32=4,0,1,0,0,0/f16=4,0,1,1,0,0')
    args = ap.parse_args()
    legs = set(args.legs.split(','))
    args.no_cg = args.no_cg or 'cg' not in legs
    args.no_cpu = args.no_cpu or 'cpu' not in legs
    args.no_c3 = args.no_c3 or 'c3' not in legs
    args.no_c4 = args.no_c4 or 'c4' not in legs
    if 'serve' not in legs:
        args.serve_requests = 0
    if args.gpus != 1:
        print(f'bench.py measures one GPU; --gpus {args.gpus} is not supported (the multi-GPU halo exchange '
              f'is not built). Run with --gpus 1.', file=sys.stderr)
        sys.exit(2)

    import grid_amd as gx
    ctx = gx.Context(0)
    if args.no_graphs:
        gx.tune_set(gx.TUNE_CG_GRAPH, 0)
    from grid_amd import autotune
    if args.knobs:
        args.no_tune = True
        order = (gx.TUNE_DHOP_MAP, gx.TUNE_DHOP_PF, gx.TUNE_DHOP_NT, gx.TUNE_DHOP_ORDER, gx.TUNE_DHOP_SHAPE,
                 gx.TUNE_DHOP_PERSIST)
        precs = {'f64': gx.F64, 'f32': gx.F32, 'f16': gx.F16}
        for part in args.knobs.split('/'):
            name, _, vals = part.rpartition('=')
            if name and name not in precs:
                ap.error(f'--knobs: unknown precision {name!r} (f64, f32, f16)')
            for k, v in zip(order, (int(x) for x in vals.split(','))):
                gx.tune_set(k, v, precs[name] if name else None)
    tuner = None if args.no_tune else autotune.Autotuner(ctx, cache_path=None)
    tuned = {}

    def tune(key, op, prec, **kw):
        if tuner is not None:
            vals = tuner.tune(op, prec, **kw)
            tuned[key] = {autotune.KNOB_NAMES[k]: v for k, v in vals.items()}
        if args.no_graphs:
            gx.tune_set(gx.TUNE_CG_GRAPH, 0)

    U = gx.random_gauge(DIMS, 1002, EPS)
    psi_o = pick(gx.gaussian_fermion(DIMS, LS, 2002), DIMS, 1)
    D = gx.FermionOp(ctx, U, DIMS, LS, 'mobius', **PARAMS)
    H = np.prod(DIMS) // 2
    res = {}
    kernel64, shape64 = None, 0
    for prec, name, c in ((gx.F64, 'f64', 16), (gx.F32, 'f32', 8), (gx.F16, 'f16', 4)):
        print(f'bench: dhop {name} ... ({time.perf_counter() - T_START:.0f} s)', file=sys.stderr, flush=True)
        tune('dhop_' + name, D, prec, dhop_iters=50, rounds=5)
        if prec == gx.F64:
            kernel64, shape64 = dhop_kernel_name(gx), gx.tune_get(gx.TUNE_DHOP_SHAPE)
        fi = D.field(prec, 1).upload(psi_o)
        fo = D.field(prec, 0)
        ms = time_dhop(ctx, D, fi, fo, args.steps, args.warmup)
        # FP16 adds a 4-byte scale per 5D site to the input read and the output written (links
        # are unscaled)
        byts = H * (24 * LS + 72) * c + (H * LS * 8 if prec == gx.F16 else 0)
        res[name] = dict(ms=ms, bytes=int(byts), gbs=byts / (ms * 1e-3) / 1e9, gflops=1320.0 * H * LS / (ms * 1e-3) / 1e9)
        if prec == gx.F64:
            # after the timed region, one streaming pass of known size (fo = 1.0 * fi: the stored
            # half field read once and written once, k_axpby): the calibration point for the HBM
            # byte counters of a profiled run
            fo.set_checkerboard_tag(1)
            gx.lib().gx_scale(fo.h, 1.0, fi.h)
            ctx.sync()
        fi.close()
        fo.close()

    cg = None
    if not args.no_cg:
        print(f'bench: cg ... ({time.perf_counter() - T_START:.0f} s)', file=sys.stderr, flush=True)
        tune('cg', D, gx.F32)
        src_full = gx.gaussian_fermion(DIMS, LS, 3003)
        fs, fsol = D.field().upload(src_full), D.field()
        st64 = D.schur_solve(fs, fsol, tol=CG_TOL_MOBIUS, maxit=20000, mixed=False)
        x64 = fsol.download()
        fsol.zero()
        stmx = D.schur_solve(fs, fsol, tol=CG_TOL_MOBIUS, maxit=20000, mixed=True)
        xmx = fsol.download()
        tune('cg_f16', D, gx.F16)
        fsol.zero()
        st16 = D.schur_solve(fs, fsol, tol=CG_TOL_MOBIUS, maxit=20000, mixed=2)
        x16 = fsol.download()
        cg = dict(cg_tol=CG_TOL_MOBIUS,
                  fp64=dict(seconds=st64['seconds'], iterations=st64['iterations'], true_residual=st64['true_residual'],
                            unprec_residual=st64['unprec_residual'],
                            meets_target=bool(st64['unprec_residual'] <= TARGET_RESIDUAL)),
                  mixed_f32=dict(seconds=stmx['seconds'], inner_iterations=stmx['inner_iterations'],
                                 outer=stmx['outer_iterations'], final_iterations=stmx['final_iterations'],
                                 true_residual=stmx['true_residual'], unprec_residual=stmx['unprec_residual'],
                                 meets_target=bool(stmx['unprec_residual'] <= TARGET_RESIDUAL),
                                 rel_err_vs_fp64=float(np.linalg.norm(xmx - x64) / np.linalg.norm(x64))),
                  mixed_f16=dict(seconds=st16['seconds'], inner_iterations=st16['inner_iterations'],
                                 outer=st16['outer_iterations'], final_iterations=st16['final_iterations'],
                                 true_residual=st16['true_residual'], unprec_residual=st16['unprec_residual'],
                                 meets_target=bool(st16['unprec_residual'] <= TARGET_RESIDUAL),
                                 rel_err_vs_fp64=float(np.linalg.norm(x16 - x64) / np.linalg.norm(x64))),
                  speedup_mixed_vs_fp64=st64['seconds'] / stmx['seconds'],
                  speedup_mixed_f16_vs_fp64=st64['seconds'] / st16['seconds'])

    c3 = None if args.no_c3 else leg('c3', lambda: bench_mixed_dwf(gx, ctx, tune=tune))
    c4 = None if args.no_c4 else leg('c4', lambda: bench_multi_rhs(gx, ctx, tune=tune))
    serve = (leg('serve', lambda: bench_serving(args.serve_requests, args.serve_rate, args.serve_slots, args.serve_mixed))
             if args.serve_requests > 0 else None)
    sweep = (leg('sweep', lambda: bench_serve_sweep(args.sweep_requests, [float(r) for r in args.sweep_rates.split(',')],
                                                    args.serve_slots, args.serve_mixed,
                                                    [float(r) for r in args.sweep_fused_rates.split(',') if r],
                                                    wide_rates=[float(r) for r in args.sweep_wide_rates.split(',') if r],
                                                    wide_slots=args.sweep_wide_slots))
             if 'sweep' in legs and args.sweep_requests > 0 else None)
    graphs = leg('graphs', lambda: bench_graphs(gx, ctx)) if 'graphs' in legs else None
    ms = leg('ms', lambda: bench_multishift(gx, ctx)) if 'ms' in legs else None
    blk = leg('block', lambda: bench_block_cgrq(gx, ctx)) if 'block' in legs else None

    large = leg('large', lambda: bench_dhop_large(gx, ctx, tune)) if 'large' in legs else None

    cpu = None
    if not args.no_cpu:
        c2_its = cg['fp64']['iterations'] if cg else 463
        cb, byts = leg('cpu', lambda: cpu_baseline(gx, U, psi_o, c2_its))
        a = cb['all_cores']
        cpu = {'value': a['gbs'], 'unit': 'GB/s', 'cores': a['threads'], 'kind': 'port',
               'sample': f'C2 DhopEO (Mobius 16^4xLs16 FP64, {byts} B per call), median of 20 calls after warm-up, '
                         f'oracle/cpu_ref.cpp (the CPU restatement of Grid\'s algorithm; Grid itself cannot be built '
                         f'here) with {a["threads"]} OpenMP threads ({cb["host_physical_cores"]} physical cores on '
                         f'the host, capped at the job\'s CPU share), and with 1 thread; red-black solves of C1 and C4 '
                         f'(one point source) at the GPU legs\' tolerances with all threads, and the first '
                         f'{C2_CPU_ITERS} iterations of the C2 solve projected to its iteration count',
               **cb}

    r = res['f64']
    traffic = traffic_from_profile(kernel64, shape64)
    out = {
        'metric': METRIC,
        'value': round(r['gbs'], 2),
        'unit': 'GB/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(r['ms'], 5),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic (random SU(3) exp(i 0.3 H) links, Gaussian source)',
        'config': {'workload': 'Mobius DWF DhopEO odd->even', 'dims': list(DIMS), 'Ls': LS, **PARAMS,
                   'parallelism': 'single GPU'},
        'roofline': {'bound': 'hbm', 'achieved': round(r['gbs'], 2), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': round(r['gbs'] / HBM_PEAK_GBS, 4), 'traffic': traffic[0],
                     'traffic_source': (f'{traffic[1]} (rocprofv3 FETCH_SIZE and WRITE_SIZE runs of bench.py --legs '
                                        'dhop with this kernel, scaled by a streaming pass)') if traffic[1] else
                                       f'no committed counter summary for {kernel64}',
                     'kernel': kernel64, 'algorithmic_bytes_per_launch': r['bytes']},
        'cpu_baseline': cpu,
        'dhop_hbm_sized': large,
        'pct_hbm_peak': round(100 * r['gbs'] / HBM_PEAK_GBS, 2),
        'gflops_1320': round(r['gflops'], 1),
        'f32': {k: (round(v, 5) if isinstance(v, float) else v) for k, v in res['f32'].items()},
        'f16': {k: (round(v, 5) if isinstance(v, float) else v) for k, v in res['f16'].items()},
        'tuned_knobs': tuned,
        'cg': cg,
        'c3_mixed_dwf': c3,
        'c4_multi_rhs': c4,
        'multishift': ms,
        'block_cgrq': blk,
        'serving': serve,
        'serving_rate_sweep': sweep,
        'cg_graph_vs_eager': graphs,
    }
    print(json.dumps(out))


if __name__ == '__main__':
    main()

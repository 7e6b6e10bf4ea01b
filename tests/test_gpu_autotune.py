"""The autotuner on the GPU: it times every hop and solver knob setting on the caller's workload,
caches the winners, and the tuned operator computes what the default one computes."""
import numpy as np
import pytest

import grid_amd as gx
from grid_amd import autotune as at
from conftest import max_site_rel_err

pytestmark = pytest.mark.gpu


def test_autotune_mobius_small(tmp_path):
    ctx = gx.Context(0)
    dims, Ls = (8, 8, 8, 8), 8
    U = gx.random_gauge(dims, 5, 0.3)
    D = gx.FermionOp(ctx, U, dims, Ls, 'mobius', mass=0.05)
    psi = gx.gaussian_fermion(dims, Ls, 6)
    lt, lz, ly, lx = dims[::-1]
    import npref as R
    po = R.pick(psi, dims, gx.ODD)
    saved = {k: gx.tune_get(k) for k in at.KNOB_NAMES}
    try:
        fi, fo = D.field(gx.F64, gx.ODD).upload(po), D.field(gx.F64, gx.EVEN)
        D.DhopEO(fi, fo, 0)
        ref = fo.download()
        tuner = at.Autotuner(ctx, cache_path=str(tmp_path / 'tune.json'))
        vals = tuner.tune(D, gx.F64, dhop_iters=5, cg_iters=16, rounds=1)
        assert set(vals) == set(at.KNOB_NAMES)
        assert all(gx.tune_get(k) == v for k, v in vals.items())
        D.DhopEO(fi, fo, 0)
        assert max_site_rel_err(fo.download(), ref) < 1e-14
        again = at.Autotuner(ctx, cache_path=str(tmp_path / 'tune.json'))
        assert again.apply('mobius', dims, Ls, gx.F64) == vals
    finally:
        for k, v in saved.items():
            gx.tune_set(k, v)
        ctx.close()


def test_autotune_sloppy_precision(tmp_path):
    """Tuning an FP32 operator sets the FP32 knob set only (the FP64 one is left alone), and an FP32
    mixed solve under the tuned knobs reaches the FP64 tolerance."""
    ctx = gx.Context(0)
    dims, Ls = (8, 8, 8, 8), 8
    U = gx.random_gauge(dims, 7, 0.3)
    D = gx.FermionOp(ctx, U, dims, Ls, 'mobius', mass=0.1)
    import npref as R
    po = R.pick(gx.gaussian_fermion(dims, Ls, 8), dims, gx.ODD)
    saved = {p: {k: gx.tune_get(k, p) for k in at.KNOB_NAMES} for p in (gx.F64, gx.F32)}
    try:
        tuner = at.Autotuner(ctx, cache_path=str(tmp_path / 'tune.json'))
        vals = tuner.tune(D, gx.F32, dhop_iters=5, cg_iters=16, rounds=1)
        assert all(gx.tune_get(k, gx.F32) == v for k, v in vals.items())
        assert all(gx.tune_get(k, gx.F64) == saved[gx.F64][k] for k in at.KNOB_NAMES)
        fs, fx = D.field(gx.F64, gx.ODD).upload(po), D.field(gx.F64, gx.ODD)
        st = D.mixed_cg(fs, fx, tol=1e-10, inner_tol=1e-5, inner_prec=gx.F32)
        assert st['converged'] and st['true_residual'] < 1e-9, st
    finally:
        for p, kv in saved.items():
            for k, v in kv.items():
                gx.tune_set(k, v, p)
        ctx.close()

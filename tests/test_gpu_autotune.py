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

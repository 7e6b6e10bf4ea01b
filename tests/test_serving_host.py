"""CPU checks of the serving host logic: request traces, the admission footprint, latency
summaries and the runtime's argument validation (no device work)."""
import numpy as np
import pytest

import grid_amd as gx
from grid_amd import serving


def test_trace_is_poisson_and_uniform():
    n, rate = 20000, 50.0
    t = serving.make_trace(n, rate, seed=3)
    arr = np.array([r['arrival'] for r in t])
    assert arr[0] == 0.0 and np.all(np.diff(arr) >= 0)
    gaps = np.diff(arr)
    assert abs(gaps.mean() * rate - 1) < 0.03 and abs(gaps.std() * rate - 1) < 0.05  # exponential
    vols = [r['dims'] for r in t]
    for v in serving.VOLUMES:
        assert abs(vols.count(v) / n - 1 / len(serving.VOLUMES)) < 0.02
    ns = np.array([r['nsrc'] for r in t])
    assert ns.min() == 1 and ns.max() == 12 and abs(ns.mean() - 6.5) < 0.1
    ms = np.array([r['mass'] for r in t])
    assert 0.05 <= ms.min() and ms.max() <= 0.3 and abs(ms.mean() - 0.175) < 0.005
    assert len({r['seed'] for r in t}) == n
    assert t == serving.make_trace(n, rate, seed=3)
    assert t != serving.make_trace(n, rate, seed=4)


def test_footprint_closed_form():
    dims, nsrc = (16, 16, 16, 32), 7
    hb = np.prod(dims) // 2 * nsrc * 12 * 16
    coef = 48 * nsrc  # the operator's per-slice Schur coefficients: 4 doubles and 4 floats per slice
    # the fused hop epilogues' slice sums: one double per slice per tile, at most one tile per 64 // nsrc sites
    spb = 64 // nsrc
    tiles = 8 * nsrc * (-(-np.prod(dims) // 2 // spb))
    assert serving.footprint(dims, nsrc, mixed=True) == 24.5 * hb + coef + tiles
    assert serving.footprint(dims, nsrc, mixed=False) == 18.0 * hb + coef + tiles
    assert serving.footprint(dims, 1, mixed=False) == 18.0 * hb / nsrc + 48


def test_summary_percentiles():
    res = [dict(status=0, latency=float(i), nsrc=2, residual=1e-11) for i in range(1, 101)]
    res.append(dict(status=3, latency=1e9, nsrc=5, residual=1.0))
    s = serving.summarize(res, {'wall': 10.0, 'resident_bytes': 0.0})
    assert s['completed'] == 100 and s['failed'] == 1
    assert s['p50_latency_s'] == pytest.approx(50.5) and s['p99_latency_s'] == pytest.approx(99.01)
    assert s['solves_per_s'] == 10.0 and s['rhs_per_s'] == 20.0 and s['max_residual'] == 1e-11


@pytest.mark.parametrize('bad', [dict(slots=0), dict(n_gauge=0), dict(budget=0.0)])
def test_serve_rejects_bad_parameters(bad):
    reqs = [dict(dims=(4, 4, 4, 4), nsrc=1, mass=0.1, gauge=0, seed=1, arrival=0.0)]
    kw = dict(slots=1, n_gauge=1, budget=1e9)
    kw.update(bad)
    with pytest.raises(gx.GxError) as e:
        serving.serve(reqs, **kw)
    assert e.value.code == 1


@pytest.mark.parametrize('req', [dict(dims=(5, 4, 4, 4)), dict(nsrc=0), dict(nsrc=17), dict(gauge=3)])
def test_serve_rejects_bad_requests(req):
    r = dict(dims=(4, 4, 4, 4), nsrc=1, mass=0.1, gauge=0, seed=1, arrival=0.0)
    r.update(req)
    with pytest.raises(gx.GxError) as e:
        serving.serve([r], slots=1, n_gauge=2, budget=1e9)
    assert e.value.code == 1

"""Solve serving on one GPU: request traces, the gx_serve runtime, and latency summaries.

A request asks for Wilson solves M x_k = b_k (k < nsrc) on one of the resident gauge fields of
its lattice, with Gaussian sources.  ``serve`` hands a trace to the native runtime (gx_serve in
libgx): FIFO admission by idle slot and device-memory footprint, one stream and host thread per
slot, so solves overlap on the device while requests keep arriving.  No Grid counterpart.
"""
import ctypes
import os

import numpy as np

from . import Request, Response, ServeParams, _DP, _check, _dims, lib

# the request mix: lattice volumes (x, y, z, t), masses, sources per request
VOLUMES = ((8, 8, 8, 8), (12, 12, 12, 12), (16, 16, 16, 16), (16, 16, 16, 32), (24, 24, 24, 24))
MASS_RANGE = (0.05, 0.3)
NSRC_RANGE = (1, 12)


def make_trace(n, rate, seed=5000, volumes=VOLUMES, mass_range=MASS_RANGE, nsrc_range=NSRC_RANGE, n_gauge=8):
    """n requests with Poisson arrivals at `rate` per second (exponential gaps from t = 0), volume
    uniform over `volumes`, mass uniform over `mass_range`, nsrc uniform over the integers of
    `nsrc_range`, gauge field uniform over n_gauge, and a distinct source seed per request."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=n) if rate > 0 else np.zeros(n)
    arrival = np.cumsum(gaps) - gaps[0] if n else gaps
    reqs = []
    for i in range(n):
        reqs.append({'dims': tuple(volumes[rng.integers(len(volumes))]),
                     'nsrc': int(rng.integers(nsrc_range[0], nsrc_range[1] + 1)),
                     'mass': float(rng.uniform(*mass_range)),
                     'gauge': int(rng.integers(n_gauge)),
                     'seed': int(seed * 1_000_003 + i),
                     'arrival': float(arrival[i])})
    return reqs


def footprint(dims, nsrc, mixed=True):
    """Device bytes a request holds while it runs (fields and operator temporaries).  mixed: 0 or
    False (FP64 CG), 1 or True (FP32 inner solves), 2 (FP16-storage inner solves)."""
    v = ctypes.c_double()
    _check(lib().gx_serve_footprint(_dims(dims), int(nsrc), int(mixed), ctypes.byref(v)))
    return v.value


def serve(reqs, slots=4, budget=64e9, n_gauge=8, gauge_seed=1000, gauge_eps=0.3, tol=1e-10, mixed=True,
          maxit=20000, realtime=True, device=0, want_solutions=False, progress=0, gauge_dir=None, fuse_rhs=0):
    """Run a trace through gx_serve.  Returns (responses, info): one dict per request (status,
    slot, iterations, residual, t_admit, t_done, footprint, digest, latency) and info with the
    wall time and resident link bytes.  With want_solutions the FP64 solutions come back as host
    arrays (Lt, Lz, Ly, Lx, nsrc, 4, 3) in responses[i]['solution'].  progress = N > 0 makes the
    runtime print a line on stderr after every N completed requests.  fuse_rhs = N > 0 lets waiting
    requests of the same lattice and gauge field join an admitted request's solve (at most N
    right-hand sides, a mass per slice); responses[i]['batch'] counts the requests of its solve."""
    n = len(reqs)
    arr = (Request * max(n, 1))()
    for i, r in enumerate(reqs):
        arr[i].dims[:] = [int(d) for d in r['dims']]
        arr[i].nsrc, arr[i].mass, arr[i].gauge = int(r['nsrc']), float(r['mass']), int(r['gauge'])
        arr[i].seed, arr[i].arrival = int(r['seed']), float(r['arrival'])
    out = (Response * max(n, 1))()
    p = ServeParams(device=device, slots=slots, budget=float(budget), n_gauge=n_gauge, gauge_seed=gauge_seed,
                    gauge_eps=gauge_eps, tol=tol, mixed=int(mixed), maxit=maxit, realtime=int(bool(realtime)),
                    progress=int(progress), gauge_dir=None if gauge_dir is None else os.fsencode(gauge_dir),
                    fuse_rhs=int(fuse_rhs))
    sols, ptrs = None, None
    if want_solutions:
        sols = []
        ptrs = (_DP * max(n, 1))()
        for i, r in enumerate(reqs):
            lx, ly, lz, lt = r['dims']
            a = np.zeros((lt, lz, ly, lx, int(r['nsrc']), 4, 3), dtype=np.complex128)
            sols.append(a)
            ptrs[i] = a.ctypes.data_as(_DP)
    wall, resident = ctypes.c_double(), ctypes.c_double()
    rc = lib().gx_serve(ctypes.byref(p), arr, n, out, ptrs, ctypes.byref(wall), ctypes.byref(resident))
    _check(rc)
    res = []
    for i, r in enumerate(reqs):
        d = out[i].as_dict()
        d['arrival'], d['dims'], d['nsrc'] = r['arrival'], tuple(r['dims']), int(r['nsrc'])
        d['latency'] = d['t_done'] - r['arrival'] if realtime else d['t_done']
        if sols is not None:
            d['solution'] = sols[i]
        res.append(d)
    return res, {'wall': wall.value, 'resident_bytes': resident.value}


def summarize(res, info):
    """Throughput and latency percentiles of a served trace (failed requests excluded from the
    percentiles and counted separately)."""
    ok = [r for r in res if r['status'] in (0,)]
    lat = np.array([r['latency'] for r in ok]) if ok else np.zeros(1)
    nsrc = sum(r['nsrc'] for r in ok)
    wall = info['wall']
    return {'requests': len(res), 'completed': len(ok), 'failed': len(res) - len(ok),
            'wall_s': wall, 'solves_per_s': len(ok) / wall if wall > 0 else 0.0,
            'rhs_per_s': nsrc / wall if wall > 0 else 0.0,
            'p50_latency_s': float(np.percentile(lat, 50)), 'p99_latency_s': float(np.percentile(lat, 99)),
            'max_residual': max((r['residual'] for r in ok), default=0.0),
            'fused_requests': sum(1 for r in ok if r.get('batch', 1) > 1),
            'resident_link_bytes': info['resident_bytes']}

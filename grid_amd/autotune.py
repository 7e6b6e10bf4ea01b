"""Launch-shape autotuner for the hop stencil and the solver loop.

The knobs are libgx's process-wide launch choices (gx_tune_set):
  GX_TUNE_DHOP_MAP    workgroup -> lattice tile map (linear, or one contiguous slab per XCD)
  GX_TUNE_DHOP_PF     software prefetch of the next leg's neighbours
  GX_TUNE_DHOP_NT     non-temporal output stores
  GX_TUNE_DHOP_ORDER  leg order (Grid's, or the two legs of a direction adjacent; Ls in {1,12,16})
  GX_TUNE_DHOP_SHAPE  tile shape: whole sites per wave in 4-wave workgroups, or W waves of 64
                      consecutive (site, s) entries each (sites per workgroup = 64 W / Ls)
  GX_TUNE_DHOP_PERSIST  the W-wave hop as a persistent grid of resident workgroups over the tiles
  GX_TUNE_DHOP_TILE   tiles of 2 x 2 rows of x with the x, y and z legs' neighbours staged in LDS
  GX_TUNE_FUSED       fused Schur operator (hop epilogues) or the unfused composition
  GX_TUNE_CG_GRAPH    solver iteration batches replayed as hipGraphs
Every knob holds one value per storage precision; tune(D, prec) measures and sets the values of
precision prec only, so the FP64 and inner-precision operators of a mixed solve keep their own.
The hop knobs are chosen by timing DhopEO on the caller's own lattice, Ls and precision (HIP
events over repeated launches, best of several rounds), by coordinate search: map, prefetch,
stores and leg order in the default tile shape, then the tile shape (W-wave shapes with and
without the persistent grid, and the LDS row tiles), then the map again in the chosen shape.  For
FP32 and FP16 storage, whose CG folds its reductions into hop epilogues, a persistent winner and
the fastest hop of the other prefetch setting are then compared by CG iteration time.  The
solver knobs are chosen by timing a fixed number of CG iterations.  Results are cached in a JSON file keyed
by device and workload, so a later
process applies them without measuring.  No Grid counterpart: Grid fixes its launch shape when it
is compiled (Grid/threads/Accelerator.h:416-461).
"""
import itertools
import json
import math
import os
import time

import numpy as np

from . import (EVEN, F16, F32, F64, ODD, TUNE_CG_GRAPH, TUNE_DHOP_MAP, TUNE_DHOP_NT, TUNE_DHOP_ORDER,
               TUNE_DHOP_PERSIST, TUNE_DHOP_PF, TUNE_DHOP_SHAPE, TUNE_DHOP_TILE, TUNE_FUSED, tune_get, tune_set)

HOP_KNOBS = (TUNE_DHOP_MAP, TUNE_DHOP_PF, TUNE_DHOP_NT, TUNE_DHOP_ORDER)
SOLVER_KNOBS = (TUNE_FUSED, TUNE_CG_GRAPH)
KNOB_NAMES = {TUNE_DHOP_MAP: 'dhop_map', TUNE_DHOP_PF: 'dhop_pf', TUNE_DHOP_NT: 'dhop_nt',
              TUNE_DHOP_ORDER: 'dhop_order', TUNE_FUSED: 'fused', TUNE_CG_GRAPH: 'cg_graph',
              TUNE_DHOP_SHAPE: 'dhop_shape', TUNE_DHOP_PERSIST: 'dhop_persist', TUNE_DHOP_TILE: 'dhop_tile'}
LAUNCH_KNOBS = (TUNE_DHOP_SHAPE, TUNE_DHOP_PERSIST, TUNE_DHOP_TILE)
COMPLEX_BYTES = {F64: 16, F32: 8, F16: 8}   # arithmetic complex of a precision (FP16 storage computes in FP32)
PREC_NAMES = {F64: 'f64', F32: 'f32', F16: 'f16'}
DEFAULT_CACHE = os.path.join(os.path.expanduser('~'), '.cache', 'grid_amd', 'tune.json')


def workload_key(device, kind, dims, Ls, prec):
    return f'{device}|{kind}|{"x".join(str(d) for d in dims)}|Ls{Ls}|{PREC_NAMES[prec]}'


def search(candidates, measure, rounds=3, margin=0.0):
    """Best candidate by the minimum over `rounds` of measure(candidate) (seconds); candidates are
    measured round-robin so drift affects them alike.  With margin > 0 the first candidate (the
    default) is kept unless another is faster by more than that fraction.  Returns
    (best, {candidate: time})."""
    best_t = {c: float('inf') for c in candidates}
    for _ in range(rounds):
        for c in candidates:
            best_t[c] = min(best_t[c], measure(c))
    best = min(candidates, key=lambda c: (best_t[c], candidates.index(c)))
    if margin > 0 and best_t[best] >= (1.0 - margin) * best_t[candidates[0]]:
        best = candidates[0]
    return best, best_t


def hop_candidates(Ls):
    """Hop knob settings to try, each a tuple of (knob, value) pairs."""
    orders = (0, 1) if Ls in (1, 12, 16) else (0,)
    return [tuple(zip(HOP_KNOBS, v)) for v in itertools.product((0, 1, 2, 3, 4), (0, 1), (0, 1), orders)]


def shape_candidates(Ls, prec=F64):
    """Tile shapes to try for extent Ls: 0 (whole sites per wave, 4 waves; Ls <= 16) and W waves for
    W = w0, 2 w0, 4 w0, ... <= 8 (w0 = Ls / gcd(Ls, 64), the library's rule for whole-site
    workgroups) whose link tile of 64 W / Ls sites fits 64 KB of LDS (Ls >= 8)."""
    out = [0] if Ls <= 16 else []
    if Ls >= 8:
        w = Ls // math.gcd(Ls, 64)
        while w <= 8:
            if (64 * w // Ls) * 73 * COMPLEX_BYTES[prec] <= 65536:
                out.append(w)
            w *= 2
    return out


def launch_candidates(shapes):
    """Launch forms: each tile shape (W > 0 with and without the persistent grid), and the LDS row
    tiles (the library falls back to the W-wave or templated form where a row tile does not fit)."""
    out = [((TUNE_DHOP_SHAPE, w), (TUNE_DHOP_PERSIST, p), (TUNE_DHOP_TILE, 0))
           for w in shapes for p in ((0, 1) if w else (0,))]
    out.append(((TUNE_DHOP_SHAPE, shapes[0]), (TUNE_DHOP_PERSIST, 0), (TUNE_DHOP_TILE, 1)))
    return out


def solver_candidates():
    return [tuple(zip(SOLVER_KNOBS, v)) for v in itertools.product((1, 0), (1, 0))]


PRECS = (F64, F32, F16)


class Knobs:
    """Set knobs (for every precision, or for `prec` only) for the duration of a with-block and
    restore every precision's previous values."""

    def __init__(self, values, prec=None):
        self.values = dict(values)
        self.prec = prec

    def __enter__(self):
        self.saved = {(k, p): tune_get(k, p) for k in self.values for p in PRECS}
        for k, v in self.values.items():
            tune_set(k, v, self.prec)
        return self

    def __exit__(self, *exc):
        for (k, p), v in self.saved.items():
            tune_set(k, v, p)


def _device_name():
    try:
        import torch
        if torch.cuda.is_available():
            return torch.cuda.get_device_properties(0).gcnArchName.split(':')[0]
    except Exception:
        pass
    return 'gfx950'


class Autotuner:
    def __init__(self, ctx, cache_path=DEFAULT_CACHE, device=None):
        self.ctx = ctx
        self.cache_path = cache_path
        self.device = device or _device_name()
        self.cache = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as f:
                self.cache = json.load(f)

    def save(self):
        if not self.cache_path:
            return
        os.makedirs(os.path.dirname(self.cache_path), exist_ok=True)
        tmp = self.cache_path + '.tmp'
        with open(tmp, 'w') as f:
            json.dump(self.cache, f, indent=1, sort_keys=True)
        os.replace(tmp, self.cache_path)

    def _time_dhop(self, D, fi, fo, iters):
        D.DhopEO(fi, fo, 0)
        self.ctx.sync()
        self.ctx.record(0)
        for _ in range(iters):
            D.DhopEO(fi, fo, 0)
        self.ctx.record(1)
        self.ctx.sync()
        return self.ctx.elapsed_ms(0, 1) * 1e-3 / iters

    def _time_cg(self, D, fs, fx, iters):
        fx.zero()
        self.ctx.sync()
        t0 = time.perf_counter()
        D.cg(fs, fx, tol=1e-30, maxit=iters)  # runs exactly maxit iterations
        self.ctx.sync()
        return (time.perf_counter() - t0) / iters

    def tune(self, D, prec=F64, dhop_iters=20, cg_iters=64, rounds=3, force=False):
        """Tune for operator D at precision prec; returns the chosen knob values (and applies them)."""
        key = workload_key(self.device, D.kind, D.dims, D.Ls, prec)
        if key in self.cache and not force:
            return self.apply_entry(self.cache[key])
        lt, lz, ly, lx = D.dims[3], D.dims[2], D.dims[1], D.dims[0]
        rng = np.random.default_rng(1)
        shape = (lt, lz, ly, lx // 2, D.Ls, 4, 3)
        psi = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        fi, fo = D.field(prec, ODD).upload(psi), D.field(prec, EVEN)

        def hop(c):
            with Knobs(c, prec):
                return self._time_dhop(D, fi, fo, dhop_iters)

        shapes = shape_candidates(D.Ls, prec)
        default_shape = ((TUNE_DHOP_SHAPE, shapes[0]), (TUNE_DHOP_PERSIST, 0), (TUNE_DHOP_TILE, 0))
        first, hop_times = search([c + default_shape for c in hop_candidates(D.Ls)], hop, rounds)
        others = tuple(kv for kv in first if kv[0] not in LAUNCH_KNOBS)
        launch = launch_candidates(shapes)
        if len(launch) > 1:
            # the first launch form (for Ls <= 16 the templated 4-wave kernel) unless another wins
            # by over 2%: the forms often run within a few percent, the size of the timing noise
            best_shape, t2 = search([others + c for c in launch], hop, rounds, margin=0.02)
            hop_times.update(t2)
            chosen = tuple(kv for kv in best_shape if kv[0] in LAUNCH_KNOBS)
            rest = tuple(kv for kv in others if kv[0] != TUNE_DHOP_MAP)
            hop_best, t3 = search([((TUNE_DHOP_MAP, m),) + rest + chosen for m in range(5)], hop, rounds)
            hop_times.update(t3)
        else:
            hop_best = first
        fs, fx = D.field(prec, ODD).upload(psi), D.field(prec, ODD)

        def cg(c):
            with Knobs(hop_best + c, prec):
                return self._time_cg(D, fs, fx, cg_iters)

        # FP32 / FP16 CG folds its reductions into hop epilogues that only the one-workgroup-per-tile
        # launches have: when the fastest hop is a persistent grid, keep it only if a CG iteration
        # under it beats one under the fastest non-persistent launch.  The epilogue kernels also
        # use registers differently from the plain hop (the FP16 ones fit fewer waves without
        # prefetch), so the fastest hop of the other prefetch setting is compared by CG time too.
        def cg_hop(h):
            with Knobs(h, prec):
                return self._time_cg(D, fs, fx, cg_iters)

        if prec != F64 and dict(hop_best).get(TUNE_DHOP_PERSIST, 0):
            plain = [c for c in hop_times if not dict(c).get(TUNE_DHOP_PERSIST, 0)]
            if plain:
                alt = min(plain, key=lambda c: hop_times[c])
                hop_best, t4 = search([hop_best, alt], cg_hop, rounds)
        if prec != F64:
            pf = dict(hop_best).get(TUNE_DHOP_PF, 0)
            other = [c for c in hop_times
                     if dict(c).get(TUNE_DHOP_PF, 0) != pf and not dict(c).get(TUNE_DHOP_PERSIST, 0)]
            if other:
                alt = min(other, key=lambda c: hop_times[c])
                hop_best, t5 = search([hop_best, alt], cg_hop, rounds)

        solver_best, solver_times = search(solver_candidates(), cg, rounds)
        entry = {'knobs': {KNOB_NAMES[k]: v for k, v in hop_best + solver_best},
                 'dhop_seconds': hop_times[hop_best],
                 'cg_iteration_seconds': solver_times[solver_best],
                 'measured': time.strftime('%Y-%m-%dT%H:%M:%S')}
        self.cache[key] = entry
        self.save()
        for f in (fi, fo, fs, fx):
            f.close()
        return self.apply_entry(entry, prec)

    @staticmethod
    def apply_entry(entry, prec=None):
        """Set a cached entry's knobs for storage precision prec (every precision when None)."""
        byname = {v: k for k, v in KNOB_NAMES.items()}
        vals = {byname[n]: int(v) for n, v in entry['knobs'].items()}
        for k, v in vals.items():
            tune_set(k, v, prec)
        return vals

    def apply(self, kind, dims, Ls, prec=F64):
        """Apply cached knobs for a workload (to the knob set of prec); returns them, or None when it
        was never tuned."""
        e = self.cache.get(workload_key(self.device, kind, tuple(dims), Ls, prec))
        return self.apply_entry(e, prec) if e else None

"""CPU checks of the autotuner's search and cache logic (no device work: the timer is injected)."""
import json

import grid_amd as gx
from grid_amd import autotune as at


def test_search_picks_minimum_over_rounds():
    calls = []

    def measure(c):
        calls.append(c)
        base = {'a': 3.0, 'b': 1.0, 'c': 2.0}[c]
        return base + (0.0 if len(calls) > 3 else 10.0)  # the first round is noisy

    best, times = at.search(['a', 'b', 'c'], measure, rounds=3)
    assert best == 'b' and times == {'a': 3.0, 'b': 1.0, 'c': 2.0}
    assert calls == ['a', 'b', 'c'] * 3  # round-robin


def test_search_breaks_ties_by_order():
    best, _ = at.search(['x', 'y'], lambda c: 1.0, rounds=2)
    assert best == 'x'


def test_search_margin_keeps_default():
    t = {'x': 1.0, 'y': 0.99, 'z': 0.97}
    assert at.search(['x', 'y'], t.get, rounds=1, margin=0.02)[0] == 'x'
    assert at.search(['x', 'z'], t.get, rounds=1, margin=0.02)[0] == 'z'
    assert at.search(['x', 'y'], t.get, rounds=1)[0] == 'y'


def test_candidates_cover_knob_space():
    c16 = at.hop_candidates(16)
    c8 = at.hop_candidates(8)
    assert len(c16) == 40 and len(c8) == 20 and len(set(c16)) == 40
    assert all(dict(c)[gx.TUNE_DHOP_ORDER] == 0 for c in c8)
    assert len(at.solver_candidates()) == 4


def test_cache_round_trip_and_apply(tmp_path):
    path = tmp_path / 'tune.json'
    key = at.workload_key('gfx950', 'mobius', (16, 16, 16, 16), 16, gx.F64)
    entry = {'knobs': {'dhop_map': 1, 'dhop_pf': 0, 'dhop_nt': 1, 'dhop_order': 0, 'fused': 1, 'cg_graph': 0}}
    path.write_text(json.dumps({key: entry}))
    saved = {k: gx.tune_get(k) for k in at.KNOB_NAMES}
    try:
        t = at.Autotuner(ctx=None, cache_path=str(path), device='gfx950')
        vals = t.apply('mobius', (16, 16, 16, 16), 16, gx.F64)
        assert vals[gx.TUNE_DHOP_MAP] == 1 and gx.tune_get(gx.TUNE_DHOP_PF) == 0
        assert gx.tune_get(gx.TUNE_CG_GRAPH) == 0
        assert t.apply('mobius', (8, 8, 8, 8), 16, gx.F64) is None
        t.cache['other'] = {'knobs': {}}
        t.save()
        assert set(json.loads(path.read_text())) == {key, 'other'}
    finally:
        for k, v in saved.items():
            gx.tune_set(k, v)


def test_knobs_context_restores():
    before = gx.tune_get(gx.TUNE_DHOP_NT)
    with at.Knobs(((gx.TUNE_DHOP_NT, 1 - before),)):
        assert gx.tune_get(gx.TUNE_DHOP_NT) == 1 - before
    assert gx.tune_get(gx.TUNE_DHOP_NT) == before


def test_shape_candidates_follow_the_library_rule():
    """Tile shapes: W a multiple of Ls / gcd(Ls, 64), W <= 8, link tile within 64 KB of LDS, and the
    sites-per-wave shape (0) only for Ls <= 16."""
    assert at.shape_candidates(4) == [0]                      # Ls < 8: templated kernel only
    assert at.shape_candidates(16, gx.F64) == [0, 1, 2, 4, 8]
    assert at.shape_candidates(8, gx.F64) == [0, 1, 2, 4]         # 8 waves: a 74 KB link tile
    assert at.shape_candidates(8, gx.F32) == [0, 1, 2, 4, 8]
    assert at.shape_candidates(12, gx.F32) == [0, 3, 6]
    assert at.shape_candidates(128) == [2, 4, 8]
    assert at.shape_candidates(24, gx.F64) == [3, 6]
    assert at.shape_candidates(20) == [5]
    assert at.shape_candidates(96) == [3, 6]
    assert at.shape_candidates(17) == []                      # odd part 17: no shape
    for Ls in (8, 12, 16, 24, 32, 48, 64):
        for w in at.shape_candidates(Ls):
            if w:
                assert (64 * w) % Ls == 0 and w <= 8


def test_launch_candidates_include_row_tiles():
    """The shape stage tries every W-wave shape with and without the persistent grid, and the LDS
    row tiles once (in the default tile shape)."""
    c = at.launch_candidates([0, 3, 6])
    d = [dict(x) for x in c]
    assert len(c) == 6 and len(set(c)) == 6
    assert sum(x[gx.TUNE_DHOP_TILE] for x in d) == 1
    assert all(x[gx.TUNE_DHOP_PERSIST] == 0 for x in d if x[gx.TUNE_DHOP_SHAPE] == 0)
    assert all(set(x) == set(at.LAUNCH_KNOBS) for x in d)


def test_knobs_per_precision():
    """A knob holds one value per storage precision: tune_set(..., prec) changes that precision's
    value only, tune_set without a precision changes all three, and Knobs restores every one."""
    k = gx.TUNE_DHOP_MAP
    saved = {p: gx.tune_get(k, p) for p in at.PRECS}
    try:
        gx.tune_set(k, 2)
        gx.tune_set(k, 4, gx.F32)
        assert (gx.tune_get(k, gx.F64), gx.tune_get(k, gx.F32), gx.tune_get(k, gx.F16)) == (2, 4, 2)
        assert gx.tune_get(k) == 2
        with at.Knobs(((k, 3),), gx.F16):
            assert (gx.tune_get(k, gx.F64), gx.tune_get(k, gx.F32), gx.tune_get(k, gx.F16)) == (2, 4, 3)
        assert (gx.tune_get(k, gx.F64), gx.tune_get(k, gx.F32), gx.tune_get(k, gx.F16)) == (2, 4, 2)
    finally:
        for p, v in saved.items():
            gx.tune_set(k, v, p)

"""CPU checks of bench.py's roofline bookkeeping: the counter traffic the bench line reports belongs
to the kernel and tile shape that ran, and every committed headline counter summary is consistent
with the bench line of the run that recorded it."""
import glob
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

SUMMARIES = sorted(glob.glob(os.path.join(ROOT, 'profiles', 'dhop_mobius_16x16x16x16_ls16_f64*_traffic.json')))


def test_traffic_matches_kernel_and_shape():
    assert SUMMARIES
    for p in SUMMARIES:
        d = json.load(open(p))
        b, src = bench.traffic_from_profile(d['kernel'], d.get('shape', 0))
        assert b == d['traffic_bytes_per_launch'] and src == os.path.relpath(p, ROOT)
    assert bench.traffic_from_profile('gx::no_such_kernel', 0) == (None, None)
    # the general shape's summary does not stand in for another tile shape of the same kernel
    assert bench.traffic_from_profile('gx::k_dhop_g<double, false, 0>', 7) == (None, None)


def test_summaries_are_consistent():
    for p in SUMMARIES:
        d = json.load(open(p))
        alg = d['algorithmic_bytes_per_launch']
        h = bench.DIMS[0] * bench.DIMS[1] * bench.DIMS[2] * bench.DIMS[3] // 2
        assert alg == h * (24 * bench.LS + 72) * 16
        cal = d['calibration']
        traffic = d['fetch_size_kib_per_launch'] * 1024 * cal['fetch_factor'] + \
            d['write_size_kib_per_launch'] * 1024 * cal['write_factor']
        assert abs(traffic - d['traffic_bytes_per_launch']) <= 1e-3 * traffic
        assert d['traffic_bytes_per_launch'] >= alg  # every compulsory byte moves at least once
        s = d['kernel_stats_same_tree']
        # the profile's average kernel time fits within the bench line's time per step
        assert s['average_us'] <= s['bench_line_ms_per_step'] * 1e3

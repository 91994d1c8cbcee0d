"""Host-side serving logic (no GPU): the solve form per job type, footprint estimates by
form, size classes of the drawn stream, and the combination of a stream served in
segments."""
import json
import os
import subprocess
import sys

import numpy as np

from icx import SOLVE_FUSED, SOLVE_MIXED, serve

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_solve_form_by_dimension():
    assert serve.solve_form(2) == SOLVE_FUSED
    assert serve.solve_form(3) == SOLVE_MIXED


def test_footprint_estimate_by_form():
    """The FP64 fused form holds the mixed form's 84 B per DOF and 16 B more (node-
    interleaved copies of x and b), no FP32 copy of the moduli (4 B per element less) and
    FP64 inverse modulus sums (4 B per node more)."""
    for ndim, div in ((2, (300, 100, 1)), (3, (40, 20, 20))):
        d0, d1, d2 = div
        nn = (d0 + 1) * (d1 + 1) * ((d2 + 1) if ndim == 3 else 1)
        ne = d0 * d1 * (d2 if ndim == 3 else 1)
        mixed = serve.footprint_estimate(ndim, div, form=SOLVE_MIXED)
        fused = serve.footprint_estimate(ndim, div, form=SOLVE_FUSED)
        n = ndim * nn
        assert fused - mixed == 4 * nn - 4 * ne + 16 * n
        assert serve.footprint_estimate(ndim, div) == (fused if ndim == 2 else mixed)


def test_drawn_stream_is_reproducible():
    """The C5 stream's job sizes come from the seed alone."""
    rng = np.random.default_rng(np.random.PCG64(serve.SEED))
    rng.exponential(1.0 / 20.0, 500)
    a = [serve.job_spec(rng, 5e3, 1.5e6) for _ in range(500)]
    rng = np.random.default_rng(np.random.PCG64(serve.SEED))
    rng.exponential(1.0 / 20.0, 500)
    b = [serve.job_spec(rng, 5e3, 1.5e6) for _ in range(500)]
    assert a == b
    assert {s[0] for s in a} == {2, 3}
    elems = [d0 * d1 * d2 for _, (d0, d1, d2) in a]
    assert 4e3 <= min(elems) and max(elems) <= 1.6e6


def test_combine_segments(tmp_path):
    """bench.py --serve-combine adds wall times and DOF-matvecs and takes the percentiles
    over every design iteration of all segments."""
    stems = []
    rng = np.random.default_rng(3)
    all_lat = []
    dmv = wall_tot = 0.0
    for k, wall in enumerate((10.0, 30.0)):
        lat = rng.uniform(1, 100, 50 + 10 * k)
        all_lat.append(lat)
        recs = np.zeros((len(lat), 10))
        recs[:, 2] = 5e4 if k == 0 else 2e6
        recs[:, 3] = 100
        recs[:, 5] = 1
        recs[:, 6] = lat
        value = 1e9 * (k + 1)
        dmv += value * wall
        wall_tot += wall
        line = {"value": value, "dtype": "x", "data": "y",
                "config": {"wall_s": wall, "offset": 100 * k, "jobs": 5, "iteration_ms_p50": 0, "iteration_ms_p99": 0,
                           "unconverged_solves": 0, "iterations_per_job": 30, "elements": [5e3, 1.5e6],
                           "arrival_rate_per_s": 20.0, "budget_bytes": 240e9, "max_resident_dofs": None,
                           "tol": 1e-8}}
        stem = str(tmp_path / ("seg%d" % k))
        open(stem + ".json", "w").write("noise line\n" + json.dumps(line) + "\n")
        np.savez(stem + ".npz", latencies_ms=lat, admission_wait_ms=rng.uniform(0, 5, 5), records=recs)
        stems.append(stem)
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--serve-combine"] + stems,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = json.loads(out.stdout.strip().splitlines()[-1])
    lat = np.concatenate(all_lat)
    assert got["config"]["design_iterations"] == len(lat) and got["config"]["jobs"] == 10
    assert abs(got["config"]["wall_s"] - wall_tot) < 1e-12
    assert abs(got["value"] - dmv / wall_tot) < 1e-6 * got["value"]
    assert abs(got["config"]["iteration_ms_p99"] - np.percentile(lat, 99)) < 1e-12
    assert set(got["config"]["latency_by_dofs"]) == {"0-100000", "1e+06-1e+12"}

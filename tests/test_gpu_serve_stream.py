"""Serving configuration: a seeded Poisson stream of fits of mixed grid sizes
served in real time, every served result checked against the FP64 oracle and
against the same fit run alone.

Stream: 9 requests at 4 req/s (seed 5003), grids 48^3-160^3 including 120^3
(no fused pipeline: the batched kernels) and a core-weighted request, each
with the first 32 rotations of its 10 or 20 degree set so that the oracle
finishes in seconds.
"""
import os
import sys

import numpy as np
import pytest

from oracle import correlator as oc
from powerfit_amd import HipCorrelator
from powerfit_amd.serve import FitServer, make_stream

from parity import assert_scan_parity

pytestmark = pytest.mark.gpu
NPROC = min(8, os.cpu_count() or 1)
SIZES = (64, 120, 48, 96, 160, 80, 128, 144, 72)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_poisson_stream_results_match_oracle_and_standalone():
    reqs = make_stream(len(SIZES), 4.0, seed=5003, sizes=SIZES, max_rotations=32,
                       core_weighted=(1,))
    assert any(not set(np.unique(r.mask)) <= {0.0, 1.0} for r in reqs)
    say("serving stream: " + ", ".join(f"{r.target.shape[0]}^3@{r.arrival:.2f}s"
                                       for r in reqs))
    server = FitServer(max_active=4, depth=2)
    res = server.run(reqs, realtime=True)
    assert sorted(res) == list(range(len(reqs)))
    assert server.used == 0
    for r in reqs:
        out = res[r.rid]
        ref = oc.parallel_scan(np.asarray(r.target, np.float64), r.template, r.mask,
                               r.rotations, laplace=r.laplace, nproc=NPROC)
        err, dec = assert_scan_parity(out.lcc, out.rot, ref)
        corr = HipCorrelator(np.asarray(r.target, np.float64), laplace=r.laplace, batch=16)
        corr.template = r.template
        corr.mask = r.mask
        corr.rotations = r.rotations
        corr.scan()
        assert np.array_equal(out.lcc, corr.lcc) and np.array_equal(out.rot, corr.rot), r.rid
        say(f"  request {r.rid}: {r.target.shape[0]}^3, {r.rotations.shape[0]} rotations, "
            f"latency {out.latency:.3f} s, max |dLCC| {err:.2e}, decisive voxels {dec}")

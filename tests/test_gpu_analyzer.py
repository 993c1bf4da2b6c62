"""Peak extraction on the GPU (pf_peak_positions, Analyzer) against the
oracle's watershed (scipy labelling, ``analyzer.py:80-96``).

Bar: identical peak positions (maps with distinct values); among equal
maxima of one component the lowest flat index; identical solutions rows and
solutions.out for an LCC map from a search."""
import numpy as np
import pytest
from scipy.ndimage import gaussian_filter

from oracle.analyzer import component_maxima, cutoffs, watershed_positions
from powerfit_amd import HipCorrelator
from powerfit_amd.analyzer import Analyzer, peak_positions, watershed_cutoffs
from powerfit_amd.rotations import rotations_for_angle
from powerfit_amd.synth import Workload, make_case

pytestmark = pytest.mark.gpu


def smooth_map(shape, seed, dtype):
    rng = np.random.default_rng(seed)
    m = gaussian_filter(rng.standard_normal(shape), 1.5, mode="wrap")
    return (m / np.abs(m).max()).astype(dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(32, 32, 32), (17, 23, 29), (1, 40, 40), (64, 64, 64)])
def test_component_maxima_match_oracle(shape, dtype):
    print(f"peaks {shape} {np.dtype(dtype).name}", flush=True)
    corr = smooth_map(shape, sum(shape), dtype)
    for c in np.asarray([0.9, 0.5, 0.1, -0.2], dtype=dtype):
        got = peak_positions(corr, [c])
        ref = sorted(np.ravel_multi_index(tuple(zip(*component_maxima(corr, c))), shape)) \
            if component_maxima(corr, c) else []
        assert list(got) == ref, (shape, c)


def test_capacity_retry_with_many_components():
    rng = np.random.default_rng(1)
    corr = rng.random((48, 48, 48)).astype(np.float32)   # thousands of small components
    c = np.float32(0.7)
    got = peak_positions(corr, [c])
    ref = component_maxima(corr, c)
    assert len(got) == len(ref) > 4096
    assert set(got.tolist()) == set(np.ravel_multi_index(tuple(zip(*ref)), corr.shape).tolist())


def test_equal_maxima_take_lowest_index():
    corr = np.zeros((4, 5, 6), np.float32)
    corr[1, 2, 1:5] = [0.5, 0.9, 0.7, 0.9]      # one component, two equal maxima
    corr[3, 0, 0] = 0.9                          # another component
    got = peak_positions(corr, [np.float32(0.4)])
    assert list(got) == sorted([np.ravel_multi_index((1, 2, 2), corr.shape),
                                np.ravel_multi_index((3, 0, 0), corr.shape)])


def test_watershed_cutoffs_equal_oracle():
    for dtype in (np.float32, np.float64):
        corr = smooth_map((16, 16, 16), 3, dtype)
        assert list(watershed_cutoffs(corr, 5)) == [dtype(c) for c in cutoffs(corr, 5)]


def test_analyzer_on_search_result(tmp_path):
    """Peaks of a real LCC map: same positions, rows and file as the reference's
    Analyzer computes from the same map (positions from the oracle)."""
    w = Workload("peaks32", 32, 2.5, 300, 7.0, 20.0, seed=7400)
    c = make_case(w)
    R = rotations_for_angle(20.0)[:96]
    corr = HipCorrelator(c.target.astype(np.float64), laplace=c.laplace, mode="auto", batch=32)
    corr.template = c.template
    corr.mask = c.mask
    corr.rotations = R
    corr.scan()
    lcc, rot = corr.lcc, corr.rot
    print(f"analyzer on 32^3 map, max lcc {lcc.max():.4f}", flush=True)
    a = Analyzer(lcc, R, rot, voxelspacing=2.5, origin=(1.0, -2.0, 3.0), z_sigma=0.3)
    ref_pos = watershed_positions(lcc, 5)
    assert a._positions == ref_pos
    ref = Analyzer.__new__(Analyzer)
    ref.__dict__.update(a.__dict__)
    ref._positions = ref_pos
    ref._generate_solutions()
    assert len(a.solutions) == len(ref_pos) > 0
    for got, exp in zip(a.solutions, ref.solutions):
        assert [float(v) for v in got] == [float(v) for v in exp]
    a.tofile(str(tmp_path / "a.out"))
    ref.tofile(str(tmp_path / "r.out"))
    assert (tmp_path / "a.out").read_text() == (tmp_path / "r.out").read_text()
    a.steps = 3
    assert a._positions == watershed_positions(lcc, 3)


def test_analyzer_equals_reference_fixture(tmp_path):
    """The whole Analyzer (GPU watershed, rows, file) on the map of the golden
    fixture written by the reference's own Analyzer
    (tests/golden/make_output_fixtures.py)."""
    import os
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    f = np.load(os.path.join(gold, "analyzer_ref.npz"))
    print("analyzer on the reference fixture map", flush=True)
    a = Analyzer(f["corr"], f["rotmat"], f["rot_ind"], steps=int(f["steps"]),
                 voxelspacing=float(f["voxelspacing"]),
                 origin=tuple(float(v) for v in f["origin"]), z_sigma=float(f["z_sigma"]))
    assert a._positions == set(tuple(int(v) for v in p) for p in f["positions"])
    assert np.array_equal(np.array([[float(v) for v in r] for r in a.solutions]), f["solutions"])
    a.tofile(str(tmp_path / "solutions.out"))
    with open(os.path.join(gold, "analyzer_ref_solutions.out")) as fh:
        assert (tmp_path / "solutions.out").read_text() == fh.read()

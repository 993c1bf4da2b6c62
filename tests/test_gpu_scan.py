"""GPU end-to-end parity of the LCC search against the FP64 oracle.

Bar: best LCC within 1e-4 absolute of the oracle everywhere; rotation index
identical wherever the oracle's winner beats its runner-up by more than 1e-4.
"""
import os
import sys

import numpy as np
import pytest

from oracle import correlator as oc
from powerfit_amd import HipCorrelator, PowerFitter, Volume
from powerfit_amd.rotations import rotations_for_angle
from powerfit_amd.synth import WORKLOADS, Workload, make_case

from parity import LCC_TOL, assert_scan_parity

pytestmark = pytest.mark.gpu
NPROC = min(8, os.cpu_count() or 1)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


SMALL = {
    "t32": Workload("t32", 32, 2.5, 300, 7.0, 20.0, seed=7000),
    "t32_lap_core": Workload("t32_lap_core", 32, 2.5, 300, 7.0, 20.0, laplace=True,
                             core_weighted=True, seed=7100),
    "t30_lap": Workload("t30_lap", 30, 2.5, 250, 7.0, 20.0, laplace=True, seed=7200),
    "t_odd": Workload("t_odd", 0, 2.5, 200, 7.0, 20.0, core_weighted=True, seed=7300,
                      shape=(27, 25, 30)),
}

_cache = {}


def case_and_oracle(name, nrot=None, angle=None):
    key = (name, nrot, angle)
    if key not in _cache:
        w = SMALL.get(name) or WORKLOADS[name]
        c = make_case(w)
        R = rotations_for_angle(angle or w.angle)
        if nrot is not None:
            R = R[:nrot]
        ref = oc.parallel_scan(c.target.astype(np.float64), c.template, c.mask, R,
                               laplace=c.laplace, nproc=NPROC)
        _cache[key] = (c, R, ref)
    return _cache[key]


def run_hip(c, R, mode="batched", batch=16):
    corr = HipCorrelator(c.target.astype(np.float64), laplace=c.laplace, mode=mode, batch=batch)
    corr.template = c.template
    corr.mask = c.mask
    corr.rotations = R
    corr.scan()
    return corr.lcc, corr.rot


@pytest.mark.parametrize("name", list(SMALL))
def test_small_full_rotation_set(name):
    c, R, ref = case_and_oracle(name)
    say(f"scan {name} shape={c.target.shape} K={len(R)} batched(16)")
    best, rot = run_hip(c, R, "batched", 16)
    err, dec = assert_scan_parity(best, rot, ref)
    say(f"  max err {err:.2e}, decisive voxels {dec}")


def test_plain_and_batch_sizes_identical():
    """Batching changes launch shapes only: results are bit-identical."""
    c, R, ref = case_and_oracle("t32")
    say("scan t32 plain vs batched(7) vs batched(16)")
    b16 = run_hip(c, R, "batched", 16)
    b7 = run_hip(c, R, "batched", 7)
    p1 = run_hip(c, R, "plain")
    for other in (b7, p1):
        assert np.array_equal(b16[0], other[0]) and np.array_equal(b16[1], other[1])
    assert_scan_parity(*p1, ref)


def test_split_ranges_accumulate():
    """Two scans over [0, K/2) and [K/2, K) without reset equal one scan."""
    c, R, ref = case_and_oracle("t32")
    say("scan t32 split ranges")
    corr = HipCorrelator(c.target.astype(np.float64), laplace=c.laplace, batch=32)
    corr.template = c.template
    corr.mask = c.mask
    corr.rotations = R
    corr.scan()
    whole = (corr.lcc.copy(), corr.rot.copy())
    corr.reset()
    h = len(R) // 2
    corr.enqueue_scan(0, h)
    corr.enqueue_scan(h, len(R) - h)
    corr.sync()
    import ctypes
    from powerfit_amd import _lib
    best = np.zeros(c.target.shape, np.float32)
    rot = np.zeros(c.target.shape, np.int32)
    _lib.check(_lib.load().pf_get_result(corr._ctx, _lib.ptr(best, _lib.FP),
                                         _lib.ptr(rot, _lib.I32P)))
    assert np.array_equal(best, whole[0]) and np.array_equal(rot, whole[1])


@pytest.mark.parametrize("name,nrot,mode", [("cpu64", 96, "batched"), ("lc96", 24, "batched"),
                                            ("search128", 12, "batched"),
                                            ("batch128", 12, "auto"), ("prec192", 6, "auto")])
def test_configs_rotation_subset(name, nrot, mode):
    """Each benchmark configuration (64^3 plain, 96^3 Laplace + core weighting,
    128^3 at 10 and 5 degrees, 192^3) on a prefix of its rotation set."""
    c, R, ref = case_and_oracle(name, nrot=nrot)
    say(f"scan {name} shape={c.target.shape} first {nrot} rotations, mode {mode}")
    best, rot = run_hip(c, R, mode, 8)
    err, dec = assert_scan_parity(best, rot, ref)
    say(f"  max err {err:.2e}, decisive voxels {dec}")


def test_perfect_fit():
    """Target = template, identity first: LCC 1 at voxel 0 with rotation 0."""
    w = Workload("pf", 32, 2.5, 300, 7.0, 20.0, seed=7400)
    c = make_case(w)
    target = c.template.copy()
    R = rotations_for_angle(20.0)[:5]
    corr = HipCorrelator(target)
    assert corr._lcc_mask[0, 0, 0] == 1
    corr.template = c.template
    corr.mask = c.mask
    corr.rotations = R
    corr.scan()
    assert abs(float(corr.lcc[0, 0, 0]) - 1.0) < 1e-5
    assert int(corr.rot[0, 0, 0]) == 0
    assert corr.lcc.max() <= 1.0 + 1e-5


def test_powerfitter_interface(tmp_path):
    c, R, ref = case_and_oracle("t30_lap")
    say("PowerFitter interface")
    pf = PowerFitter(Volume(c.target.astype(np.float64)), laplace=c.laplace)
    pf._rotations = R
    pf._template = Volume(c.template)
    pf._mask = Volume(c.mask)
    pf.directory = str(tmp_path)
    pf.scan()
    assert pf._lcc.shape == c.target.shape and pf._lcc.dtype == np.float32
    assert pf._rot.dtype == np.int32
    assert_scan_parity(pf._lcc, pf._rot, ref)


def test_scan_before_setup_raises():
    c = make_case(SMALL["t32"])
    corr = HipCorrelator(c.target.astype(np.float64))
    with pytest.raises(ValueError):
        corr.scan()


FUSED_CASES = [("cpu64", 96), ("lc96", 24), ("search128", 12)]


@pytest.mark.parametrize("mode", ["mix_fwd", "mix_z", "mix_inv", "fused"])
@pytest.mark.parametrize("name,nrot", FUSED_CASES)
def test_fused_stages_vs_oracle(name, nrot, mode):
    """Each fused stage alone (the rest unfused), then all three fused."""
    c, R, ref = case_and_oracle(name, nrot=nrot)
    say(f"scan {name} mode={mode} first {nrot} rotations")
    best, rot = run_hip(c, R, mode, 16)
    err, dec = assert_scan_parity(best, rot, ref)
    say(f"  max err {err:.2e}, decisive voxels {dec}")


def test_fused_without_binary_shortcut():
    c, R, ref = case_and_oracle("cpu64", nrot=96)
    say("scan cpu64 fused, W^2 transformed separately")
    corr = HipCorrelator(c.target.astype(np.float64), laplace=c.laplace, mode="fused", batch=16,
                         binary_shortcut=False)
    corr.template = c.template
    corr.mask = c.mask
    corr.rotations = R
    corr.scan()
    assert_scan_parity(corr.lcc, corr.rot, ref)


def test_fused_batch_sizes_identical():
    """Chunking and batch boundaries do not change any per-rotation value."""
    c, R, ref = case_and_oracle("cpu64", nrot=96)
    say("scan cpu64 fused batch 5 / 16 / 40")
    outs = [run_hip(c, R, "fused", b) for b in (5, 16, 40)]
    for other in outs[1:]:
        assert np.array_equal(outs[0][0], other[0]) and np.array_equal(outs[0][1], other[1])


def test_tuning_does_not_change_results():
    """Per-block rotation chunks change launch shapes only."""
    c, R, ref = case_and_oracle("cpu64", nrot=96)
    say("scan cpu64 fused with chunk sizes (1,1) (3,5) (16,16) auto")
    outs = []
    for cz, cl in ((1, 1), (3, 5), (16, 16), (0, 0)):
        corr = HipCorrelator(c.target.astype(np.float64), laplace=c.laplace, mode="fused",
                             batch=32)
        corr.set_tuning(cz, cl)
        corr.template = c.template
        corr.mask = c.mask
        corr.rotations = R
        corr.scan()
        outs.append((corr.lcc.copy(), corr.rot.copy()))
    for other in outs[1:]:
        assert np.array_equal(outs[0][0], other[0]) and np.array_equal(outs[0][1], other[1])
    assert_scan_parity(*outs[0], ref)


def test_autotune_picks_a_working_shape(tmp_path, monkeypatch):
    from powerfit_amd.autotune import autotune
    monkeypatch.setenv("POWERFIT_AMD_TUNE", str(tmp_path / "tune.json"))
    c, R, ref = case_and_oracle("cpu64", nrot=96)
    corr = HipCorrelator(c.target.astype(np.float64), laplace=c.laplace, mode="fused")
    corr.template = c.template
    corr.mask = c.mask
    corr.rotations = R
    t = autotune(corr, batches=(16, 32), chunks=((0, 0), (4, 4)), nrot=64)
    assert t["batch"] in (16, 32) and t["rate"] > 0 and not t["cached"]
    assert autotune(corr, batches=(16, 32), chunks=((0, 0), (4, 4)), nrot=64)["cached"]
    corr.scan()
    assert_scan_parity(corr.lcc, corr.rot, ref)


@pytest.mark.parametrize("mode", ["fused", "batched"])
def test_graph_replays_match_eager(mode):
    """A captured scan replayed several times, with eager work and syncs in
    between, gives the eager result every time."""
    c, R, ref = case_and_oracle("cpu64", nrot=96)
    say(f"graph capture cpu64 mode={mode}")
    corr = HipCorrelator(c.target.astype(np.float64), laplace=c.laplace, mode=mode, batch=16)
    corr.template = c.template
    corr.mask = c.mask
    corr.rotations = R
    corr.scan()
    eager = (corr.lcc.copy(), corr.rot.copy())
    for rep in range(3):
        corr.reset()
        corr.enqueue_scan_graph(0, len(R))
        lcc, rot = corr.result()
        assert np.array_equal(lcc, eager[0]) and np.array_equal(rot, eager[1]), rep
        corr.scan()                      # eager work between replays
        assert np.array_equal(corr.lcc, eager[0])
    assert_scan_parity(*eager, ref)


@pytest.mark.parametrize("n", [128, 192])
def test_graph_replays_match_eager_multi_launch_stages(n):
    """Graph replays where K1 runs one launch per grid (128, 192) and K3 keeps
    the variance in the work buffer (192): every replay equals the eager scan."""
    w = WORKLOADS["search128"] if n == 128 else NEW_SIZES[192]
    c = make_case(w)
    R = rotations_for_angle(w.angle)[:32]
    say(f"graph capture {n}^3 fused, 32 rotations")
    corr = HipCorrelator(c.target.astype(np.float64), laplace=c.laplace, mode="fused", batch=16)
    corr.template = c.template
    corr.mask = c.mask
    corr.rotations = R
    corr.scan()
    eager = (corr.lcc.copy(), corr.rot.copy())
    for rep in range(3):
        corr.reset()
        corr.enqueue_scan_graph(0, len(R))
        lcc, rot = corr.result()
        assert np.array_equal(lcc, eager[0]) and np.array_equal(rot, eager[1]), rep
        corr.scan()
        assert np.array_equal(corr.lcc, eager[0])


def test_serving_results_equal_standalone():
    """Fits admitted together and interleaved on the device give exactly the
    result each gives alone."""
    from powerfit_amd.serve import FitRequest, FitServer
    say("serving: three fits alone and in the mix")
    reqs = []
    for i, (name, nrot) in enumerate([("cpu64", 40), ("t_odd", 40), ("lc96", 24)]):
        w = SMALL.get(name) or WORKLOADS[name]
        cs = make_case(w)
        R = rotations_for_angle(w.angle)[:nrot]
        reqs.append(FitRequest(i, cs.target, cs.template, cs.mask, R, cs.laplace, 0.0))
    alone = {}
    for r in reqs:
        corr = HipCorrelator(r.target.astype(np.float64), laplace=r.laplace, batch=8)
        corr.template = r.template
        corr.mask = r.mask
        corr.rotations = r.rotations
        corr.scan()
        alone[r.rid] = (corr.lcc.copy(), corr.rot.copy())
    server = FitServer(max_active=3, batch=8, depth=2)
    res = server.run(reqs, realtime=False)
    assert sorted(res) == [0, 1, 2] and server.used == 0
    for rid, (lcc, rot) in alone.items():
        assert np.array_equal(res[rid].lcc, lcc) and np.array_equal(res[rid].rot, rot), rid


def test_serving_keeps_target_spectra_resident():
    """Requests on the same map reuse its idle context (spectra stay in HBM),
    with results identical to standalone fits; idle contexts stay counted
    against the budget until released."""
    from powerfit_amd.serve import FitRequest, FitServer
    say("serving: four fits on two maps, contexts reused")
    maps = []
    for name in ("cpu64", "lc96"):
        w = SMALL.get(name) or WORKLOADS[name]
        maps.append((make_case(w), rotations_for_angle(w.angle)))
    reqs = []
    for i in range(4):
        cs, R = maps[i % 2]
        Ri = R[:24] if i < 2 else R[24:56]
        reqs.append(FitRequest(i, cs.target, cs.template, cs.mask, Ri, cs.laplace, 0.0,
                               f"map{i % 2}"))
    server = FitServer(max_active=1, batch=8, depth=2)
    res = server.run(reqs, realtime=False)
    assert sorted(res) == [0, 1, 2, 3] and server.reused == 2
    assert server.used > 0 and len(server.idle) == 2
    for r in reqs:
        corr = HipCorrelator(r.target.astype(np.float64), laplace=r.laplace, batch=8)
        corr.template = r.template
        corr.mask = r.mask
        corr.rotations = r.rotations
        corr.scan()
        assert np.array_equal(res[r.rid].lcc, corr.lcc), r.rid
        assert np.array_equal(res[r.rid].rot, corr.rot), r.rid
    server.release_idle()
    assert server.used == 0 and not server.idle


def test_serving_admission_respects_budget():
    from powerfit_amd.serve import FitRequest, FitServer, ctx_bytes
    w = WORKLOADS["cpu64"]
    cs = make_case(w)
    R = rotations_for_angle(w.angle)[:16]
    need = ctx_bytes(cs.target.shape, 8)
    reqs = [FitRequest(i, cs.target, cs.template, cs.mask, R, False, 0.0) for i in range(3)]
    server = FitServer(hbm_budget=int(need * 1.5), max_active=8, batch=8)
    peak = []
    orig = server._pump

    def pump():
        peak.append(len(server.active))
        orig()
    server._pump = pump
    res = server.run(reqs, realtime=False)
    assert len(res) == 3 and max(peak) == 1     # budget admits one fit at a time
    for r in res.values():
        assert np.array_equal(r.lcc, res[0].lcc)


NEW_SIZES = {
    48: Workload("f48", 48, 2.0, 600, 8.0, 20.0, seed=7500),
    72: Workload("f72", 72, 1.8, 2000, 7.0, 20.0, laplace=True, core_weighted=True, seed=7600),
    80: Workload("f80", 80, 1.5, 2500, 6.0, 20.0, seed=7700),
    144: Workload("f144", 144, 1.2, 6000, 5.0, 20.0, laplace=True, seed=7800),
    160: Workload("f160", 160, 1.2, 8000, 5.0, 20.0, seed=7900),
    192: Workload("f192", 192, 1.0, 10000, 5.0, 20.0, core_weighted=True, seed=8000),
}


@pytest.mark.parametrize("n", sorted(NEW_SIZES))
def test_fused_serving_sizes_vs_oracle(n):
    w = NEW_SIZES[n]
    c = make_case(w)
    R = rotations_for_angle(w.angle)[:48]
    ref = oc.parallel_scan(c.target.astype(np.float64), c.template, c.mask, R,
                           laplace=c.laplace, nproc=NPROC)
    say(f"scan {n}^3 fused, 48 rotations")
    corr = HipCorrelator(c.target.astype(np.float64), laplace=c.laplace, mode="fused", batch=16)
    assert corr.fused_available
    corr.template = c.template
    corr.mask = c.mask
    corr.rotations = R
    corr.scan()
    err, dec = assert_scan_parity(corr.lcc, corr.rot, ref)
    say(f"  max err {err:.2e}, decisive voxels {dec}")


BINARY_SIZES = {
    72: Workload("b72", 72, 1.8, 2000, 7.0, 20.0, seed=7650),
    96: Workload("b96", 96, 1.5, 3000, 6.0, 20.0, seed=7950),
    192: Workload("b192", 192, 1.0, 10000, 5.0, 20.0, seed=8050),
}


@pytest.mark.parametrize("n", sorted(BINARY_SIZES))
def test_fused_binary_mask_sizes_vs_oracle(n):
    """Sizes whose other GPU cases use weighted masks: the 0/1-mask form of
    their first kernel (two-plane at 72, one launch per grid at 192)."""
    w = BINARY_SIZES[n]
    c = make_case(w)
    assert set(np.unique(c.mask)) <= {0.0, 1.0}
    R = rotations_for_angle(w.angle)[:24]
    ref = oc.parallel_scan(c.target.astype(np.float64), c.template, c.mask, R,
                           laplace=c.laplace, nproc=NPROC)
    say(f"scan {n}^3 fused, 0/1 mask, 24 rotations")
    best, rot = run_hip(c, R, "fused", 8)
    err, dec = assert_scan_parity(best, rot, ref)
    say(f"  max err {err:.2e}, decisive voxels {dec}")


SCHEDULE_CASES = [("f48", 48), ("cpu64", 64), ("f72", 72), ("f80", 80), ("lc96", 96),
                  ("search128", 128), ("f144", 144), ("f160", 160), ("f192", 192)]


@pytest.mark.parametrize("name,n", SCHEDULE_CASES)
def test_alternative_z_schedule_vs_oracle(name, n):
    """Every z-stage schedule (radix plan x column tile, and the register z
    stage with either register plan) at every fused size: oracle parity,
    agreement with the default to FP32 rounding, and the tile width alone
    changes nothing."""
    w = NEW_SIZES[n] if n in NEW_SIZES else WORKLOADS[name]
    c = make_case(w)
    R = rotations_for_angle(w.angle)[:24]
    ref = oc.parallel_scan(c.target.astype(np.float64), c.template, c.mask, R,
                           laplace=c.laplace, nproc=NPROC)
    say(f"scan {n}^3 fused, z schedules 0-5, 24 rotations")
    outs = []
    for zs in range(6):
        corr = HipCorrelator(c.target.astype(np.float64), laplace=c.laplace, mode="fused",
                             batch=8)
        corr.set_schedule(zs)
        corr.template = c.template
        corr.mask = c.mask
        corr.rotations = R
        corr.scan()
        err, dec = assert_scan_parity(corr.lcc, corr.rot, ref)
        say(f"  schedule {zs}: max err {err:.2e}, decisive voxels {dec}")
        outs.append((corr.lcc.copy(), corr.rot.copy()))
    for zs in (1, 3, 4, 5):
        assert np.abs(outs[0][0] - outs[zs][0]).max() < 1e-5
    for plan in (0, 1):   # schedules plan and plan + 2 differ in the tile width only
        assert np.array_equal(outs[plan][0], outs[plan + 2][0])
        assert np.array_equal(outs[plan][1], outs[plan + 2][1])


LARGE_MIX = [(160, "mix_fwd"), (160, "mix_inv"), (192, "mix_fwd"), (192, "mix_inv")]


@pytest.mark.parametrize("n,mode", LARGE_MIX)
def test_large_grid_stage_mix_vs_oracle(n, mode):
    """The large grids' K1 (mask values parked in the work buffer, one launch
    per grid) and K3 (variance parked in the consumed B plane, column sweeps)
    next to the unfused stages, which read and write every grid themselves."""
    w = NEW_SIZES[n]
    c = make_case(w)
    R = rotations_for_angle(w.angle)[:8]
    ref = oc.parallel_scan(c.target.astype(np.float64), c.template, c.mask, R,
                           laplace=c.laplace, nproc=NPROC)
    say(f"scan {n}^3 mode={mode}, 8 rotations")
    best, rot = run_hip(c, R, mode, 4)
    err, dec = assert_scan_parity(best, rot, ref)
    say(f"  max err {err:.2e}, decisive voxels {dec}")


@pytest.mark.parametrize("n", [160, 192])
def test_large_grid_chunks_and_weighted_mask(n):
    """Chunks of 1 and several rotations per K3 block, and the W^2 grid
    transformed separately, give the same maps (chunks) and oracle parity."""
    w = NEW_SIZES[n]
    c = make_case(w)
    R = rotations_for_angle(w.angle)[:12]
    ref = oc.parallel_scan(c.target.astype(np.float64), c.template, c.mask, R,
                           laplace=c.laplace, nproc=NPROC)
    say(f"scan {n}^3 fused, K3 chunks 1 and 5, W^2 separate")
    outs = []
    for cz, cl in ((1, 1), (2, 5)):
        corr = HipCorrelator(c.target.astype(np.float64), laplace=c.laplace, mode="fused",
                             batch=12)
        corr.set_tuning(cz, cl)
        corr.template = c.template
        corr.mask = c.mask
        corr.rotations = R
        corr.scan()
        outs.append((corr.lcc.copy(), corr.rot.copy()))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    assert_scan_parity(outs[0][0], outs[0][1], ref)
    corr = HipCorrelator(c.target.astype(np.float64), laplace=c.laplace, mode="fused", batch=6,
                         binary_shortcut=False)
    corr.template = c.template
    corr.mask = c.mask
    corr.rotations = R
    corr.scan()
    err, dec = assert_scan_parity(corr.lcc, corr.rot, ref)
    say(f"  max err {err:.2e}, decisive voxels {dec}")

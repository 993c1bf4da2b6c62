"""Launch-shape and FFT-schedule autotuner for the LCC search.

Times short scans of the correlator's own inputs over candidate rotation-batch
sizes and per-block rotation chunks of the fused kernels, then, at the best
of those, the other schedules of the fused z stage (alternative FFT radix
plan, alternative column tile width, both, and the run-time planned
pipeline's register z stage with either register plan; for the run-time
planned pipeline its alternative register plan), and keeps the fastest.  Results are cached per (device, grid, mask kind, mode, precision) in a JSON
file (``$POWERFIT_AMD_TUNE`` or ``~/.cache/powerfit_amd/tune.json``).  Batch
and chunk sizes do not change the result (every rotation's LCC map is computed
identically and merged with an order-independent max); the FFT schedule
changes it only by FP32 rounding.  With split precision the fused z stage
has one schedule (the one whose first stage runs on the matrix cores), so only
batch and chunks are tuned, and those choices are cached separately.
"""
import json
import os
import time

import numpy as np

DEFAULT_BATCHES = (32, 64, 128, 256, 512)
DEFAULT_CHUNKS = ((0, 0), (4, 4), (8, 8), (16, 16))


def _cache_path():
    return os.environ.get("POWERFIT_AMD_TUNE",
                          os.path.join(os.path.expanduser("~"), ".cache", "powerfit_amd",
                                       "tune.json"))


def _device_name():
    try:
        import torch
        return torch.cuda.get_device_name(0)
    except Exception:  # noqa: BLE001 - name is only a cache key
        return "gpu"


def _key(corr):
    binary = bool(np.all((corr.mask == 0) | (corr.mask == 1)))
    precision = getattr(corr, "precision", "fp32")
    return f"{_device_name()}|{'x'.join(map(str, corr.target.shape))}|" \
           f"{'binary' if binary else 'weighted'}|{corr.mode}|{precision}"


def _time(corr, batch, chunks, nrot, repeats=2, z_schedule=0):
    corr.batch = batch
    corr.set_tuning(*chunks)
    corr.set_schedule(z_schedule)
    K = corr.rotations.shape[0]
    n = min(nrot, K)
    corr.reset()
    corr.enqueue_scan(0, min(batch, n))       # warm the shapes
    corr.sync()
    best = float("inf")
    for _ in range(repeats):
        corr.reset()
        t = time.perf_counter()
        corr.enqueue_scan(0, n)
        corr.sync()
        best = min(best, time.perf_counter() - t)
    return n / best


def autotune(corr, batches=DEFAULT_BATCHES, chunks=DEFAULT_CHUNKS, nrot=None, use_cache=True,
             verbose=False, schedules=(0, 1, 2, 3, 4, 5)):
    """Pick (batch, chunk_z, chunk_lcc, z_schedule) for ``corr`` (a ready
    HipCorrelator).

    Applies the choice to ``corr`` and returns
    ``{"batch", "chunk_z", "chunk_lcc", "z_schedule", "rate", "trials"}``.
    """
    path = _cache_path()
    key = _key(corr)
    cache = {}
    if use_cache and os.path.exists(path):
        try:
            with open(path) as fh:
                cache = json.load(fh)
        except (OSError, ValueError):
            cache = {}
    if use_cache and key in cache:
        hit = cache[key]
        corr.batch = hit["batch"]
        corr.set_tuning(hit["chunk_z"], hit["chunk_lcc"])
        corr.set_schedule(hit.get("z_schedule", 0))
        return dict(hit, cached=True)
    nrot = nrot or 2 * max(batches)
    trials = []
    for b in batches:
        for cz, cl in chunks:
            if cz > b or cl > b:
                continue
            rate = _time(corr, b, (cz, cl), nrot)
            trials.append({"batch": b, "chunk_z": cz, "chunk_lcc": cl, "rate": rate})
            if verbose:
                print(f"autotune {key}: batch {b} chunks ({cz},{cl}) -> {rate:.0f} rot/s",
                      flush=True)
    best = max(trials, key=lambda t: t["rate"])
    best["z_schedule"] = 0
    # split precision runs the fused z stage's default schedule only; the
    # run-time planned z stage reads bit 0 only (its register plan)
    if (corr.fused_available and corr.mode in ("auto", "fused")
            and getattr(corr, "precision", "fp32") != "split"):
        z_trials = [zs for zs in schedules if zs != 0]
    elif corr.fused_rt_available and corr.mode in ("auto", "fused_rt"):
        z_trials = [zs for zs in schedules if zs == 1]
    else:
        z_trials = []
    if z_trials:
        for zs in z_trials:
            rate = _time(corr, best["batch"], (best["chunk_z"], best["chunk_lcc"]), nrot,
                         z_schedule=zs)
            trials.append({"batch": best["batch"], "chunk_z": best["chunk_z"],
                           "chunk_lcc": best["chunk_lcc"], "z_schedule": zs, "rate": rate})
            if verbose:
                print(f"autotune {key}: z schedule {zs} -> {rate:.0f} rot/s", flush=True)
            if rate > best["rate"]:
                best = dict(trials[-1])
    corr.batch = best["batch"]
    corr.set_tuning(best["chunk_z"], best["chunk_lcc"])
    corr.set_schedule(best["z_schedule"])
    out = dict(best, trials=len(trials))
    if use_cache:
        cache[key] = out
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            with open(path, "w") as fh:
                json.dump(cache, fh, indent=1)
        except OSError:
            pass
    return dict(out, cached=False)

"""Benchmark: rotations/s of the exhaustive LCC search at 128^3 on one MI355X.

Workload (default ``search128``): a 128^3 float32 target (1.2 A voxels) made
from a seeded 12,000-atom synthetic chain blurred at 5 A, placed at a random
pose with N(0, 0.05 max) noise; template and threshold mask from the same
chain; the 10-degree rotation set (c48n309, K = 7416).  All inputs are
resident in HBM before timing starts.

A step is one batch of ``--batch`` rotations through the whole chain (rotate
-> 3 x R2C -> conj-multiply -> 3 x C2R -> LCC + take-best).  The timed region
covers ``--steps`` steps (default: the whole rotation set, rounded down to
whole batches), bracketed by device synchronisations.  Kernel durations come
from HIP events recorded on the search's own stream during the timed region.

Prints one JSON line.  ``--gpus`` above 1 is refused: this engine runs on one
GPU.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def alg_bytes(kernel, V, H, binary=False):
    """(bytes per rotation, bytes per launch) a kernel class must move
    (f32 = 4 B, complex64 = 8 B).  With a 0/1 mask the fused path transforms W
    once (FFT(W^2) = FFT(W)): 2 xy-spectra instead of 3.  The per-launch part
    is what a launch reads or writes once whatever its rotation count (target
    spectra, lcc mask, the key read and write)."""
    nsp = 2 if binary else 3
    per_rot = {
        "rotate": 8 * V + 12 * V,          # read template+mask, write P, W, W^2
        "r2c_x": 12 * V + 24 * H,          # read 3 real grids, write 3 half spectra
        "c2c_fwd": 48 * H,                 # one strided pass over 3 spectra (r + w)
        "conj_mul": 48 * H,                # 3 spectra r + w
        "c2c_inv": 48 * H,
        "c2r_x": 24 * H + 12 * V,
        "lcc": 12 * V,                     # G, A, B
        # fused kernels (xy-spectra in numpy layout between them)
        "fwd_plane": 8 * V + 8 * nsp * H,  # gather template+mask, write the xy-spectra
        "zcorr": 8 * nsp * H + 24 * H,     # xy-spectra in, 3 correlation spectra out
        "inv_plane": 24 * H,               # 3 spectra in
    }[kernel]
    per_launch = {"conj_mul": 16 * H, "lcc": V + 16 * V, "zcorr": 16 * H,
                  "inv_plane": V + 16 * V}.get(kernel, 0)  # F1,F2 / mask, keys r+w
    return per_rot, per_launch


def measured_traffic(workload, mode, kernel, rotations_per_launch, binary):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (profiles/<workload>_fused_traffic.json, bytes per rotation),
    at the average rotations per launch; None when no summary covers this
    configuration."""
    path = os.path.join(ROOT, "profiles", f"{workload}_fused_traffic.json")
    if mode not in ("fused", "auto") or not os.path.exists(path):
        return None
    with open(path) as fh:
        prof = json.load(fh)
    k = prof.get("kernels", {}).get(kernel)
    if k is None or prof.get("binary_mask") != binary:
        return None
    return k["bytes_per_rotation"] * rotations_per_launch


def path_bytes_per_rotation(V, H):
    """76 V + 120 H: the reference GPU dataflow's bytes per rotation."""
    return 76 * V + 120 * H


def run_scan(corr, K, B, steps, first0=0):
    """Enqueue `steps` batches of B rotations (wrapping around the set)."""
    first = first0
    for _ in range(steps):
        left = B
        while left:
            n = min(left, K - first)
            corr.enqueue_scan(first, n)
            left -= n
            first = (first + n) % K
    return first


# ---- the CPU leg: the only place bench.py runs the oracle (FP64 numpy
# restatement of the reference's CPU path) -- timed as the CPU baseline, and
# as the checker of GPU results on samples
def cpu_workers():
    """Host cores for the CPU baseline: the cores this process may run on,
    at most 16 (a one-GPU share of the host; os.cpu_count() reports the whole
    machine)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, R, nproc, per_worker):
    """The oracle (FP64 numpy, the reference's CPU fallback path) with nproc
    forked workers and the reference's contiguous chunking over the first
    nproc * per_worker rotations; returns (rotations, seconds)."""
    from oracle import correlator as oc
    n = min(nproc * per_worker, R.shape[0])
    done, secs = oc.timed_parallel_scan(case.target.astype(np.float64), case.template, case.mask,
                                        R[:n], laplace=case.laplace, nproc=nproc)
    return done, secs


def cpu_baseline_check(case, Rs, lcc, rot, nproc=1, ref=None):
    """Oracle result on rotations Rs (or `ref`, computed before) and the GPU
    (lcc, rot) compared with it."""
    if ref is None:
        from oracle import correlator as oc
        if nproc > 1:
            ref = oc.parallel_scan(case.target.astype(np.float64), case.template, case.mask, Rs,
                                   laplace=case.laplace, nproc=nproc)
        else:
            ref = oc.scan(case.target.astype(np.float64), case.template, case.mask, Rs,
                          laplace=case.laplace)
    decisive = ref.margin() > 1e-4
    return {"max_abs_lcc_err": float(np.abs(lcc - ref.best).max()),
            "index_mismatches": int(((rot != ref.rot.astype(np.int64)) & decisive).sum()),
            "decisive_voxels": int(decisive.sum()), "tolerance": 1e-4}


def precision_report(args, HipCorrelator, case, R, w):
    """FP32 vs split-precision (MFMA first stage) unfused passes on a rotation
    prefix, and the fused pipeline where the grid has one (FP32, and split
    where its z stage has the MFMA form): rotations/s with inputs resident,
    and every result checked against the FP64 oracle in the CPU leg."""
    n = min(args.precision_rotations, R.shape[0])
    Rs = R[:n]
    rows = {}
    results = {}
    runs = [("fp32", "fp32", "batched"), ("split", "split", "batched")]
    probe = HipCorrelator(case.target.astype(np.float64), laplace=case.laplace, device=0,
                          mode="batched", batch=1)
    if probe.fused_available:
        runs.append(("fp32_fused", "fp32", "fused"))
    del probe
    probe = HipCorrelator(case.target.astype(np.float64), laplace=case.laplace, device=0,
                          mode="batched", batch=1, precision="split")
    if probe.fused_available:
        runs.append(("split_fused", "split", "fused"))
    del probe
    for key, prec, mode in runs:
        corr = HipCorrelator(case.target.astype(np.float64), laplace=case.laplace, device=0,
                             mode=mode, batch=args.precision_batch, precision=prec)
        corr.template = case.template
        corr.mask = case.mask
        corr.rotations = Rs
        corr.scan()                      # warm-up, also the result checked below
        results[key] = (corr.lcc.copy(), corr.rot.copy())
        best = float("inf")
        for _ in range(2):
            corr.reset()
            corr.sync()
            t = time.perf_counter()
            corr.enqueue_scan(0, n)
            corr.sync()
            best = min(best, time.perf_counter() - t)
        rows[key] = {"rotations_per_s": n / best, "path": f"{mode}, batch {args.precision_batch}"}
        del corr
    from oracle import correlator as oc
    ref = oc.parallel_scan(case.target.astype(np.float64), case.template, case.mask, Rs,
                           laplace=case.laplace, nproc=cpu_workers())
    for key in rows:
        rows[key].update(cpu_baseline_check(case, Rs, *results[key], ref=ref))
    nz, ny, nx = case.target.shape
    out = {"workload": f"{w.name}: {nz}x{ny}x{nx}, first {n} rotations of the {w.angle} deg set",
           **rows,
           "split_vs_fp32_rate": rows["split"]["rotations_per_s"] / rows["fp32"]["rotations_per_s"]}
    return out


def device_copy_gbs(torch, mib=1024, reps=10):
    """Measured HBM bandwidth of a plain device-to-device copy (bytes read +
    written per second), the practical ceiling the roofline peak is set
    against."""
    n = (mib << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t
    del a, b
    return 2.0 * (mib << 20) * reps / secs / 1e9


def serving_bench(args):
    """Serving configuration: a Poisson stream of mixed-size fits through
    FitServer (throughput, p50/p99 latency overall and per grid size), then
    the same stream with every rotation set cut to a prefix, served again,
    each served map checked against the FP64 oracle on that prefix."""
    import dataclasses

    from powerfit_amd.serve import FitServer, make_stream, summarize

    stream = make_stream(args.serve_requests, args.serve_rate, seed=5003)
    print(f"# stream of {len(stream)} requests ready", file=sys.stderr, flush=True)

    def serve(max_rot):
        reqs = [dataclasses.replace(r, rotations=r.rotations[:max_rot] if max_rot else r.rotations)
                for r in stream]
        server = FitServer(max_active=args.serve_max_active, depth=args.serve_depth)
        t = time.perf_counter()
        res = server.run(reqs)
        return reqs, res, time.perf_counter() - t, server

    reqs, res, wall, server = serve(0)
    out = {"metric": "serving: rotations/s and latency of a Poisson fit stream",
           "config": {"workload": "serving: Poisson stream (seed 5003), 48-160 cubed maps, "
                                  "500-20,000-atom chains at 4-12 A, 10 or 20 deg sets",
                      "requests": args.serve_requests, "rate_per_s": args.serve_rate,
                      "max_active": args.serve_max_active, "depth": args.serve_depth},
           "summary": summarize(res, wall, args.serve_rate),
           "requests": [{"request": r.rid, "n": r.shape[0], "rotations": r.rotations,
                         "arrival_s": round(r.arrival, 4),
                         "queued_s": round(r.admitted - r.arrival, 4),
                         "latency_s": round(r.latency, 4)}
                        for r in sorted(res.values(), key=lambda r: r.rid)]}
    print(json.dumps(out), flush=True)
    if args.serve_check > 0:
        # checker: the FP64 oracle (the reference's CPU path) on each request's prefix
        from oracle import correlator as oc
        reqs, res, wall, _ = serve(args.serve_check)
        rows = []
        for r in reqs:
            ref = oc.parallel_scan(np.asarray(r.target, np.float64), r.template, r.mask,
                                   r.rotations, laplace=r.laplace, nproc=cpu_workers())
            row = {"request": r.rid, "n": int(r.target.shape[0]),
                   "rotations": int(r.rotations.shape[0])}
            row.update(cpu_baseline_check(None, None, res[r.rid].lcc, res[r.rid].rot, ref=ref))
            rows.append(row)
            print(json.dumps(row), flush=True)
        print(json.dumps({"serving_check": {
            "prefix_rotations": args.serve_check, "requests": len(rows),
            "max_abs_lcc_err": max(x["max_abs_lcc_err"] for x in rows),
            "index_mismatches": sum(x["index_mismatches"] for x in rows),
            "tolerance": 1e-4}}), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=0,
                    help="rotations per launch (0 = autotune batch and chunk sizes)")
    ap.add_argument("--chunks", default=None,
                    help="rotations per block of the fused z and LCC kernels, 'cz,cl' "
                         "(with --batch; default automatic)")
    ap.add_argument("--z-schedule", type=int, default=None,
                    help="schedule of the fused z stage (with --batch; default: the "
                         "library's default for the grid)")
    ap.add_argument("--mode", default="fused", choices=["fused", "fused_rt", "batched", "plain", "auto"])
    ap.add_argument("--workload", default="search128")
    ap.add_argument("--plain-rotations", type=int, default=256,
                    help="rotations timed on the plain one-rotation-per-launch path")
    ap.add_argument("--parity-rotations", type=int, default=4,
                    help="rotations in the FP64-oracle parity sample (0 = skip)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run a 256 MiB device-to-device copy first (calibrates HBM byte "
                         "counters under a profiler)")
    ap.add_argument("--cpu-per-worker", type=int, default=64,
                    help="rotations per CPU worker in the CPU baseline sample: the first "
                         "64 P rotations with P workers (0 = skip)")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "split"],
                    help="FFT arithmetic of the measured path: fp32, or split (first radix-16 "
                         "stage of the unfused y/z passes as an FP16 hi+lo GEMM on MFMA)")
    ap.add_argument("--precision-report", action="store_true",
                    help="also compare FP32 and split-precision (MFMA first stage) FFT passes "
                         "on --precision-workload against the FP64 oracle")
    ap.add_argument("--precision-workload", default="prec192")
    ap.add_argument("--precision-rotations", type=int, default=32)
    ap.add_argument("--precision-batch", type=int, default=8)
    ap.add_argument("--serving", action="store_true",
                    help="measure the serving configuration (Poisson stream of mixed-size fits) "
                         "instead of the search workload")
    ap.add_argument("--serve-requests", type=int, default=40)
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--serve-max-active", type=int, default=8)
    ap.add_argument("--serve-depth", type=int, default=2)
    ap.add_argument("--serve-check", type=int, default=16,
                    help="serve the stream again with every rotation set cut to this prefix and "
                         "check each served map against the FP64 oracle (0 = skip)")
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        return 2

    if args.serving:
        return serving_bench(args)

    import torch
    from powerfit_amd import HipCorrelator, _lib
    from powerfit_amd.rotations import rotations_for_angle
    from powerfit_amd.synth import WORKLOADS, make_case

    if args.calib_copy:
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(2):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    w = WORKLOADS[args.workload]
    case = make_case(w)
    R = rotations_for_angle(w.angle)
    K = R.shape[0]
    nz, ny, nx = case.target.shape
    V = nz * ny * nx
    H = nz * ny * (nx // 2 + 1)

    corr = HipCorrelator(case.target.astype(np.float64), laplace=case.laplace, device=0,
                         mode=args.mode, batch=max(1, args.batch), precision=args.precision)
    corr.template = case.template
    corr.mask = case.mask
    corr.rotations = R
    tuning = None
    if args.mode == "plain":
        B = 1
    elif args.batch > 0:
        B = args.batch
        if args.chunks:
            cz, cl = (int(v) for v in args.chunks.split(","))
            corr.set_tuning(cz, cl)
        if args.z_schedule is not None:
            corr.set_schedule(args.z_schedule)
    else:
        from powerfit_amd.autotune import autotune
        tuning = autotune(corr, use_cache=False)
        B = corr.batch
    steps = args.steps if args.steps is not None else max(1, K // B)

    # warmup
    corr.reset()
    run_scan(corr, K, B, args.warmup)
    corr.sync()
    torch.cuda.synchronize()

    # timed region
    corr.reset()
    corr.profile(True)
    corr.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run_scan(corr, K, B, steps)
    corr.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ktimes = corr.kernel_times()
    corr.profile(False)
    rot_done = steps * B
    value = rot_done / elapsed

    # dominant kernel roofline: algorithmic bytes of every launch of the
    # kernel in the timed region (per rotation x rotations run, plus the
    # per-launch part x launches) over the kernel's summed event time
    dom = max(ktimes, key=lambda k: ktimes[k][0])
    dom_ms, dom_n = ktimes[dom]
    binary = bool(np.all((case.mask == 0) | (case.mask == 1)))
    per_rot, per_launch = alg_bytes(dom, V, H, binary)
    dom_bytes = per_rot * rot_done + per_launch * dom_n
    achieved = dom_bytes / (dom_ms / 1e3) / 1e9
    path_gbs = value * path_bytes_per_rotation(V, H) / 1e9
    copy_gbs = device_copy_gbs(torch)

    # plain one-rotation-per-launch path on the same inputs
    plain_value = None
    if args.mode != "plain" and args.plain_rotations > 0:
        plain = HipCorrelator(case.target.astype(np.float64), laplace=case.laplace, device=0,
                              mode="plain", batch=1)
        plain.template = case.template
        plain.mask = case.mask
        plain.rotations = R
        plain.reset()
        run_scan(plain, K, 1, 4)
        plain.sync()
        t = time.perf_counter()
        run_scan(plain, K, 1, args.plain_rotations)
        plain.sync()
        plain_value = args.plain_rotations / (time.perf_counter() - t)
        del plain

    # parity on a sample: the FP64 oracle (CPU leg) vs the same path on the same rotations
    parity = None
    if args.parity_rotations > 0:
        best_full = np.zeros(case.target.shape, np.float32)
        rot_full = np.zeros(case.target.shape, np.int32)
        corr.reset()
        run_scan(corr, K, B, max(1, K // B))
        corr.sync()
        _lib.check(_lib.load().pf_get_result(corr._ctx, _lib.ptr(best_full, _lib.FP),
                                             _lib.ptr(rot_full, _lib.I32P)))
        peak_rot = int(rot_full.reshape(-1)[np.argmax(best_full)])
        pick = sorted(set(range(min(args.parity_rotations - 1, K))) | {peak_rot})
        Rs = R[pick]
        sample = HipCorrelator(case.target.astype(np.float64), laplace=case.laplace, device=0,
                               mode=args.mode, batch=max(1, min(B, len(pick))),
                               precision=args.precision)
        sample.template = case.template
        sample.mask = case.mask
        sample.rotations = Rs
        sample.scan()
        parity = {"rotations": pick,
                  **cpu_baseline_check(case, Rs, sample.lcc, sample.rot.astype(np.int64)),
                  "full_scan_peak_lcc": float(best_full.max())}
        del sample

    cpu = None
    if args.cpu_per_worker > 0:
        nproc = cpu_workers()
        done, secs = cpu_baseline(case, R, nproc, args.cpu_per_worker)
        cpu = {"value": done / secs, "unit": "rotations/s", "cores": nproc, "kind": "port",
               "cpu_model": cpu_model(),
               "sample": f"prefix-timed: first {done} rotations ({args.cpu_per_worker} per "
                         f"worker) of the {K}-rotation set, {nproc} forked FP64 numpy workers "
                         f"with the reference's contiguous chunking, setup excluded "
                         f"({secs:.1f} s wall)"}

    out = {
        "metric": "rotations/sec (LCC search, 128^3 grid) + HBM roofline %, chosen vs plain FP32",
        "value": value,
        "unit": "rotations/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded chain template, noisy target; no network for EM maps)",
        "config": {"workload": f"{w.name}: {nz}x{ny}x{nx} f32 target, {w.voxelspacing} A voxels, "
                               f"{w.atoms}-atom chain at {w.resolution} A, {w.angle} deg "
                               f"rotation set (K={K})",
                   "grid": [nz, ny, nx], "rotations": K, "batch": B, "mode": args.mode,
                   "autotuned": None if tuning is None else
                   {k: tuning.get(k) for k in ("batch", "chunk_z", "chunk_lcc", "z_schedule",
                                                "trials")},
                   "laplace": w.laplace, "core_weighted": w.core_weighted,
                   "precision": args.precision},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "launches": dom_n, "rotations": rot_done,
                     "kernel_ms_per_launch": dom_ms / dom_n,
                     "alg_bytes_per_rotation": per_rot, "alg_bytes_per_launch_fixed": per_launch,
                     "alg_bytes_per_launch": dom_bytes / dom_n,
                     "traffic": measured_traffic(args.workload, args.mode, dom, rot_done / dom_n,
                                                 binary),
                     "device_copy_gbs": copy_gbs, "frac_of_device_copy": achieved / copy_gbs},
        "path_roofline": {"bytes_per_rotation": path_bytes_per_rotation(V, H),
                          "achieved": path_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": path_gbs / HBM_PEAK_GBS},
        "kernels_ms_per_rotation": {k: ktimes[k][0] / rot_done for k in ktimes
                                    if ktimes[k][1]},
        "plain_value": plain_value,
        "speedup_vs_plain": (value / plain_value) if plain_value else None,
        "cpu_baseline": cpu,
        "parity": parity,
    }
    if args.precision_report:
        pw = WORKLOADS[args.precision_workload]
        out["precision"] = precision_report(args, HipCorrelator, make_case(pw),
                                            rotations_for_angle(pw.angle), pw)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())

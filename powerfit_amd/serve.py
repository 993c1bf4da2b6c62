"""Continuous-batching request loop for LCC searches on one GPU.

Each admitted fit gets its own correlator context: its own HIP stream, its
target spectra resident in HBM for the whole fit.  Requests are admitted as
slots and memory free up — a fit is admitted when the bytes its context will
allocate (``pf_ctx_bytes``) fit in the remaining HBM budget and a slot is
free — so fits of mixed grid sizes share the device.  The loop interleaves
the active fits batch by batch: each keeps up to ``depth`` batches in flight
on its stream, and a progress mark after every batch (``pf_ctx_mark``) lets
the loop queue the next batch as soon as one finishes, so a fit's stream
never runs dry while it has rotations left.  Work from several fits overlaps
on the device, and a request arriving mid-way starts as soon as it is
admitted instead of waiting for the queue to drain.  A fit's result is read back when its stream is idle; its
context is then released and the memory returned to the budget — or, when
the request names its map (``target_key``), kept resident with that map's
spectra: a later request on the same map takes the idle context and only
loads its template, mask and rotations.  Idle contexts count against the
budget and are released, least recently used first, when a new fit needs the
memory.

Each fit's result is computed exactly as in a standalone search: its
rotations go through the same kernels, and the per-voxel merge is an
order-independent max, so a result is identical alone and in the mix.

    python -m powerfit_amd.serve --requests 12 --rate 4

prints one JSON line with throughput and p50/p99 latency.
"""
import argparse
import ctypes
import json
import time
from collections import OrderedDict
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .correlator import HipCorrelator


@dataclass
class FitRequest:
    """One fit: map, template, mask, rotation set, options, arrival time (s)."""
    rid: int
    target: np.ndarray
    template: np.ndarray
    mask: np.ndarray
    rotations: np.ndarray
    laplace: bool = False
    arrival: float = 0.0
    target_key: object = None   # equal keys: the same map (its spectra can stay resident)


@dataclass
class FitResult:
    rid: int
    lcc: np.ndarray
    rot: np.ndarray
    arrival: float
    admitted: float
    finished: float
    rotations: int
    shape: tuple

    @property
    def latency(self):
        return self.finished - self.arrival


@dataclass
class _Active:
    req: FitRequest
    corr: HipCorrelator
    bytes: int
    admitted: float
    batch: int
    next_rot: int = 0
    queued: list = field(default_factory=list)   # progress tickets of batches in flight


def ctx_bytes(shape, batch, mode="auto", flags=0, nrot=0):
    """Peak device bytes of a context for `shape` (admission control)."""
    nz, ny, nx = shape
    b = _lib.load().pf_ctx_bytes(nz, ny, nx, int(batch), _lib.MODES[mode], int(flags), int(nrot))
    if b < 0:
        _lib.check(int(b), "pf_ctx_bytes")
    return int(b)


def fit_batch(shape, nrot, budget, max_active, mode="auto", cap=256):
    """Rotations per launch for one fit: the largest power of two up to `cap`
    (256: the measured best or within 1.5 % of it at every fused size) whose
    context fits an equal share (budget / max_active) of the HBM budget, and
    no more than the fit's rotations."""
    share = budget // max(1, max_active)
    b = cap
    while b > 1 and ctx_bytes(shape, b, mode, 0, nrot) > share:
        b //= 2
    return max(1, min(b, int(nrot)))


class FitServer:
    """Admit, interleave and complete LCC searches on one device."""

    def __init__(self, device=0, hbm_budget=64 << 30, max_active=8, batch=0, depth=2,
                 mode="auto", keep_targets=True):
        """batch: rotations per launch for every fit, or 0 to pick per fit
        (``fit_batch``: up to 256 within an equal share of the budget)."""
        self.device = device
        self.keep_targets = keep_targets   # reuse idle contexts of requests' named maps
        self.hbm_budget = int(hbm_budget)
        self.max_active = max_active
        self.batch = batch
        self.depth = depth
        self.mode = mode
        self.used = 0
        self.active = []
        self.waiting = []
        self.results = {}
        self.idle = OrderedDict()   # (target_key, shape, laplace) -> [(corr, bytes)], LRU first
        self.reused = 0

    def warmup(self):
        """Load the library and start the HIP runtime on the device (a
        long-running server does this once, before its first request)."""
        lib = _lib.load()
        count = ctypes.c_int(0)
        _lib.check(lib.pf_device_count(ctypes.byref(count)), "pf_device_count")
        if self.device >= count.value:
            raise RuntimeError(f"device {self.device} not present ({count.value} visible)")

    def _key(self, req):
        if req.target_key is None or not self.keep_targets:
            return None
        return (req.target_key, tuple(req.target.shape), bool(req.laplace))

    def _take_idle(self, key):
        if key is None or key not in self.idle:
            return None
        ctxs = self.idle[key]
        held = ctxs.pop()
        if not ctxs:
            del self.idle[key]
        return held

    def _evict_idle(self, need):
        """Release idle contexts, least recently used first, until `need`
        more bytes fit in the budget (or none is left)."""
        while self.idle and self.used + need > self.hbm_budget:
            key, ctxs = next(iter(self.idle.items()))
            corr, nbytes = ctxs.pop(0)
            if not ctxs:
                del self.idle[key]
            self.used -= nbytes
            del corr

    def release_idle(self):
        """Release every idle context (their memory returns to the budget)."""
        self._evict_idle(float("inf"))

    def _batch_for(self, req):
        if self.batch > 0:
            return min(self.batch, req.rotations.shape[0])
        return fit_batch(req.target.shape, req.rotations.shape[0], self.hbm_budget,
                         self.max_active, self.mode)

    def _admit(self, now):
        while self.waiting and len(self.active) < self.max_active:
            req = self.waiting[0]
            nrot = req.rotations.shape[0]
            held = self._take_idle(self._key(req))
            if held is not None:
                corr, need = held           # its bytes are already counted
                batch = corr.batch
                grown = ctx_bytes(req.target.shape, batch, self.mode, 0, nrot)
                if grown > need:            # a longer rotation set: count the difference
                    self.used += grown - need
                    need = grown
                self.reused += 1
            else:
                batch = self._batch_for(req)
                need = ctx_bytes(req.target.shape, batch, self.mode, 0, nrot)
                if need > self.hbm_budget:
                    raise MemoryError(f"request {req.rid} needs {need} bytes, over the budget")
                self._evict_idle(need)
                if self.used + need > self.hbm_budget:
                    break
                corr = HipCorrelator(np.asarray(req.target, np.float64), laplace=req.laplace,
                                     device=self.device, mode=self.mode, batch=batch)
                self.used += need
            self.waiting.pop(0)
            corr.template = req.template
            corr.mask = req.mask
            corr.rotations = req.rotations
            corr.reset()
            self.active.append(_Active(req, corr, need, now, batch))

    def _pump(self):
        """Retire finished batches, then top every active fit up to `depth`
        batches in flight."""
        for a in self.active:
            while a.queued and a.corr.reached(a.queued[0]):
                a.queued.pop(0)
            K = a.req.rotations.shape[0]
            while a.next_rot < K and len(a.queued) < self.depth:
                n = min(a.batch, K - a.next_rot)
                a.corr.enqueue_scan(a.next_rot, n)
                a.next_rot += n
                a.queued.append(a.corr.mark())

    def _complete(self, now_fn):
        still = []
        for a in self.active:
            if a.next_rot >= a.req.rotations.shape[0] and a.corr.done():
                lcc, rot = a.corr.result()
                self.results[a.req.rid] = FitResult(
                    a.req.rid, lcc.copy(), rot.copy(), a.req.arrival, a.admitted, now_fn(),
                    a.req.rotations.shape[0], a.req.target.shape)
                key = self._key(a.req)
                if key is None:
                    self.used -= a.bytes
                    del a.corr
                else:                       # stays resident for the next fit on this map
                    self.idle.setdefault(key, []).append((a.corr, a.bytes))
                    self.idle.move_to_end(key)
            else:
                still.append(a)
        self.active = still

    def run(self, requests, realtime=True):
        """Serve `requests` (arrival times in seconds from the start when
        realtime, else all at once).  Returns {rid: FitResult}."""
        pending = sorted(requests, key=lambda r: r.arrival)
        self.warmup()
        t0 = time.perf_counter()

        def now():
            return time.perf_counter() - t0

        if not realtime:
            for r in pending:
                r.arrival = 0.0
        while pending or self.waiting or self.active:
            t = now()
            while pending and pending[0].arrival <= t:
                self.waiting.append(pending.pop(0))
            self._admit(t)
            self._pump()
            self._complete(now)
            if not self.active and not self.waiting and pending:
                time.sleep(max(0.0, min(0.01, pending[0].arrival - now())))
            elif self.active and all(a.next_rot >= a.req.rotations.shape[0] or
                                     len(a.queued) >= self.depth for a in self.active):
                time.sleep(0.0001)
        return self.results


def make_stream(n, rate, seed=5003, sizes=None, angles=(10.0, 20.0), maps=0,
                max_rotations=None, core_weighted=()):
    """Poisson stream of synthetic fit requests (500-20,000-atom chains into
    48^3-160^3 maps at 4-12 A, 10 or 20 degree rotation sets; arrivals and
    draws from `seed`).  With maps > 0 the requests draw from that many shared
    maps (target_key set), as when several fits run against the same cryo-EM
    map.  `sizes` as a list is drawn from; as a tuple of length n it is taken
    in order.  max_rotations truncates each rotation set to a prefix; requests
    whose index is in `core_weighted` use a core-weighted mask."""
    from .rotations import rotations_for_angle
    from .synth import Workload, make_case
    rng = np.random.default_rng(seed)
    sizes = sizes or [48, 64, 72, 80, 96, 120, 128, 144, 160]
    fixed = isinstance(sizes, tuple) and len(sizes) == n

    def draw_case(tag, k, i=None):
        N = int(sizes[i]) if fixed and i is not None else int(rng.choice(sizes))
        res = float(rng.uniform(4.0, 12.0))
        atoms = int(rng.integers(500, 20001))
        return make_case(Workload(tag, N, res / 4.0, atoms, res, 10.0,
                                  core_weighted=i in core_weighted, seed=seed * 10 + k * 7))

    shared = [draw_case(f"map{j}", 100000 + j) for j in range(maps)]
    t = 0.0
    reqs = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        if maps:
            j = int(rng.integers(0, maps))
            c, key = shared[j], f"map{j}"
        else:
            c, key = draw_case(f"req{i}", i, i), None
        angle = float(rng.choice(angles))
        R = rotations_for_angle(angle)
        if max_rotations:
            R = R[:max_rotations]
        reqs.append(FitRequest(i, c.target, c.template, c.mask, R, c.laplace, t, key))
    return reqs


def summarize(results, wall, rate):
    """Throughput, p50/p99 latency overall and per grid size."""
    res = list(results.values())
    lat = np.array([r.latency for r in res])
    rots = sum(r.rotations for r in res)
    per_size = {}
    for n in sorted({r.shape[0] for r in res}):
        ls = np.array([r.latency for r in res if r.shape[0] == n])
        per_size[str(n)] = {"requests": int(ls.size), "p50_s": float(np.percentile(ls, 50)),
                            "p99_s": float(np.percentile(ls, 99))}
    return {"requests": len(res), "wall_s": wall, "rotations": rots,
            "rotations_per_s": rots / wall, "requests_per_s": len(res) / wall,
            "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)),
            "per_size": per_size, "arrival_rate": rate}


def main(argv=None):
    ap = argparse.ArgumentParser(description="serve a Poisson stream of LCC fits")
    ap.add_argument("--requests", type=int, default=12)
    ap.add_argument("--rate", type=float, default=4.0)
    ap.add_argument("--seed", type=int, default=5003)
    ap.add_argument("--batch", type=int, default=0,
                    help="rotations per launch (0: per fit, up to 256 within its budget share)")
    ap.add_argument("--depth", type=int, default=2, help="batches in flight per fit")
    ap.add_argument("--max-active", type=int, default=8)
    ap.add_argument("--budget-gib", type=float, default=64.0)
    ap.add_argument("--max-rotations", type=int, default=0,
                    help="truncate every rotation set to this prefix (0: whole sets)")
    ap.add_argument("--maps", type=int, default=0,
                    help="draw the requests from this many shared maps (0: one map per request)")
    ap.add_argument("--no-keep-targets", action="store_true",
                    help="release every context after its fit (no resident map spectra)")
    ap.add_argument("--detail", action="store_true",
                    help="also print one JSON line per request (size, rotations, times)")
    args = ap.parse_args(argv)
    reqs = make_stream(args.requests, args.rate, args.seed, maps=args.maps,
                       max_rotations=args.max_rotations or None)
    server = FitServer(hbm_budget=int(args.budget_gib * (1 << 30)), max_active=args.max_active,
                       batch=args.batch, depth=args.depth,
                       keep_targets=not args.no_keep_targets)
    t = time.perf_counter()
    res = server.run(reqs)
    wall = time.perf_counter() - t
    if args.detail:
        for r in sorted(res.values(), key=lambda r: r.rid):
            print(json.dumps({"request": r.rid, "n": r.shape[0], "rotations": r.rotations,
                              "arrival_s": round(r.arrival, 4),
                              "queued_s": round(r.admitted - r.arrival, 4),
                              "latency_s": round(r.latency, 4)}))
    out = summarize(res, wall, args.rate)
    out.update({"maps": args.maps, "contexts_reused": server.reused, "depth": args.depth,
                "batch": args.batch or "per fit"})
    print(json.dumps(out))


if __name__ == "__main__":
    main()

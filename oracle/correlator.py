"""ORACLE (test infrastructure): FP64 restatement of the CPU LCC search.

Follows ``src/powerfit_em/powerfitter.py``:

* ``prepare_target``      ``BaseCorrelator.__init__``/``_get_lcc_mask``
                          (:169-180) and the target spectra of
                          ``CPUCorrelator.__init__`` (:262-277).
* ``prepare_template``    ``mask`` setter + ``_laplace_filter`` +
                          ``_normalize_template`` (:190-220).
* ``laplace_wrap``        ``scipy.ndimage.laplace(mode='wrap')`` restated as
                          the closed form sum over axes of
                          roll(+1) + roll(-1) - 2x (pinned against scipy).
* ``lcc_for_rotation``    ``_translational_scan``/``_get_lcc``
                          (:335-373, FFTs :375-393), with ``conj_multiply``
                          (``_powerfit.pyx:45-53``) and ``calc_lcc``
                          (``_powerfit.pyx:56-72``) inline.  FFTs are
                          ``numpy.fft.rfftn/irfftn``: the reference's own CPU
                          fallback (:19-20, 312-315), so the library is the
                          oracle.
* ``scan``                ``CPUCorrelator.scan`` (:317-333): strict ``>``
                          take-best from 0/0, plus the runner-up LCC per voxel
                          so that "index identical where the winner's margin
                          exceeds the tolerance" is decidable.
* ``parallel_scan``       ``PowerFitter._cpu_scan``/``_combine`` (:94-163):
                          P forked workers over contiguous chunks of K//P
                          rotations (remainder to the last), merged in job
                          order with strict ``>`` and index offsets.

Parity pins: tests/test_oracle_correlator.py (perfect-fit LCC = 1 at flat
index 0, ``tests/test_powerfitter.py:67-79``; normalisation mean 0 / std 1;
brute-force direct-space LCC over all cyclic translations on tiny grids;
closed forms of conj-multiply and calc_lcc, ``tests/test__powerfit.py``;
chunked merge equals the single-process scan).
"""
import multiprocessing as mp
import os

import numpy as np

from .rotate import rotate_grid3d


def laplace_wrap(a):
    """Discrete Laplacian with periodic boundaries (scipy ``laplace`` wrap)."""
    out = np.zeros_like(a, dtype=np.float64)
    for ax in range(a.ndim):
        out += np.roll(a, 1, axis=ax) + np.roll(a, -1, axis=ax) - 2.0 * a
    return out


def prepare_target(target, laplace=False):
    """Return (T, lcc_mask, rmax, Tf, F1, F2) as ``BaseCorrelator``/``CPUCorrelator``."""
    target = np.asarray(target, dtype=np.float64)
    T = target / target.max()
    lcc_mask = (T > T.max() * 0.05).astype(np.uint8)
    rmax = min(target.shape) // 2
    Tf = laplace_wrap(T) if laplace else T
    F1 = np.fft.rfftn(Tf)
    F2 = np.fft.rfftn(Tf ** 2)
    return T, lcc_mask, rmax, Tf, F1, F2


def prepare_template(template, mask, laplace=False):
    """Return (P, W, N) as the ``template``/``mask`` setters leave them."""
    P = np.array(template, dtype=np.float64, copy=True)
    W = np.array(mask, dtype=np.float64, copy=True)
    if P.shape != W.shape:
        raise ValueError("Shape of the mask is different from target.")
    ind = W != 0
    N = int(ind.sum())
    if N == 0:
        raise ValueError("Zero-filled mask is not allowed.")
    if laplace:
        P = laplace_wrap(P)
    P *= W
    P[ind] -= P[ind].mean()
    P[ind] /= P[ind].std()
    P *= W
    return P, W, N


def lcc_from_rotated(rot_template, rot_mask, N, F1, F2, lcc_mask):
    """``_get_lcc`` (:349-373): LCC map from already rotated template and mask."""
    shape = rot_template.shape
    rot_mask2 = rot_mask * rot_mask
    ft_t = np.fft.rfftn(rot_template)
    ft_m = np.fft.rfftn(rot_mask)
    ft_m2 = np.fft.rfftn(rot_mask2)
    axes = tuple(range(len(shape)))
    gcc = np.fft.irfftn(np.conj(ft_t) * F1, s=shape, axes=axes)
    ave = np.fft.irfftn(np.conj(ft_m) * F1, s=shape, axes=axes)
    ave2 = np.fft.irfftn(np.conj(ft_m2) * F2, s=shape, axes=axes)
    ave2 *= N
    lcc = np.zeros(shape)
    m = np.asarray(lcc_mask).astype(bool)
    with np.errstate(divide="ignore", invalid="ignore"):
        lcc[m] = gcc[m] / np.sqrt(ave2[m] - ave[m] ** 2)
    return lcc


def rotate_pair(P, W, rotmat, rmax):
    """``_rotate_grids`` (:339-347): template trilinear, mask nearest, into zeros."""
    rot_template = np.zeros(P.shape)
    rot_mask = np.zeros(W.shape)
    rotate_grid3d(P, rotmat, rmax, rot_template, False)
    rotate_grid3d(W, rotmat, rmax, rot_mask, True)
    return rot_template, rot_mask


def lcc_for_rotation(P, W, N, rotmat, rmax, F1, F2, lcc_mask):
    """Per-voxel LCC map for one rotation (zeros where ``lcc_mask`` is 0)."""
    rot_template, rot_mask = rotate_pair(P, W, rotmat, rmax)
    return lcc_from_rotated(rot_template, rot_mask, N, F1, F2, lcc_mask)


class ScanResult:
    """best LCC, rotation index (as float64 like the CPU reference) and runner-up."""

    def __init__(self, best, rot, second):
        self.best = best
        self.rot = rot
        self.second = second

    def margin(self):
        return self.best - self.second


def scan(target, template, mask, rotations, laplace=False, first_index=0):
    """``CPUCorrelator.scan`` over ``rotations`` with runner-up tracking.

    The runner-up of a voxel is the best LCC of every candidate other than the
    winner, where the initial state (LCC 0, index 0) counts as a candidate.
    """
    T, lcc_mask, rmax, Tf, F1, F2 = prepare_target(target, laplace)
    P, W, N = prepare_template(template, mask, laplace)
    rotations = np.asarray(rotations, dtype=np.float64).reshape(-1, 3, 3)
    best = np.zeros(T.shape)
    rot = np.zeros(T.shape)
    second = np.full(T.shape, -np.inf)
    for n in range(rotations.shape[0]):
        s = lcc_for_rotation(P, W, N, rotations[n], rmax, F1, F2, lcc_mask)
        up = s > best
        rest = ~up & (s > second)
        second[up] = best[up]
        second[rest] = s[rest]
        best[up] = s[up]
        rot[up] = n + first_index
    return ScanResult(best, rot, second)


def _worker(args):
    target, template, mask, rotations, laplace, offset = args
    r = scan(target, template, mask, rotations, laplace, first_index=offset)
    return r.best, r.rot, r.second


def chunks(nrot, nproc):
    """The reference's contiguous split: K//P each, remainder to the last."""
    per = nrot // nproc
    out = []
    for n in range(nproc):
        lo = n * per
        hi = nrot if n == nproc - 1 else lo + per
        out.append((lo, hi))
    return out


def merge(parts):
    """``_combine``: strict ``>`` in job order; runner-up merged alongside."""
    best = np.zeros_like(parts[0][0])
    rot = np.zeros_like(parts[0][1])
    second = np.full_like(parts[0][0], -np.inf)
    for pb, pr, ps in parts:
        up = pb > best
        # runner-up of the union: the larger of the loser's best and both runner-ups
        new_second = np.where(up, np.maximum(best, ps), np.maximum(second, pb))
        best = np.where(up, pb, best)
        rot = np.where(up, pr, rot)
        second = new_second
    return ScanResult(best, rot, second)


def parallel_scan(target, template, mask, rotations, laplace=False, nproc=None):
    """``PowerFitter._cpu_scan``: fork ``nproc`` workers over contiguous chunks."""
    rotations = np.asarray(rotations, dtype=np.float64).reshape(-1, 3, 3)
    nproc = nproc or os.cpu_count()
    nproc = max(1, min(nproc, rotations.shape[0]))
    jobs = [(target, template, mask, rotations[lo:hi], laplace, lo)
            for lo, hi in chunks(rotations.shape[0], nproc)]
    if nproc == 1:
        parts = [_worker(jobs[0])]
    else:
        ctx = mp.get_context("fork")
        with ctx.Pool(nproc) as pool:
            parts = pool.map(_worker, jobs)
    return merge(parts)


def _timed_worker(args, barrier, queue):
    target, template, mask, rotations, laplace, offset = args
    T, lcc_mask, rmax, Tf, F1, F2 = prepare_target(target, laplace)
    P, W, N = prepare_template(template, mask, laplace)
    best = np.zeros(T.shape)
    rot = np.zeros(T.shape)
    barrier.wait()
    import time
    t0 = time.perf_counter()
    for n in range(rotations.shape[0]):
        s = lcc_for_rotation(P, W, N, rotations[n], rmax, F1, F2, lcc_mask)
        up = s > best
        best[up] = s[up]
        rot[up] = n + offset
    t1 = time.perf_counter()
    queue.put((t0, t1, rotations.shape[0]))


def timed_parallel_scan(target, template, mask, rotations, laplace=False, nproc=None):
    """Wall time of the rotation loop of ``PowerFitter._cpu_scan`` (setup excluded).

    ``nproc`` forked workers take the reference's contiguous chunks, build
    their spectra, meet at a barrier, then scan; returns (rotations, seconds)
    from the first loop start to the last loop end.
    """
    rotations = np.asarray(rotations, dtype=np.float64).reshape(-1, 3, 3)
    nproc = max(1, min(nproc or os.cpu_count(), rotations.shape[0]))
    ctx = mp.get_context("fork")
    barrier = ctx.Barrier(nproc)
    queue = ctx.Queue()
    procs = []
    for lo, hi in chunks(rotations.shape[0], nproc):
        args = (target, template, mask, rotations[lo:hi], laplace, lo)
        p = ctx.Process(target=_timed_worker, args=(args, barrier, queue))
        p.start()
        procs.append(p)
    res = [queue.get() for _ in procs]
    for p in procs:
        p.join()
        if p.exitcode != 0:
            raise RuntimeError(f"oracle worker failed with exit code {p.exitcode}")
    t0 = min(r[0] for r in res)
    t1 = max(r[1] for r in res)
    return sum(r[2] for r in res), t1 - t0

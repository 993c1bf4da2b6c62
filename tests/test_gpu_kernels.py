"""GPU parity of the kernel-level C-ABI drop-ins against the oracle / closed forms."""
import ctypes
import os
import sys

import numpy as np
import pytest

from oracle.rotate import rotate_grid3d as oracle_rotate
from powerfit_amd import _lib
from powerfit_amd.rotations import load_set, quat_to_rotmat

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return torch


def hip_rotate(hip, grid, R, radius, out, nearest):
    grid = np.ascontiguousarray(grid, dtype=np.float64)
    R = np.ascontiguousarray(R, dtype=np.float64)
    _lib.check(hip.pf_rotate_grid3d(_lib.ptr(grid, _lib.DP), *grid.shape, _lib.ptr(R, _lib.DP),
                                    int(radius), _lib.ptr(out, _lib.DP), *out.shape,
                                    int(nearest)), "pf_rotate_grid3d")


def test_rotate_known_answers(hip):
    say("rotate known answers (4,5,6) r=2 and (5,6,7) r=3")
    grid = np.zeros((4, 5, 6))
    for idx in [(0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 0, 2), (0, 0, -1), (-1, 0, 0)]:
        grid[idx] = 1
    out = np.zeros_like(grid)
    hip_rotate(hip, grid, np.eye(3), 2, out, True)
    assert np.array_equal(out, grid)
    out.fill(0)
    hip_rotate(hip, grid, np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1.0]]), 2, out, False)
    answer = np.zeros_like(out)
    for idx in [(0, 0, 0), (0, 1, 0), (0, 1, -1), (0, 2, 0), (0, -1, 0), (-1, 0, 0)]:
        answer[idx] = 1
    assert np.allclose(answer, out)
    t = np.zeros((5, 6, 7))
    t[0, 0, 0:3] = 1
    out = np.zeros_like(t)
    hip_rotate(hip, t, np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1.0]]), 3, out, False)
    answer = np.zeros_like(t)
    answer[0, :3, 0] = 1
    assert np.allclose(out, answer)


def test_rotate_leaves_outside_untouched(hip):
    say("rotate: voxels outside the sphere keep their value")
    g = np.random.default_rng(0).standard_normal((10, 10, 10))
    out = np.full_like(g, 7.0)
    hip_rotate(hip, g, np.eye(3), 3, out, False)
    ref = np.full_like(g, 7.0)
    oracle_rotate(g, np.eye(3), 3, ref, False)
    assert np.array_equal(out, ref)


def test_rotate_golden_vectors(hip):
    """Against vectors written by the reference's compiled _extensions.c."""
    d = np.load(os.path.join(HERE, "golden", "rotate_ref.npz"))
    for c in range(int(d["ncases"])):
        say(f"rotate golden case {c}")
        ref = d[f"case{c}_out"]
        out = np.zeros(ref.shape)
        hip_rotate(hip, d[f"case{c}_grid"], d[f"case{c}_rot"], int(d[f"case{c}_radius"]), out,
                   int(d[f"case{c}_nearest"]))
        if int(d[f"case{c}_nearest"]):
            assert np.array_equal(out, ref), c
        else:
            assert np.allclose(out, ref, rtol=0, atol=1e-13), c


def test_rotate_bit_exact_vs_oracle(hip):
    """Same FP64 evaluation order, no contraction: identical to the oracle."""
    rng = np.random.default_rng(1)
    for setname in ("c48n309.npy", "c600v.npy", "c48u1.npy"):
        R = quat_to_rotmat(load_set(setname)[:, :4])
        for shape in [(16, 16, 16), (17, 18, 19), (12, 20, 14), (32, 32, 32)]:
            g = rng.standard_normal(shape)
            m = rng.integers(0, 5, shape).astype(float)
            r = min(shape) // 2
            for k in rng.choice(len(R), min(8, len(R)), replace=False):
                say(f"rotate vs oracle {setname} shape={shape} k={k}")
                for nearest, src in ((0, g), (1, m)):
                    a = np.zeros(shape)
                    b = np.zeros(shape)
                    hip_rotate(hip, src, R[k], r, a, nearest)
                    oracle_rotate(src, R[k], r, b, nearest)
                    assert np.array_equal(a, b), (setname, shape, k, nearest)


def test_rotate_rejects_bad_radius(hip):
    g = np.zeros((6, 6, 6))
    out = np.zeros_like(g)
    rc = hip.pf_rotate_grid3d(_lib.ptr(g, _lib.DP), 6, 6, 6, _lib.ptr(np.eye(3), _lib.DP), 5,
                              _lib.ptr(out, _lib.DP), 6, 6, 6, 0)
    assert rc == -1


FFT_SHAPES = [(8, 8, 8), (5, 6, 7), (4, 5, 6), (1, 1, 8), (3, 1, 1), (2, 3, 2), (16, 12, 10),
              (9, 15, 25), (7, 11, 13), (24, 20, 18), (48, 48, 48), (64, 64, 64),
              (96, 96, 96), (30, 42, 98), (128, 128, 128)]


@pytest.mark.parametrize("shape", FFT_SHAPES)
def test_rfftn_irfftn_vs_numpy(hip, shape):
    torch = torch_dev()
    say(f"fft shape={shape}")
    rng = np.random.default_rng(hash(shape) % 2**32)
    batch = 2
    x = rng.standard_normal((batch,) + shape).astype(np.float32)
    ref = np.fft.rfftn(x.astype(np.float64), axes=(1, 2, 3))
    dx = torch.from_numpy(x).cuda()
    dX = torch.zeros(ref.shape, dtype=torch.complex64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    _lib.check(hip.pf_rfftn_dev(dx.data_ptr(), dX.data_ptr(), *shape, batch, stream), "rfftn")
    torch.cuda.synchronize()
    X = dX.cpu().numpy()
    scale = np.abs(ref).max()
    assert np.abs(X - ref).max() / scale < 2e-6
    # inverse of an arbitrary (not Hermitian-consistent) half spectrum: numpy's
    # irfftn semantics (C2C over z, y; C2R over x ignoring DC/Nyquist imag)
    Y = (rng.standard_normal(ref.shape) + 1j * rng.standard_normal(ref.shape)).astype(np.complex64)
    yref = np.fft.irfftn(Y.astype(np.complex128), s=shape, axes=(1, 2, 3))
    dY = torch.from_numpy(Y).cuda()
    dy = torch.zeros((batch,) + shape, dtype=torch.float32, device="cuda")
    _lib.check(hip.pf_irfftn_dev(dY.data_ptr(), dy.data_ptr(), *shape, batch, stream), "irfftn")
    torch.cuda.synchronize()
    y = dy.cpu().numpy()
    assert np.abs(y - yref).max() / np.abs(yref).max() < 2e-6


def test_conj_multiply(hip):
    torch = torch_dev()
    rng = np.random.default_rng(2)
    n = 100003
    a = (rng.random(n) + 1j * rng.random(n)).astype(np.complex64)
    b = (rng.random(n) + 1j * rng.random(n)).astype(np.complex64)
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    do = torch.zeros(n, dtype=torch.complex64, device="cuda")
    _lib.check(hip.pf_conj_multiply_dev(da.data_ptr(), db.data_ptr(), do.data_ptr(), n,
                                        torch.cuda.current_stream().cuda_stream), "conj")
    torch.cuda.synchronize()
    assert np.allclose(do.cpu().numpy(), a.conj() * b, rtol=1e-6, atol=1e-6)


def test_calc_lcc(hip):
    torch = torch_dev()
    rng = np.random.default_rng(3)
    n = 50001
    gcc = rng.random(n).astype(np.float32)
    ave = rng.random(n).astype(np.float32)
    ave2 = (2 * ave.astype(np.float64) ** 2).astype(np.float32)
    ave2[:3] = [ave[0] ** 2, 0.0, ave[2] ** 2 * 0.5]   # var = 0, < 0, < 0 special cases
    gcc[:3] = [1.0, 1.0, 1.0]
    mask = (rng.random(n) > 0.2).astype(np.uint8)
    mask[:3] = 1
    lcc = np.full(n, -5.0, np.float32)
    d = [torch.from_numpy(v).cuda() for v in (gcc, ave, ave2, mask, lcc)]
    _lib.check(hip.pf_calc_lcc_dev(*[t.data_ptr() for t in d], n,
                                   torch.cuda.current_stream().cuda_stream), "calc_lcc")
    torch.cuda.synchronize()
    got = d[4].cpu().numpy()
    with np.errstate(divide="ignore", invalid="ignore"):
        expect = np.where(mask > 0, gcc.astype(np.float64) /
                          np.sqrt(ave2.astype(np.float64) - ave.astype(np.float64) ** 2), -5.0)
    fin = np.isfinite(expect)
    assert np.allclose(got[fin], expect[fin], rtol=1e-6)
    assert np.array_equal(np.isnan(got), np.isnan(expect))
    assert np.array_equal(np.isposinf(got), np.isposinf(expect))

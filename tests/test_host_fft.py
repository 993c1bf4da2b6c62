"""The device FFT butterflies, built for the host and run sequentially, match numpy."""
import ctypes
import os

import numpy as np
import pytest

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "powerfit_amd",
                   "libpf_fft_host.so")
FP = ctypes.POINTER(ctypes.c_float)


@pytest.fixture(scope="module")
def host():
    if not os.path.exists(LIB):
        pytest.fail("libpf_fft_host.so missing: run python -m powerfit_amd.build")
    return ctypes.CDLL(LIB)


SIZES = list(range(1, 65)) + [72, 75, 81, 96, 98, 120, 125, 128, 135, 143, 160, 192, 256, 343]


@pytest.mark.parametrize("n", SIZES)
def test_stockham_matches_numpy(host, n):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    for d in (-1, 1):
        out = np.zeros(n, np.complex64)
        assert host.pf_host_fft(n, d, x.ctypes.data_as(FP), out.ctypes.data_as(FP)) >= 0
        ref = np.fft.fft(x.astype(np.complex128)) if d < 0 else np.fft.ifft(x) * n
        assert np.abs(out - ref).max() / np.abs(ref).max() < 1e-6


def test_plans(host):
    buf = (ctypes.c_int * 16)()
    for n, expect in [(128, [8, 8, 2]), (96, [8, 4, 3]), (192, [8, 8, 3]), (160, [8, 4, 5]),
                      (48, [8, 2, 3]), (7, [7]), (1, [])]:
        ns = host.pf_host_plan(n, buf, 16)
        assert list(buf[:ns]) == expect
        assert int(np.prod(expect)) == n


# (N, T) of every register group FFT the fused kernels instantiate: rows (N/2, TX),
# columns and z (N, TY / TZ) for N in 48, 64, 72, 80, 96, 128, plus a few others
GROUP_PAIRS = [(24, 2), (32, 4), (36, 3), (40, 4), (48, 4), (64, 8), (72, 6), (80, 8),
               (96, 8), (128, 16), (128, 8), (64, 4), (96, 4), (160, 8), (192, 8), (100, 5)]


@pytest.mark.parametrize("n,t", GROUP_PAIRS)
def test_group_fft_matches_numpy(host, n, t):
    """Register result (group_fft) and the exchange buffer after the last stage
    (what group_fft_ex leaves in LDS) both equal numpy's transform."""
    rng = np.random.default_rng(n * 100 + t)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    for d in (-1, 1):
        ref = np.fft.fft(x.astype(np.complex128)) if d < 0 else np.fft.ifft(x) * n
        for fn in (host.pf_host_group_fft, host.pf_host_group_fft_ex):
            out = np.zeros(n, np.complex64)
            ns = fn(n, t, d, x.ctypes.data_as(FP), out.ctypes.data_as(FP))
            assert ns >= 1, (n, t)
            assert np.abs(out - ref).max() / np.abs(ref).max() < 2e-6, (fn, n, t, d)


@pytest.mark.parametrize("n,t,pad", [(32, 4, 8), (48, 4, 4), (64, 8, 8), (80, 4, 8), (96, 4, 16),
                                     (96, 8, 16)])
def test_group_fft_padded_exchanges_match_numpy(host, n, t, pad):
    """The row FFTs' padded LDS exchanges (element p at p + p // pad) give the
    same transform: in registers, and in natural order in the exchange buffer
    for the form that leaves its result there; a dropped or misplaced slot
    would scramble the output."""
    f = host.pf_host_group_fft_pad
    rng = np.random.default_rng(n * 100 + t)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    for d in (-1, 1):
        ref = np.fft.fft(x.astype(np.complex128)) if d < 0 else np.fft.ifft(x) * n
        for to_ex in (0, 1):
            out = np.zeros(n, np.complex64)
            slots = f(n, t, pad, to_ex, d, x.ctypes.data_as(FP), out.ctypes.data_as(FP))
            assert slots == n + (n - 1) // pad
            assert np.abs(out - ref).max() / np.abs(ref).max() < 1e-6


def _smooth(n):
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


CT_LENGTHS = [n for n in range(8, 257) if _smooth(n)]


@pytest.mark.parametrize("n", CT_LENGTHS)
def test_compile_time_length_passes_match_numpy(host, n):
    """The compile-time-length LDS passes of the run-time planned pipeline
    (every 2,3,5,7-smooth length from 8 to 256), run stage by stage on the
    host, equal numpy's transform in both directions; their plans use radices
    up to 16: one stage up to 16 points, two wherever a two-factor split
    into such radices exists, else three."""
    rng = np.random.default_rng(n + 7)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    buf = (ctypes.c_int * 8)()
    ns = host.pf_host_ct_plan(n, buf, 8)
    radices = list(buf[:ns])
    assert int(np.prod(radices)) == n and max(radices) <= 16
    two = any(n % a == 0 and n // a <= 16 for a in range(2, 17))
    assert ns == (1 if n <= 16 else 2 if two else 3), radices
    for d in (-1, 1):
        out = np.zeros(n, np.complex64)
        assert host.pf_host_ct_fft(n, d, x.ctypes.data_as(FP), out.ctypes.data_as(FP)) == ns
        ref = np.fft.fft(x.astype(np.complex128)) if d < 0 else np.fft.ifft(x) * n
        assert np.abs(out - ref).max() / np.abs(ref).max() < 2e-6, (n, d, radices)


@pytest.mark.parametrize("slots", [16, 8])
@pytest.mark.parametrize("n", [n for n in CT_LENGTHS if n >= 16])
def test_register_fft_with_stagewise_butterflies_matches_numpy(host, n, slots):
    """The register FFTs of the run-time planned z stage (each stage with its
    own butterflies per thread; the default plan of 16 register slots and the
    alternative of 8, z schedule bit 0, where a length has one) equal numpy's
    transform, forward and inverse, in both stage orders; the forward plan's
    output pattern is the reversed plan's input pattern and vice versa (the z
    stage multiplies and transforms back without an exchange).  A wrong
    idle-slot test or a misplaced output would fail here."""
    rng = np.random.default_rng(n + 11)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    info = (ctypes.c_int * 4)()
    probe = np.zeros(n, np.complex64)
    if host.pf_host_flex_fft(n, slots, -1, 0, x.ctypes.data_as(FP), probe.ctypes.data_as(FP),
                             info) == -2:
        assert slots == 8
        pytest.skip(f"no {slots}-slot register plan for length {n}")
    for d in (-1, 1):
        ref = np.fft.fft(x.astype(np.complex128)) if d < 0 else np.fft.ifft(x) * n
        for rev in (0, 1):
            out = np.full(n, np.nan, np.complex64)
            assert host.pf_host_flex_fft(n, slots, d, rev, x.ctypes.data_as(FP),
                                         out.ctypes.data_as(FP), info) == 0
            assert np.abs(out - ref).max() / np.abs(ref).max() < 2e-6, (n, d, rev, list(info))
    t, e, ns, same = list(info)
    assert same == 1 and e <= slots and t in (1, 2, 4, 8, 16, 32) and 1 <= ns <= 3


def test_alternative_register_plans_cover_the_cli_lengths(host):
    """The lengths the command line's grids take most (70, 75, 84, 96, 98,
    105, 112, 120, 126, 128) all have the alternative 8-slot plan."""
    info = (ctypes.c_int * 4)()
    for n in (70, 75, 84, 96, 98, 105, 112, 120, 126, 128):
        x = np.ones(n, np.complex64)
        out = np.zeros(n, np.complex64)
        assert host.pf_host_flex_fft(n, 8, -1, 0, x.ctypes.data_as(FP),
                                     out.ctypes.data_as(FP), info) == 0, n

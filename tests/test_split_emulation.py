"""The split-precision DFT stage, emulated in NumPy: per-column power-of-two
scaling, FP16 hi + lo operands, Fh Xh + Fh Xl + Fl Xh accumulated in FP32.

Pins the method the MFMA stage implements (pf_split.hip): its error against
the FP64 16-point DFT stays at FP32 level relative to the column's largest
input, and plausible mistakes (dropping a lo part, skipping the scaling) fail."""
import numpy as np


def dft16_real(sign):
    k = np.arange(16)
    ang = 2 * np.pi * np.outer(k, k) / 16
    c, d = np.cos(ang), sign * np.sin(ang)
    return np.block([[c, -d], [d, c]])        # rows Re/Im y, columns Re/Im x


def split16(v):
    hi = v.astype(np.float16)
    lo = (v - hi.astype(np.float32)).astype(np.float16)
    return hi.astype(np.float32), lo.astype(np.float32)


def emulate(x, sign, use_lo=True, scale=True):
    """x: (16, ncols) complex; returns (16, ncols) complex64 like the kernel."""
    F = dft16_real(sign).astype(np.float32)
    X = np.concatenate([x.real, x.imag]).astype(np.float32)        # (32, ncols)
    m = np.abs(X).max(axis=0)
    e = np.where(m > 0, np.frexp(m)[1], 0) if scale else np.zeros(X.shape[1], int)
    Xs = X * np.ldexp(np.float32(1), -e)[None, :].astype(np.float32)
    Fh, Fl = split16(F)
    Xh, Xl = split16(Xs)
    acc = (Fh @ Xh).astype(np.float32)
    if use_lo:
        acc = acc + (Fh @ Xl).astype(np.float32) + (Fl @ Xh).astype(np.float32)
    acc = acc * np.ldexp(np.float32(1), e)[None, :]
    return acc[:16] + 1j * acc[16:]


def cases():
    rng = np.random.default_rng(16)
    x = rng.standard_normal((16, 512)) + 1j * rng.standard_normal((16, 512))
    x *= 10.0 ** rng.uniform(-6, 8, 512)[None, :]          # wide range across columns
    x[:, :8] *= 10.0 ** rng.uniform(-3, 0, (16, 8))         # and within some columns
    return x


def rel_err(y, x, sign):
    ref = np.fft.fft(x, axis=0) if sign < 0 else np.fft.ifft(x, axis=0) * 16
    return (np.abs(y - ref).max(axis=0) / np.abs(x).max(axis=0)).max()


def test_split_stage_is_fp32_accurate():
    x = cases()
    for sign in (-1, 1):
        assert rel_err(emulate(x, sign), x, sign) < 3e-6


def test_dropping_lo_or_scaling_fails():
    x = cases()
    assert rel_err(emulate(x, -1, use_lo=False), x, -1) > 1e-4
    # without the scaling the large columns overflow FP16
    with np.errstate(over="ignore", invalid="ignore"):
        err = rel_err(emulate(x, -1, scale=False), x, -1)
    assert not np.isfinite(err) or err > 1e-2

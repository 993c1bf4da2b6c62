"""Host-side checks of powerfit_amd.volume (argument handling, grids), no GPU."""
import numpy as np
import pytest

from powerfit_amd import volume


def test_argument_errors_before_any_device_call():
    out = np.zeros((4, 4, 4))
    with pytest.raises(ValueError):
        volume.blur_points(np.zeros((2, 3)), np.ones(3), 1.0, out)
    with pytest.raises(ValueError):
        volume.blur_points(np.zeros((3, 3)), np.ones(2), 1.0, out)
    with pytest.raises(ValueError):
        volume.blur_points(np.zeros((3, 3)), np.ones(3), 1.0, out.astype(np.float32))
    with pytest.raises(ValueError):
        volume.dilate_points(np.zeros((3, 3)), np.ones(2), out)
    with pytest.raises(ValueError):
        volume.dilate_points(np.zeros((3, 3)), np.ones(3), out[:, :, ::2])
    with pytest.raises(ValueError):
        volume.structure_to_shape(np.zeros((3, 3)), 4.0, shape="surface")
    with pytest.raises(ValueError):
        volume.structure_to_shape_like(volume.Volume(np.zeros((4, 4, 4))), np.zeros((3, 3)),
                                       resolution=4.0, weights=np.ones(2))


def test_volume_and_sigma_helpers():
    v = volume.Volume(np.zeros((4, 5, 6)), voxelspacing=2.0, origin=(1.0, 2.0, 4.0))
    assert v.shape == (4, 5, 6)
    assert np.array_equal(v.dimensions, [12.0, 10.0, 8.0])
    assert np.array_equal(v.start, [0.5, 1.0, 2.0])
    assert volume.zeros_like(v).shape == v.shape
    assert volume.sigma_to_res(volume.res_to_sigma(6.0)) == pytest.approx(6.0)
    assert volume.res_to_sigma(np.sqrt(2.0) * np.pi) == pytest.approx(1.0)


def _example():
    array = np.zeros((3, 4, 5))
    array[0, 0, 0] = 1.1
    array[1, 2, 3] = 2.2
    array[2, 3, 4] = 3.3
    return volume.Volume(array)


def test_reference_volume_file_cases(tmp_path):
    """The reference's own volume tests (tests/test_volume.py:10-54)."""
    v = _example()
    fn = tmp_path / "example_volume.mrc"
    v.tofile(str(fn))
    assert fn.exists() and fn.stat().st_size == 1024 + 4 * 60
    assert v.shape == (3, 4, 5)
    assert np.array_equal(v.dimensions, (5.0, 4.0, 3.0))
    assert np.array_equal(v.start, (0, 0, 0))
    r = volume.Volume.fromfile(str(fn))
    assert r.shape == (3, 4, 5)
    assert np.array_equal(r.start, (0, 0, 0))
    assert np.array_equal(r.dimensions, (5.0, 4.0, 3.0))
    np.testing.assert_array_almost_equal(r.array, v.array)
    assert r.array.dtype == np.float64


@pytest.mark.parametrize("ext", ["ccp4", "map", "mrc"])
def test_map_round_trip_keeps_origin_and_values(tmp_path, ext):
    rng = np.random.default_rng(0)
    arr = rng.standard_normal((6, 7, 8)).astype(np.float32)
    origin = (2.4, -1.2, 3.6) if ext != "mrc" else (0.37, -5.5, 12.25)
    v = volume.Volume(arr, voxelspacing=1.2, origin=origin)
    fn = str(tmp_path / f"m.{ext}")
    v.tofile(fn)
    r = volume.Volume.fromfile(fn)
    assert r.shape == arr.shape and r.voxelspacing == pytest.approx(1.2, rel=1e-6)
    assert np.array_equal(r.array, arr.astype(np.float64))
    np.testing.assert_allclose(r.origin, origin, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("dtype,mode,read_as", [(np.int8, 0, np.int8), (np.int16, 1, np.int32),
                                                (np.int32, 1, np.int32)])
def test_integer_modes(tmp_path, dtype, mode, read_as):
    arr = (np.arange(60).reshape(3, 4, 5) - 30).astype(dtype)
    fn = str(tmp_path / "i.mrc")
    volume.Volume(arr).tofile(fn)
    raw = open(fn, "rb").read()
    assert np.frombuffer(raw[12:16], "<i4")[0] == mode
    r = volume.Volume.fromfile(fn)
    assert r.array.dtype == read_as and np.array_equal(r.array, arr.astype(read_as))


def _raw_map(arr_file, axes, endian="<", nsymbt=0, spacing=(1.0, 1.0, 1.0), angles=90.0,
             nstart=(0, 0, 0)):
    """A map file written field by field: arr_file is (sections, rows, columns)."""
    ns, nr, nc = arr_file.shape
    counts = (nc, nr, ns)
    m = [0, 0, 0]
    for i, a in enumerate(axes):
        m[a - 1] = counts[i]
    h = np.zeros((), dtype=volume._header_dtype(endian))
    h["n"] = counts
    h["mode"] = 2
    h["nstart"] = nstart
    h["m"] = m
    h["cell"] = [m[a] * spacing[a] for a in range(3)]
    h["angles"] = (angles, 90.0, 90.0)
    h["axes"] = axes
    h["nsymbt"] = nsymbt
    h["map"] = b"MAP "
    h["machst"] = (0x44, 0x41, 0, 0) if endian == "<" else (0x11, 0x11, 0, 0)
    return (h.tobytes() + b"\0" * nsymbt
            + np.ascontiguousarray(arr_file, dtype=endian + "f4").tobytes())


def test_big_endian_extended_header_and_permuted_axes(tmp_path):
    rng = np.random.default_rng(4)
    zyx = rng.standard_normal((3, 4, 5))                    # (z, y, x)
    # columns along z, rows along x, sections along y: file array (y, x, z)
    fn = tmp_path / "p.ccp4"
    fn.write_bytes(_raw_map(np.transpose(zyx, (1, 2, 0)), (3, 1, 2), endian=">", nsymbt=80,
                            nstart=(7, 5, 6)))
    r = volume.Volume.fromfile(str(fn))
    np.testing.assert_allclose(r.array, zyx, rtol=1e-6)
    # nstart is (columns=z, rows=x, sections=y) -> origin (x, y, z) = (5, 6, 7)
    assert np.array_equal(r.origin, (5.0, 6.0, 7.0))


def test_map_header_errors(tmp_path):
    a = np.zeros((2, 2, 2))
    bad = tmp_path / "a.ccp4"
    bad.write_bytes(_raw_map(a, (1, 2, 3), angles=60.0))
    with pytest.raises(RuntimeError):
        volume.Volume.fromfile(str(bad))
    bad.write_bytes(_raw_map(a, (1, 2, 3), spacing=(1.0, 1.5, 1.0)))
    with pytest.raises(RuntimeError):
        volume.Volume.fromfile(str(bad))
    raw = bytearray(_raw_map(a, (1, 2, 3)))
    raw[212] = 0
    bad.write_bytes(bytes(raw))
    with pytest.raises(RuntimeError):
        volume.Volume.fromfile(str(bad))
    with pytest.raises(ValueError):
        volume.read_map(str(tmp_path / "x.pdb"))
    with pytest.raises(RuntimeError):
        volume.Volume(a).tofile(str(tmp_path / "x.xyz"))


def test_trim_extend_resample_closed_forms():
    arr = np.zeros((10, 12, 14))
    arr[3:5, 6, 2:9] = 2.0
    v = volume.Volume(arr, voxelspacing=1.5, origin=(1.0, 2.0, 3.0))
    t = volume.trim(v, 1.0, margin=2)
    assert t.shape == (6, 5, 11)                       # z 1..6, y 4..8, x 0..10
    assert np.array_equal(t.array, arr[1:7, 4:9, 0:11])
    assert np.allclose(t.origin, (1.0 + 0 * 1.5, 2.0 + 4 * 1.5, 3.0 + 1 * 1.5))
    with pytest.raises(ValueError):
        volume.trim(v, 2.0)
    e = volume.extend(t, (8, 8, 12))
    assert e.shape == (8, 8, 12) and np.array_equal(e.array[:6, :5, :11], t.array)
    assert e.array[6:].max() == 0 and e.origin == t.origin
    r = volume.resample(v, 0.5)
    assert r.shape == (5, 6, 7) and r.voxelspacing == 3.0


def test_nearest_multiple2357_brute_force():
    def smooth(n):
        return all(p in (2, 3, 5, 7) for p in _factors(n))
    for n in range(1, 400):
        got = volume.nearest_multiple2357(n)
        assert smooth(got) and all(not smooth(k) for k in range(n, got))
        assert volume.is_multiple2357(n) == smooth(n)
    assert not volume.is_multiple2357(0) and volume.nearest_multiple2357(0) == 1


def _factors(n):
    out, p = [], 2
    while n > 1 and p * p <= n:
        while n % p == 0:
            out.append(p)
            n //= p
        p += 1
    if n > 1:
        out.append(n)
    return out

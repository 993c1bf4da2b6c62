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

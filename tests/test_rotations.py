"""Rotation sets (rotations.py:30-61, 88-131 of the reference)."""
import numpy as np
import pytest

from powerfit_amd import rotations as rot


def test_quat_to_rotmat_closed_forms():
    R = rot.quat_to_rotmat(np.array([[1.0, 0, 0, 0]]))
    assert np.array_equal(R[0], np.eye(3))
    c = np.cos(np.pi / 4)
    R = rot.quat_to_rotmat(np.array([[c, 0, 0, c]]))  # +90 deg about z
    assert np.array_equal(R[0], np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1.0]]))
    R = rot.quat_to_rotmat(np.array([[c, c, 0, 0]]))  # +90 deg about x
    assert np.array_equal(R[0], np.array([[1, 0, 0], [0, 0, -1], [0, 1, 0.0]]))
    # non-unit quaternion: s = 2 / |q|^2 normalises
    assert np.allclose(rot.quat_to_rotmat(np.array([[2.0, 0, 0, 0]]))[0], np.eye(3))
    assert np.array_equal(rot.quat_to_rotmat(np.zeros((1, 4)))[0], np.eye(3))


@pytest.mark.parametrize("name", sorted(rot.SHA256))
def test_shipped_sets(name):
    qw = rot.load_set(name)
    assert qw.shape == (rot.ROT_SETS[name][0], 5)
    R = rot.quat_to_rotmat(qw[:, :4])
    assert np.array_equal(R[0], np.eye(3))           # row 0 is the identity
    assert np.allclose(np.linalg.det(R), 1, atol=1e-7)
    assert np.allclose(R @ R.transpose(0, 2, 1), np.eye(3), atol=1e-7)
    assert np.array_equal(R, np.around(R, 8))


def test_nearest_set_rule():
    assert rot.nearest_set(10) == "c48n309.npy"
    assert rot.nearest_set(5) == "c48u2219.npy"
    assert rot.nearest_set(20) == "c48u27.npy"
    assert rot.nearest_set(360) == "E.npy"
    q, w, alpha = rot.proportional_orientations(10)
    assert q.shape == (7416, 4) and alpha == 9.72


def test_unshipped_set_is_a_clear_error():
    with pytest.raises(FileNotFoundError):
        rot.proportional_orientations(4.0)

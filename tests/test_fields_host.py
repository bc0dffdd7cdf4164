"""Host side of the field maps (lumina_amd/fields.py), on the CPU: the plane
basis and observation grid of fields.rs:51-141 and the circular-dichroism
sum of fields.rs:236-253."""
import numpy as np
import pytest

from lumina_amd import fields as F


@pytest.mark.parametrize("normal", [(0, 0, 1), (0, 0, -2), (1, 0, 0), (0.95, 0.1, 0.2),
                                    (0.3, -0.7, 0.5), (0, 1, 0)])
def test_plane_basis_is_orthonormal_and_right_handed(normal):
    u, v = F.plane_basis(normal)
    n = np.asarray(normal, float) / np.linalg.norm(normal)
    assert abs(np.linalg.norm(u) - 1) < 1e-15 and abs(np.linalg.norm(v) - 1) < 1e-15
    assert abs(np.dot(u, n)) < 1e-15 and abs(np.dot(v, n)) < 1e-15 and abs(np.dot(u, v)) < 1e-15
    assert np.allclose(np.cross(u, v), n, atol=1e-15)
    # the seed is x unless |n_x| >= 0.9 (fields.rs:115-119): u lies in the
    # plane spanned by n and the seed
    seed = np.array([1.0, 0, 0]) if abs(n[0]) < 0.9 else np.array([0, 1.0, 0])
    assert abs(np.dot(np.cross(n, seed), u)) < 1e-15


def test_plane_basis_special_cases():
    assert F.plane_basis((0, 0, 1)) == ([1.0, 0.0, 0.0], [0.0, 1.0, 0.0])
    u, v = F.plane_basis((1, 0, 0))
    assert u == [0.0, 1.0, 0.0] and v == [0.0, 0.0, 1.0]


def test_plane_points_grid_and_extent():
    plane = F.NearFieldPlane((5.0, -1.0, 2.0), (0, 0, 1), 10.0, 4.0, 6, 3)
    pos, extent = F.plane_points(plane)
    assert pos.shape == (18, 3)
    xs = 5.0 + np.linspace(-10.0, 10.0, 6)
    ys = -1.0 + np.array([-4.0, 0.0, 4.0])
    assert np.allclose(pos[:, 0], np.tile(xs, 3), atol=1e-14)      # x fastest
    assert np.allclose(pos[:, 1], np.repeat(ys, 6), atol=1e-14)
    assert np.all(pos[:, 2] == 2.0)
    assert extent == (-5.0, 15.0, -5.0, 3.0)
    # one point per axis: dx = dy = 0, the lower-left corner
    one, _ = F.plane_points(F.NearFieldPlane((0, 0, 0), (0, 0, 1), 3.0, 2.0, 1, 1))
    assert np.array_equal(one, [[-3.0, -2.0, 0.0]])


def test_circular_dichroism_sum():
    rng = np.random.default_rng(0)
    px = rng.standard_normal((7, 3)) + 1j * rng.standard_normal((7, 3))
    py = rng.standard_normal((7, 3)) + 1j * rng.standard_normal((7, 3))
    k = 0.0123
    want = k * (np.sum(px[:, 1].imag) - np.sum(py[:, 0].imag))
    assert abs(F.compute_circular_dichroism(px, py, k) - want) < 1e-15
    # p_x under y-polarisation equal to p_y under x-polarisation: every term
    # cancels exactly
    same = px.copy()
    same[:, 0] = px[:, 1]
    assert F.compute_circular_dichroism(px, same, k) == 0.0

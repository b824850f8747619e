"""Pins the CPU oracle of point_cloud_extract (cuda_point_cloud.cu:22-70)
and its float64 twin point_cloud_from_depth (point_cloud_utility.cpp:76-115):
round trips through the projection, planes stay planes, and the two agree."""
import numpy as np
from scipy.spatial.transform import Rotation

import oracle_ffi as O
from recon import scenes


def _pose(rng):
    q = Rotation.from_rotvec(rng.normal(0, 0.5, 3)).as_quat()
    return np.array([q[3], q[0], q[1], q[2], *rng.normal(0, 0.3, 3)])


def test_round_trip_projection():
    rng = np.random.default_rng(2)
    rows, cols = 37, 53
    intr = (61.0, 59.0, 26.5, 18.0)
    depth = rng.uniform(0.3, 4.0, (rows, cols)).astype(np.float32)
    depth[rng.random(depth.shape) < 0.1] = 0.0
    T = _pose(rng)
    xyz, valid, _ = O.depth_to_points(depth, intr, T)
    R = Rotation.from_quat([T[1], T[2], T[3], T[0]]).as_matrix()
    pc = (xyz.astype(np.float64) - T[4:]) @ R  # back to the camera frame
    v, u = np.divmod(np.arange(rows * cols), cols)
    m = valid.astype(bool)
    assert np.array_equal(m, depth.ravel() > 0)
    np.testing.assert_allclose(pc[m, 2], depth.ravel()[m], rtol=3e-6, atol=3e-6)
    np.testing.assert_allclose(pc[m, 0] * intr[0] / pc[m, 2] + intr[2], u[m], atol=2e-3)
    np.testing.assert_allclose(pc[m, 1] * intr[1] / pc[m, 2] + intr[3], v[m], atol=2e-3)
    assert (xyz[~m] == 0).all()


def test_plane_points_are_coplanar():
    rows, cols = 48, 64
    T = scenes.look_at_pose((0.3, -0.2, 1.5), (0.0, 0.1, 0.0))
    depth = np.full((rows, cols), 1.25, np.float32)  # fronto-parallel plane
    xyz, valid, _ = O.depth_to_points(depth, (50.0, 50.0, 31.5, 23.5), [1.0, 0, 0, 0, 0, 0, 0])
    assert valid.all()
    np.testing.assert_allclose(xyz[:, 2], 1.25, atol=0)
    # camera->world transform of those points: one plane, normal = camera axis
    xyz2, _, _ = O.depth_to_points(depth, (50.0, 50.0, 31.5, 23.5), T)
    p = xyz2.astype(np.float64)
    centred = p - p.mean(0)
    sing = np.linalg.svd(centred, compute_uv=False)
    assert sing[2] < 1e-5 * sing[0]


def test_float32_kernel_agrees_with_float64_path():
    rng = np.random.default_rng(4)
    depth = rng.uniform(0.5, 3.0, (30, 40)).astype(np.float32)
    depth[::3, ::4] = 0.0
    T = _pose(rng)
    intr = (45.0, 47.0, 19.5, 14.5)
    xyz, valid, _ = O.depth_to_points(depth, intr, T)
    p64, idx = O.point_cloud_from_depth_f64(depth, intr, T, dmin=0.0, dmax=1e30)
    assert np.array_equal(np.flatnonzero(valid), idx)
    np.testing.assert_allclose(xyz[idx], p64, atol=2e-6 * (1 + np.abs(p64).max()))


def test_colours():
    depth = np.ones((4, 5), np.float32)
    depth[1, 2] = 0.0
    rgb = np.arange(4 * 5 * 3, dtype=np.uint8).reshape(4, 5, 3)
    _, valid, col = O.depth_to_points(depth, (5.0, 5.0, 2.0, 1.5), [1.0, 0, 0, 0, 0, 0, 0], rgb=rgb)
    m = valid.astype(bool)
    assert np.array_equal(col[m], rgb.reshape(-1, 3)[m]) and (col[~m] == 0).all()
    _, _, white = O.depth_to_points(depth, (5.0, 5.0, 2.0, 1.5), [1.0, 0, 0, 0, 0, 0, 0])
    assert (white[m] == 255).all()

"""The synthetic workload generator: exact ray-cast depth against closed
forms, bit-identical frames for any thread count, and noise statistics."""
import numpy as np

from recon import scenes


def test_look_at_pose():
    c, target = np.array([1.2, -0.4, 0.7]), np.array([0.1, 0.2, 0.0])
    pose = scenes.look_at_pose(c, target)
    R, t = scenes.pose_matrix(pose)
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)
    assert abs(np.linalg.det(R) - 1) < 1e-12
    np.testing.assert_allclose(-R.T @ t, c, atol=1e-12)  # camera centre
    p = R @ target + t
    assert p[2] > 0 and abs(p[0]) < 1e-12 and abs(p[1]) < 1e-12  # on the optical axis
    up = R @ np.array([0, 0, 1.0])
    assert up[1] < 0  # world up points to image up (camera -y)


def test_plane_depth_is_constant_z():
    w = scenes.plain_small(num_frames=1)
    w.prims = [(scenes.PRIM_PLANE_Z, (0.0,))]
    w.poses = np.array([scenes.look_at_pose((0.0, 0.0, 2.0), (0.0, 0.0, 0.0), up=(0, 1, 0))])
    d = w.render(noise=False)[0]
    np.testing.assert_allclose(d, 2.0, atol=1e-6)


def test_sphere_depth_on_axis():
    w = scenes.plain_small(num_frames=1)
    w.prims = [(scenes.PRIM_SPHERE, (0.0, 0.0, 0.0, 0.3))]
    w.intrinsics = (525.0, 525.0, 320.0, 240.0)
    w.poses = np.array([scenes.look_at_pose((1.5, 0.0, 0.0), (0.0, 0.0, 0.0))])
    d = w.render(noise=False)[0]
    assert abs(d[240, 320] - 1.2) < 1e-6
    assert d[0, 0] == 0.0  # miss


def test_deterministic_for_any_thread_count():
    w = scenes.plain_small(num_frames=2)
    a = w.render(nthreads=1)
    b = w.render(nthreads=7)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    # frame identity, not position in the call, drives the noise
    c = w.render(frames=[1])
    assert np.array_equal(c[0].view(np.uint32), a[1].view(np.uint32))


def test_noise_statistics():
    w = scenes.plain_small(num_frames=1)
    w.prims = [(scenes.PRIM_PLANE_Z, (0.0,))]
    w.poses = np.array([scenes.look_at_pose((0.0, 0.0, 2.0), (0.0, 0.0, 0.0), up=(0, 1, 0))])
    w.noise = (0.01, 0.0, 0.05, 11)
    d = w.render()[0]
    valid = d > 0
    assert abs((~valid).mean() - 0.05) < 0.005
    err = d[valid] - 2.0
    assert abs(err.mean()) < 5e-4 and abs(err.std() - 0.01) < 5e-4


def test_workloads():
    big = scenes.plain_large(num_frames=300)
    assert big.grid == (512, 512, 512) and big.num_frames == 300 and len(big.prims) == 41
    r = np.linalg.norm([(-scenes.pose_matrix(p)[0].T @ scenes.pose_matrix(p)[1])[:2] for p in big.poses],
                       axis=1)
    assert r.min() >= 1.2 - 1e-9 and r.max() <= 1.8 + 1e-9
    vols = scenes.batched_uneven()
    assert len(vols) == 64
    assert {v.grid[0] for v in vols} <= {64, 96, 128, 192, 256}
    assert all(1 <= v.num_frames <= 40 for v in vols)
    assert {v.image for v in vols} <= {(320, 240), (640, 480), (1280, 720)}
    small = scenes.plain_small()
    assert small.num_frames == 10 and abs(small.truncation - 0.032) < 1e-12

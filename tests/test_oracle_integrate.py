"""Pins the CPU oracle of cu_tsdf_integrate (tsdf_volume.cu:144-190 under the
build contract) against closed forms and invariants computed independently
(float64 numpy + scipy rotations), so that a dropped term, a wrong sign or
index, a transposed rotation or a rounding-instead-of-truncation pixel index
fails here."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import oracle_ffi as O
from recon import scenes


def _vol(grid, vs, off, trunc, maxw=128.0, intr=(500.0, 500.0, 319.5, 239.5)):
    return O.volume(grid, [vs] * 3, off, trunc, maxw, intr)


def _centres(vol):
    g = list(vol.grid)
    vs = np.float64(np.float32(vol.voxel_size[0]))
    off = [np.float64(np.float32(o)) for o in vol.offset]
    z, y, x = np.meshgrid(*[(np.arange(g[i]) + 0.5) * vs + off[i] for i in (2, 1, 0)], indexing="ij")
    return np.stack([x, y, z], -1)


def _cam(pose, pts):
    """Camera coordinates of world points, rotation from scipy."""
    q = np.asarray(pose[:4], np.float64)
    R = Rotation.from_quat([q[1], q[2], q[3], q[0]]).as_matrix()  # scipy: x,y,z,w
    return pts @ R.T + np.asarray(pose[4:7])


def _pixels(p, intr, cols, rows):
    fx, fy, cx, cy = intr
    z = p[..., 2]
    with np.errstate(all="ignore"):
        su = p[..., 0] * fx / z + cx
        sv = p[..., 1] * fy / z + cy
    ok = (z > 0) & (su > -1) & (su < cols) & (sv > -1) & (sv < rows)
    # distance of s to the nearest integer: decisions there are float-sensitive
    amb = np.minimum(np.abs(su - np.round(su)), np.abs(sv - np.round(sv))) < 1e-3
    return su, sv, ok, amb


def test_fronto_parallel_plane_rotated_camera():
    """Constant depth Z seen by a rotated camera: every voxel it updates gets
    D = min(Z - z_cam, trunc), W = 1; z_cam from scipy's rotation."""
    rng = np.random.default_rng(3)
    grid, vs, off, trunc = (24, 20, 16), 0.02, (-0.24, -0.2, -0.16), 0.05
    vol = _vol(grid, vs, off, trunc)
    for _ in range(3):
        rot = Rotation.from_rotvec(rng.normal(0, 0.4, 3))
        q = rot.as_quat()  # x,y,z,w
        pose = np.array([q[3], q[0], q[1], q[2], *rng.normal(0, 0.05, 2), 0.8])
        Z = 0.82
        depth = np.full((480, 640), Z, np.float32)
        D, W = O.init(vol)
        n, mask = O.integrate_f32(vol, D, W, depth, pose, mask=True)
        p = _cam(pose, _centres(vol))
        su, sv, ok, amb = _pixels(p, (500.0, 500.0, 319.5, 239.5), 640, 480)
        sdf = Z - p[..., 2]
        expect_mask = ok & (sdf >= -trunc)
        near = amb | (np.abs(sdf + trunc) < 1e-5)
        assert n > 1000
        assert np.array_equal(mask.astype(bool)[~near], expect_mask[~near])
        m = mask.astype(bool)
        np.testing.assert_allclose(D[m], np.minimum(sdf[m], trunc), atol=2e-6)
        assert (W[m] == 1).all() and (W[~m] == 0).all() and (D[~m] == np.float32(trunc)).all()


def test_pixel_index_is_truncation_of_projection():
    """Depth map encoding its pixel: D + z_cam recovers the depth of pixel
    ((int)(x fx / z + cx), (int)(y fy / z + cy)), not of a rounded or
    transposed index."""
    grid, vs, off = (30, 26, 10), 0.01, (-0.15, -0.13, 0.9)
    trunc = 50.0  # never clamps
    cols, rows = 64, 48
    intr = (40.0, 41.0, 31.5, 23.5)
    vol = _vol(grid, vs, off, trunc, intr=intr)
    u, v = np.meshgrid(np.arange(cols), np.arange(rows))
    depth = (5.0 + 0.01 * u + 0.0001 * v).astype(np.float32)
    pose = np.array([1.0, 0, 0, 0, 0.003, -0.002, 0.0])
    D, W = O.init(vol)
    n, mask = O.integrate_f32(vol, D, W, depth, pose, mask=True)
    p = _cam(pose, _centres(vol))
    su, sv, ok, amb = _pixels(p, intr, cols, rows)
    m = mask.astype(bool)
    assert n > 0 and np.array_equal(m[~amb], ok[~amb])
    sel = m & ~amb
    iu, iv = np.trunc(su[sel]).astype(int), np.trunc(sv[sel]).astype(int)
    np.testing.assert_allclose(D[sel] + p[..., 2][sel], depth[iv, iu], atol=2e-5)
    # a rounded index would differ for about half of the voxels
    ru = np.round(su[sel]).astype(int).clip(0, cols - 1)
    assert (np.abs(depth[iv, ru] - depth[iv, iu]) > 1e-3).mean() > 0.3


def test_running_average_is_the_mean_and_weight_caps():
    """Frames of constant depths Z_k at the identity pose: while k <= cap the
    TSDF is the arithmetic mean of the clamped sdf values, W = k; past the
    cap W stays at the cap."""
    grid, vs, off, trunc = (8, 8, 12), 0.01, (-0.04, -0.04, 0.5), 0.03
    cap = 4.0
    vol = _vol(grid, vs, off, trunc, maxw=cap)
    pose = [1.0, 0, 0, 0, 0, 0, 0]
    Zs = [0.55, 0.56, 0.52, 0.58, 0.57, 0.53]
    D, W = O.init(vol)
    z = _centres(vol)[..., 2]
    for k, Z in enumerate(Zs, start=1):
        O.integrate_f32(vol, D, W, np.full((480, 640), Z, np.float32), pose)
        if k <= cap:
            vals = [np.minimum(Zj - z, trunc) for Zj in Zs[:k]]
            upd = np.all([Zj - z >= -trunc for Zj in Zs[:k]], axis=0)
            np.testing.assert_allclose(D[upd], np.mean(vals, axis=0)[upd], atol=3e-6)
            assert (W[upd] == k).all()
    assert W.max() == cap


def test_invariants_on_sphere_scene():
    w = scenes.plain_small(num_frames=4)
    depth = w.render()
    vol = O.volume_from_config(w.config())
    D, W = O.init(vol)
    masks = []
    for f in range(w.num_frames):
        _, m = O.integrate_f32(vol, D, W, depth[f], w.poses[f], mask=True)
        masks.append(m)
    trunc = np.float32(vol.truncation)
    assert (np.abs(D) <= trunc).all()
    assert np.array_equal(W, np.sum(masks, axis=0).astype(np.float32))
    untouched = W == 0
    assert (D[untouched] == trunc).all()
    assert 0.05 < (~untouched).mean() < 0.9


def test_float64_twin_within_1e6():
    w = scenes.plain_small(num_frames=3)
    depth = w.render()
    vol = O.volume_from_config(w.config())
    D, W = O.init(vol)
    D64, W64 = D.astype(np.float64), W.astype(np.float64)
    for f in range(w.num_frames):
        n32 = O.integrate_f32(vol, D, W, depth[f], w.poses[f])
        n64, flips = O.integrate_f64(vol, D64, W64, depth[f], w.poses[f])
        assert n32 == n64
        assert flips < 1e-4 * n32  # decisions an independent float64 pass changes
    assert np.abs(D - D64).max() < 1e-6
    assert np.array_equal(W.astype(np.float64), W64)


def test_pose_matrix_matches_scipy():
    rng = np.random.default_rng(5)
    for _ in range(10):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        pose = np.array([*q, 0.1, -0.2, 0.3])
        R, t = O.pose_matrix_f32(pose)
        Rs = Rotation.from_quat([q[1], q[2], q[3], q[0]]).as_matrix()
        np.testing.assert_allclose(R, Rs, atol=1e-6)
        np.testing.assert_allclose(t, pose[4:], atol=1e-7)


def test_behind_camera_is_never_updated():
    # the volume lies behind the camera: the reference would mirror it into
    # the image (camera_util.cu:7-19); the contract requires camera z > 0
    vol = _vol((10, 10, 10), 0.02, (-0.1, -0.1, -0.6), 1.0)
    D, W = O.init(vol)
    n = O.integrate_f32(vol, D, W, np.full((480, 640), 0.5, np.float32), [1.0, 0, 0, 0, 0, 0, 0])
    assert n == 0 and (W == 0).all()


@pytest.mark.parametrize("bad", [np.nan, np.inf, -1.0, 0.0, 24.5])
def test_invalid_depth_values_skip(bad):
    vol = _vol((8, 8, 8), 0.01, (-0.04, -0.04, 0.5), 0.03)
    D, W = O.init(vol)
    n = O.integrate_f32(vol, D, W, np.full((480, 640), bad, np.float32), [1.0, 0, 0, 0, 0, 0, 0])
    assert n == 0

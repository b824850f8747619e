"""Pins the oracle's z-buffer render (oracle_render_points, following
cuda_point_cloud.cu:229-353 and filter_image_if :89-99) without retyping its
loop: a fronto-parallel plane sampled at pixel centres renders to its own
depth at every pixel (projection, intrinsics and pose direction), the
selection rule equals an order-free formulation (smallest z, then smallest
index, via lexsort), and the reference's edge rules hold (truncation toward
zero, z == 0 skipped, negative z kept, FLT_MAX filter)."""
import numpy as np
import pytest

import oracle_ffi as O
from recon import scenes

INTR = (527.01, 521.5, 320.25, 239.75)


def _world_from_camera(cam, pose):
    R, t = scenes.pose_matrix(pose)
    return (cam - t) @ R        # R^T (c - t), row vectors


def test_plane_at_pixel_centres_renders_its_depth():
    cols, rows = 64, 48
    fx, fy, cx, cy = INTR[0], INTR[1], 31.5, 23.25
    pose = scenes.look_at_pose((0.3, -0.2, 0.1), (0.0, 0.4, 1.5), up=(0, -1, 0))
    u, v = np.meshgrid(np.arange(cols) + 0.5, np.arange(rows) + 0.5)
    near, far = 1.25, 2.0
    pts, cols_rgb = [], []
    for z, rgb in ((far, (10, 20, 30)), (near, (200, 100, 50))):
        cam = np.stack([(u - cx) * z / fx, (v - cy) * z / fy, np.full_like(u, z)], -1).reshape(-1, 3)
        pts.append(_world_from_camera(cam, pose))
        cols_rgb.append(np.tile(np.array(rgb, np.uint8), (cam.shape[0], 1)))
    xyz = np.concatenate(pts).astype(np.float32)
    rgb = np.concatenate(cols_rgb)
    perm = np.random.default_rng(0).permutation(xyz.shape[0])
    depth, color = O.render_points(xyz[perm], pose, (fx, fy, cx, cy), (cols, rows), colors=rgb[perm])
    np.testing.assert_allclose(depth, near, rtol=1e-5)
    assert (color == np.array([200, 100, 50], np.uint8)).all()
    # wrong direction (camera->world) or swapped principal point fail this
    inv = np.r_[pose[0], -pose[1:4], 0, 0, 0]
    R, t = scenes.pose_matrix(pose)
    inv[4:] = -(R.T @ t)
    d_inv, _ = O.render_points(xyz, inv, (fx, fy, cx, cy), (cols, rows))
    assert not np.allclose(d_inv, near, rtol=1e-5)
    d_sw, _ = O.render_points(xyz, pose, (fx, fy, cy, cx), (cols, rows))
    assert (d_sw == 0).any()


def _lexsort_render(xyz, pose, intr, image, valid=None, colors=None):
    """Order-free z-buffer: per pixel the point of least z, ties to the least
    index.  Pixel coordinates in float32 in the reference's operation order,
    checked against a float64 projection."""
    R32, t32 = O.pose_matrix_f32(pose)
    p = xyz.astype(np.float32)
    cam = [((R32[k, 0] * p[:, 0] + R32[k, 1] * p[:, 1]) + R32[k, 2] * p[:, 2]) + t32[k] for k in range(3)]
    X, Y, Z = cam
    fx, fy, cx, cy = [np.float32(v) for v in intr]
    cols, rows = image
    with np.errstate(all="ignore"):
        u = (X * fx) / Z + cx
        v = (Y * fy) / Z + cy
        ok = (Z != 0) & (Z < np.finfo(np.float32).max) & (u > -1) & (u < cols) & (v > -1) & (v < rows)
    if valid is not None:
        ok &= valid.astype(bool)
    idx = np.nonzero(ok)[0]
    ix = np.trunc(u[idx]).astype(np.int64)
    iy = np.trunc(v[idx]).astype(np.int64)
    pix = iy * cols + ix
    order = np.lexsort((idx, Z[idx], pix))
    pix_s = pix[order]
    first = np.r_[True, pix_s[1:] != pix_s[:-1]]
    win = idx[order][first]
    depth = np.zeros(rows * cols, np.float32)
    color = np.zeros((rows * cols, 3), np.uint8)
    depth[pix_s[first]] = Z[win]
    color[pix_s[first]] = colors[win] if colors is not None else 255
    # float64 projection of the winners agrees with their pixel; truncation
    # toward zero gives pixel 0 the interval (-1, 1)
    R, t = scenes.pose_matrix(pose)
    c64 = xyz[win].astype(np.float64) @ R.T + t
    for k, (f, c, cell) in enumerate(((intr[0], intr[2], pix_s[first] % cols),
                                      (intr[1], intr[3], pix_s[first] // cols))):
        q = c64[:, k] * f / c64[:, 2] + c
        lo = np.where(cell == 0, -1.0, cell)
        assert np.all(q > lo - 1e-3) and np.all(q < cell + 1 + 1e-3)
    return depth.reshape(rows, cols), color.reshape(rows, cols, 3)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_selection_equals_order_free_zbuffer(seed):
    rng = np.random.default_rng(seed)
    n = 20000
    cols, rows = 40, 30
    pose = scenes.look_at_pose((0.0, 0.0, -1.0), (0.0, 0.0, 1.0), up=(0, -1, 0))
    xyz = rng.uniform([-1.0, -0.8, -0.2], [1.0, 0.8, 2.0], (n, 3)).astype(np.float32)
    # some points behind the camera (camera z in (-1, 0)): the reference keeps them
    xyz[:300] = rng.uniform([-0.3, -0.2, -2.0], [0.3, 0.2, -1.05], (300, 3))
    # duplicated positions with different colours: ties go to the lower index
    dup = rng.integers(0, n, 3000)
    xyz[rng.integers(0, n, 3000)] = xyz[dup]
    colors = rng.integers(0, 256, (n, 3)).astype(np.uint8)
    valid = (rng.random(n) > 0.1).astype(np.uint8)
    intr = (60.0, 55.0, 20.0, 15.0)
    d, c = O.render_points(xyz, pose, intr, (cols, rows), valid=valid, colors=colors)
    de, ce = _lexsort_render(xyz, pose, intr, (cols, rows), valid=valid, colors=colors)
    assert np.array_equal(d.view(np.uint32), de.view(np.uint32))
    assert np.array_equal(c, ce)
    assert (d < 0).any() and (d > 0).any()      # points behind the camera are kept (quirk)


def test_edge_rules():
    intr = (10.0, 10.0, 0.0, 0.0)
    ident = [1, 0, 0, 0, 0, 0, 0]
    fmax = np.finfo(np.float32).max
    pts = np.array([
        [-0.05, 0.0, 1.0],     # u = -0.5 -> column 0 (truncation toward zero)
        [-0.1, 0.2, 1.0],      # u = -1 -> off the image
        [0.39, 0.2, 2.0],      # u = 1.95 -> column 1
        [0.0, 0.1, 0.0],       # z == 0 -> skipped
        [0.0, 0.3, fmax],      # z == FLT_MAX never beats the fill
        [0.0, 0.0, np.nan],    # NaN skipped
        [0.2, 0.2, 1.0],       # u = 2, v = 2 -> off a 2-high image
        [0.3, 0.0, 1.0],       # u = 3 -> off a 3-wide image
    ], np.float32)
    d, c = O.render_points(pts, ident, intr, (3, 2), colors=np.arange(24, dtype=np.uint8).reshape(8, 3))
    expect = np.zeros((2, 3), np.float32)
    expect[0, 0] = 1.0         # point 0
    expect[1, 1] = 2.0         # point 2 (v = 1.0 -> row 1); point 6 has v = 2, off a 2-high image
    assert np.array_equal(d, expect), d
    assert list(c[0, 0]) == [0, 1, 2] and list(c[1, 1]) == [6, 7, 8] and (c[0, 2] == 0).all()
    # a point exactly at the principal point behind the camera renders with negative depth
    d2, _ = O.render_points(np.array([[0, 0, -3.0]], np.float32), ident, (5.0, 5.0, 1.0, 1.0), (3, 3))
    assert d2[1, 1] == -3.0 and np.count_nonzero(d2) == 1
    # empty cloud: all zeros
    d3, c3 = O.render_points(np.zeros((0, 3), np.float32), ident, intr, (4, 4))
    assert (d3 == 0).all() and (c3 == 0).all()

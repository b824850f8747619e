"""Parity of the HIP depth->point-cloud kernel with the CPU oracle of
point_cloud_extract (cuda_point_cloud.cu:22-70): bit-exact positions,
validity and colours."""
import numpy as np
import pytest

from helpers import assert_bitexact

import oracle_ffi as O
import recon
from recon import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("image", [(640, 480), (37, 23), (1280, 720)])
def test_points_exact(cuda, image):
    w = scenes.plain_small()
    w.image = image
    fx = 0.82 * image[0]
    w.intrinsics = (fx, fx * 1.01, image[0] / 2 - 0.5, image[1] / 2 - 0.5)
    depth = w.render(frames=[3])[0]
    depth[::7, ::5] = -0.5
    rng = np.random.default_rng(1)
    rgb = rng.integers(0, 256, (image[1], image[0], 3), dtype=np.uint8)
    dev = cuda.from_numpy(depth).cuda()
    rgb_dev = cuda.from_numpy(rgb).cuda()
    T = w.poses[3]
    xyz, valid, col = recon.depth_to_points(dev, w.intrinsics, T, rgb=rgb_dev, with_colors=True)
    oxyz, ovalid, ocol = O.depth_to_points(depth, w.intrinsics, T, rgb=rgb)
    assert np.array_equal(valid.cpu().numpy(), ovalid)
    assert_bitexact(xyz.cpu().numpy(), oxyz, "xyz")
    assert np.array_equal(col.cpu().numpy(), ocol)
    # white when no colour image is given
    _, _, col2 = recon.depth_to_points(dev, w.intrinsics, T, with_colors=True)
    c2 = col2.cpu().numpy()
    v = ovalid.astype(bool)
    assert (c2[v] == 255).all() and (c2[~v] == 0).all()


def test_points_pitched(cuda):
    w = scenes.tiny()
    depth = w.render(frames=[0])[0]
    H, W = depth.shape
    buf = cuda.zeros((H, W + 7), dtype=cuda.float32, device="cuda")
    buf[:, :W] = cuda.from_numpy(depth).cuda()
    xyz, valid = recon.depth_to_points(buf[:, :W], w.intrinsics, w.poses[0])
    oxyz, ovalid, _ = O.depth_to_points(depth, w.intrinsics, w.poses[0])
    assert np.array_equal(valid.cpu().numpy(), ovalid)
    assert_bitexact(xyz.cpu().numpy(), oxyz, "xyz")

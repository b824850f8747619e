"""Parity of the HIP z-buffer render (recon_points_render, replacing
cudaPointCloud::project_to_depth / project_to_color_depth and
filter_image_if, cuda_point_cloud.cu:229-353, 89-99) with the CPU oracle:
depth bits and colours identical, for clouds extracted from rendered depth
(many points per pixel), random clouds with the edge values the reference's
loop meets (z == 0, negative z, FLT_MAX, inf, NaN, duplicates), pitched
outputs, masks and an empty cloud."""
import numpy as np
import pytest

import oracle_ffi as O
import recon
from helpers import assert_bitexact
from recon import scenes

pytestmark = pytest.mark.gpu


def _inverse(pose):
    R, t = scenes.pose_matrix(pose)
    return np.r_[pose[0], -np.asarray(pose[1:4]), -(R.T @ t)]


def _check(torch, xyz, pose, intr, image, valid=None, colors=None, pitch_pad=0, what=""):
    cols, rows = image
    dx = torch.from_numpy(np.ascontiguousarray(xyz, np.float32)).cuda()
    dv = torch.from_numpy(valid).cuda() if valid is not None else None
    dc = torch.from_numpy(colors).cuda() if colors is not None else None
    dbuf = torch.full((rows, cols + pitch_pad), 7.0, dtype=torch.float32, device="cuda")
    cbuf = torch.full((rows, cols + pitch_pad, 3), 7, dtype=torch.uint8, device="cuda")
    d, c = recon.render_points(dx, pose, intr, image, valid=dv, colors=dc, with_colors=True,
                               depth_out=dbuf[:, :cols], color_out=cbuf[:, :cols])
    torch.cuda.synchronize()
    od, oc = O.render_points(xyz, pose, intr, image, valid=valid, colors=colors)
    assert_bitexact(d.cpu().numpy(), od, f"{what} depth")
    assert np.array_equal(c.cpu().numpy(), oc), f"{what} colour"
    if pitch_pad:  # the padding is untouched
        assert (dbuf[:, cols:] == 7.0).all() and (cbuf[:, cols:] == 7).all()
    return od


def test_extracted_cloud(cuda):
    """Five frames back-projected on the device, rendered from a sixth pose."""
    w = scenes.plain_small(num_frames=6)
    depth = w.render()
    cols, rows = w.image
    rng = np.random.default_rng(5)
    rgb = rng.integers(0, 256, (rows, cols, 3)).astype(np.uint8)
    xyz, valid, colors = [], [], []
    for f in range(5):
        p, v, c = recon.depth_to_points(cuda.from_numpy(depth[f]).cuda(), w.intrinsics,
                                        _inverse(w.poses[f]), rgb=cuda.from_numpy(rgb).cuda(),
                                        with_colors=True)
        xyz.append(p.cpu().numpy())
        valid.append(v.cpu().numpy())
        colors.append(c.cpu().numpy())
    xyz, valid, colors = np.concatenate(xyz), np.concatenate(valid), np.concatenate(colors)
    print(f"render {xyz.shape[0]} extracted points", flush=True)
    od = _check(cuda, xyz, w.poses[5], w.intrinsics, w.image, valid=valid, colors=colors,
                what="extracted")
    assert (od > 0).mean() > 0.3
    # without mask and colours; the invalid points sit at the world origin
    _check(cuda, xyz, w.poses[5], w.intrinsics, w.image, what="unmasked")
    # the frame's own pose: most pixels return to their depth (truncation
    # moves a pixel by at most one)
    od0 = _check(cuda, xyz[:rows * cols], w.poses[0], w.intrinsics, w.image,
                 valid=valid[:rows * cols], pitch_pad=13, what="own pose")
    m = (depth[0] > 0) & (od0 > 0)
    assert np.mean(np.abs(od0[m] - depth[0][m]) < 1e-5) > 0.4


@pytest.mark.parametrize("seed", [11, 12])
def test_random_cloud_edge_values(cuda, seed):
    rng = np.random.default_rng(seed)
    n = 200000
    cols, rows = 97, 61
    pose = scenes.look_at_pose((0.1, -0.05, -1.0), (0.0, 0.0, 1.0), up=(0, -1, 0))
    xyz = rng.uniform([-1.0, -0.8, -0.3], [1.0, 0.8, 2.0], (n, 3)).astype(np.float32)
    xyz[:2000] = rng.uniform([-0.3, -0.2, -2.0], [0.3, 0.2, -1.05], (2000, 3))  # behind
    xyz[rng.integers(0, n, 20000)] = xyz[rng.integers(0, n, 20000)]             # duplicates
    fmax = np.finfo(np.float32).max
    special = rng.integers(0, n, 60)
    xyz[special[0:10], 2] = np.nan
    xyz[special[10:20], 0] = np.inf
    xyz[special[20:30], 2] = np.inf
    xyz[special[30:40], 2] = -np.inf
    xyz[special[40:50], 2] = fmax
    xyz[special[50:60]] = 0.0
    colors = rng.integers(0, 256, (n, 3)).astype(np.uint8)
    valid = (rng.random(n) > 0.05).astype(np.uint8)
    intr = (120.0, 110.0, 48.3, 30.1)
    print(f"render random seed {seed}", flush=True)
    od = _check(cuda, xyz, pose, intr, (cols, rows), valid=valid, colors=colors, pitch_pad=5,
                what="random")
    assert (od < 0).any() and (od > 0).any()
    # points whose camera z is exactly 0 (on the camera plane): skipped
    ident = [1, 0, 0, 0, 0, 0, 0]
    plane = rng.uniform(-1, 1, (1000, 3)).astype(np.float32)
    plane[:, 2] = 0.0
    _check(cuda, plane, ident, intr, (cols, rows), what="z == 0")


def test_empty_and_repeated_calls(cuda):
    intr = (100.0, 100.0, 16.0, 12.0)
    ident = [1, 0, 0, 0, 0, 0, 0]
    zb = cuda.empty(32 * 24, dtype=cuda.int64, device="cuda")
    empty = cuda.zeros((0, 3), dtype=cuda.float32, device="cuda")
    d, c = recon.render_points(empty, ident, intr, (32, 24), with_colors=True, zbuffer=zb)
    cuda.cuda.synchronize()
    assert (d == 0).all() and (c == 0).all()
    pts = np.array([[0.0, 0.0, 1.0], [0.01, 0.0, 2.0]], np.float32)
    d1 = recon.render_points(cuda.from_numpy(pts).cuda(), ident, intr, (32, 24), zbuffer=zb)
    cuda.cuda.synchronize()
    od, _ = O.render_points(pts, ident, intr, (32, 24), want_color=False)
    assert_bitexact(d1.cpu().numpy(), od, "two points")
    # a second call with fewer points does not see the first call's keys
    d2 = recon.render_points(cuda.from_numpy(pts[1:]).cuda(), ident, intr, (32, 24), zbuffer=zb)
    cuda.cuda.synchronize()
    od2, _ = O.render_points(pts[1:], ident, intr, (32, 24), want_color=False)
    assert_bitexact(d2.cpu().numpy(), od2, "one point")

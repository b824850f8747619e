"""Parity of the HIP integrate kernel (librecon.so) with the CPU oracle.

Bar: D and W bit-exact with the float32 restatement of cu_tsdf_integrate
(tsdf_volume.cu:144-190 under the build contract), D within 1e-6 absolute of
the float64 decision-matched oracle, update counts equal.
"""
import numpy as np
import pytest

from helpers import assert_bitexact, gpu_run, oracle_run

from recon import scenes

pytestmark = pytest.mark.gpu

F64_TOL = 1e-6  # absolute, metres, for volumes within a few metres of the origin


def _check(w, depth, fused=False, pitch_pad=0, f64=True, f64_tol=F64_TOL):
    cfg = w.config()
    o = oracle_run(cfg, depth, w.poses, f64=f64)
    g = gpu_run(cfg, depth, w.poses, fused=fused, pitch_pad=pitch_pad, count=True)
    assert_bitexact(g["W"], o["W"], "W")
    assert_bitexact(g["D"], o["D"], "D")
    if fused:
        assert int(g["counts"][0]) == int(sum(o["counts"]))
    else:
        assert g["counts"].tolist() == [int(c) for c in o["counts"]]
    if f64:
        err = np.abs(g["D"].astype(np.float64) - o["D64"]).max()
        assert err <= f64_tol, err
        assert np.array_equal(g["W"].astype(np.float64), o["W64"])
    g["vol"].close()
    return o


def test_tiny_exact(cuda):
    w = scenes.tiny()
    o = _check(w, w.render())
    assert sum(o["counts"]) > 0


def test_plain_small_exact(cuda):
    w = scenes.plain_small()
    o = _check(w, w.render())
    assert min(o["counts"]) > 100000


def test_plain_small_fused_equals_per_frame(cuda):
    w = scenes.plain_small()
    _check(w, w.render(), fused=True, f64=False)


@pytest.mark.parametrize("grid", [(37, 29, 23), (66, 8, 9), (4, 4, 4), (1, 1, 1), (130, 5, 7)])
def test_ragged_grids(cuda, grid):
    # gx not a multiple of 4 takes the scalar path; gy, gz not multiples of 4
    w = scenes.tiny(grid=grid, voxel_size=0.6 / max(grid), image=(40, 30), num_frames=4)
    _check(w, w.render())


def test_pitched_depth(cuda):
    w = scenes.tiny(grid=(48, 40, 32), voxel_size=0.02, image=(50, 38), num_frames=3)
    _check(w, w.render(), pitch_pad=13)


def test_weight_cap_and_many_frames(cuda):
    # more frames than max_weight (2): W saturates, D keeps averaging with (W+1)
    w = scenes.tiny(num_frames=20)
    w.poses = np.repeat(w.poses[:2], 10, axis=0)
    o = _check(w, w.render())
    assert o["W"].max() == 2.0


def test_fused_more_than_16_frames(cuda):
    w = scenes.tiny(grid=(32, 32, 32), voxel_size=0.02, num_frames=37)
    _check(w, w.render(), fused=True, f64=False)


def test_invalid_depth_values(cuda):
    w = scenes.tiny(grid=(32, 24, 20), voxel_size=0.03, num_frames=3)
    depth = w.render()
    depth[0, ::3, ::2] = np.nan
    depth[1, 1::4] = np.inf
    depth[2, :, :5] = -1.0
    depth[2, :, 5:9] = 30.0  # beyond max_depth 24
    _check(w, depth, f64=False)


def test_empty_depth_leaves_volume(cuda):
    w = scenes.tiny(num_frames=2)
    depth = np.zeros((2, w.image[1], w.image[0]), np.float32)
    o = _check(w, depth)
    assert sum(o["counts"]) == 0
    assert (o["W"] == 0).all()


def _random_pose_workload(seed, grid=(96, 80, 72), vs=0.01, offset=None, inside=False):
    """Cameras at random positions (inside the volume when `inside`) looking
    at random points of the volume, random-valued depth maps."""
    rng = np.random.default_rng(seed)
    w = scenes.tiny(grid=grid, voxel_size=vs, image=(160, 120), num_frames=6)
    if offset is not None:
        w.offset = offset
    lo = np.array(w.offset)
    hi = lo + np.array(grid) * vs
    poses = []
    for _ in range(w.num_frames):
        if inside:
            c = lo + rng.uniform(0.2, 0.8, 3) * (hi - lo)
        else:
            c = (lo + hi) / 2 + rng.normal(0, 1, 3) * (hi - lo)
        target = lo + rng.uniform(0, 1, 3) * (hi - lo)
        poses.append(scenes.look_at_pose(c, target, up=(0.3, 0.2, 1.0)))
    w.poses = np.array(poses)
    depth = rng.uniform(0.0, 2.0 * np.linalg.norm(hi - lo), (w.num_frames, 120, 160)).astype(np.float32)
    depth[rng.random(depth.shape) < 0.3] = 0.0
    # patches of constant depth exercise the block bounds
    depth[:, 20:60, 40:100] = rng.uniform(0.05, 1.0)
    return w, depth


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_cull_random_poses_outside(cuda, seed):
    w, depth = _random_pose_workload(seed)
    _check(w, depth)


@pytest.mark.parametrize("seed", [4, 5])
def test_cull_camera_inside_volume(cuda, seed):
    w, depth = _random_pose_workload(seed, inside=True)
    _check(w, depth)


def test_cull_far_from_origin(cuda):
    # 70 m from the origin a float32 ulp is 7.6e-6 m, so the float64 bound
    # scales with the coordinates: 4 ulp of the largest coordinate
    w, depth = _random_pose_workload(6, vs=0.02, offset=(50.0, -70.0, 30.0), inside=True)
    _check(w, depth, f64_tol=4 * 2.0 ** -23 * 72.0)
    w, depth = _random_pose_workload(7, vs=0.02, offset=(50.0, -70.0, 30.0))
    _check(w, depth, fused=True, f64=False)


def test_cull_fused_random(cuda):
    w, depth = _random_pose_workload(8, inside=True)
    w.poses = np.concatenate([w.poses] * 4)
    depth = np.concatenate([depth] * 4)
    _check(w, depth, fused=True, f64=False)


@pytest.mark.parametrize("gz", [256, 258])
def test_cull_segment_limits(cuda, gz):
    # 2048 x 1024 x gz voxels (4 GB of D and W): 64 x 128 x gz/2 wave-tiles of
    # 32 x 8 x 2.  gz = 256 gives 4096 x 256 tiles, the most cull segments of
    # 256 tiles; gz = 258 gives more, so segments of 512 tiles.  Per-frame and
    # fused calls, bit-exact, update counts equal.
    w = scenes.tiny(grid=(2048, 1024, gz), voxel_size=0.001, image=(320, 240), num_frames=2)
    cfg = w.config()
    depth = w.render()
    o = oracle_run(cfg, depth, w.poses)
    assert min(o["counts"]) > 100_000_000
    for fused in (False, True):
        g = gpu_run(cfg, depth, w.poses, fused=fused, count=True)
        assert_bitexact(g["W"], o["W"], "W")
        assert_bitexact(g["D"], o["D"], "D")
        if fused:
            assert int(g["counts"][0]) == int(sum(o["counts"]))
        else:
            assert g["counts"].tolist() == [int(c) for c in o["counts"]]
        g["vol"].close()
        del g


def test_touch_counter_is_union_per_sweep(cuda):
    """recon_tsdf_set_touch_counter: per-frame calls count each updated voxel
    once (= the update count); a fused call counts, per sweep of 16 frames,
    the voxels that any of its frames updates (the oracle's masks OR-ed)."""
    import recon
    import oracle_ffi as O

    w = scenes.tiny(grid=(40, 36, 28), voxel_size=0.025, image=(64, 48), num_frames=21)
    depth = w.render()
    cfg = w.config()
    ov = O.volume_from_config(cfg)
    D, W = O.init(ov)
    masks = []
    for f in range(w.num_frames):
        n, m = O.integrate_f32(ov, D, W, depth[f], w.poses[f], mask=True)
        assert int((m != 0).sum()) == n
        masks.append(m != 0)
    dev = cuda.from_numpy(depth).cuda()
    vol = recon.TSDFVolume(cfg)
    t = cuda.zeros(w.num_frames, dtype=cuda.int64, device="cuda")
    for f in range(w.num_frames):
        vol.set_touch_counter(t[f:f + 1])
        vol.integrate(dev[f], w.poses[f])
    vol.set_touch_counter(None)
    vol.reset()
    u = cuda.zeros(1, dtype=cuda.int64, device="cuda")
    vol.set_touch_counter(u)
    vol.integrate_frames(dev, w.poses)
    vol.set_touch_counter(None)
    cuda.cuda.synchronize()
    assert t.cpu().numpy().tolist() == [int(m.sum()) for m in masks]
    union = sum(int(np.logical_or.reduce(masks[a:a + 16]).sum()) for a in range(0, w.num_frames, 16))
    assert int(u.item()) == union > 0
    Dg, Wg = vol.download()
    assert_bitexact(Wg, W, "fused W")
    assert_bitexact(Dg, D, "fused D")
    vol.close()

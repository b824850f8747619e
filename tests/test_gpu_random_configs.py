"""A sweep of random volume and camera configurations through every path,
each checked against the oracle: per-frame, fused and colour integrate
(D, W, colour and update counts bit-exact; D within 1e-6 of the float64
oracle), the triangle soup and the welded mesh (bit-exact).  Grids of 1-96
voxels per axis (not multiples of the tile or chunk sizes), anisotropic
voxels, offsets off the origin, images of 16-200 x 12-160 pixels with
off-centre principal points, 1-5 frames, weight caps from 1.5 to 128.
The same sweep runs WEIGHT_BY_VARIANCE configurations, unbounded sparse
brick volumes, the differentiable forward, and depth -> points with the
z-buffer render of the resulting cloud."""
import numpy as np
import pytest

import oracle_ffi as O
import recon
from helpers import assert_bitexact, gpu_run, oracle_run
from recon import scenes

pytestmark = pytest.mark.gpu


def _case(seed):
    rng = np.random.default_rng(7000 + seed)
    grid = tuple(int(v) for v in rng.integers(1, 97, 3))
    vs = tuple(float(v) for v in rng.uniform(0.004, 0.03, 3))
    ext = np.array(grid) * np.array(vs)
    offset = tuple(float(v) for v in rng.uniform(-0.3, 0.3, 3) - ext / 2)
    cols, rows = int(rng.integers(16, 201)), int(rng.integers(12, 161))
    f = float(rng.uniform(0.6, 1.4)) * cols
    intr = (f, f * float(rng.uniform(0.9, 1.1)), cols / 2 - 0.5 + float(rng.uniform(-5, 5)),
            rows / 2 - 0.5 + float(rng.uniform(-5, 5)))
    nf = int(rng.integers(1, 6))
    poses = scenes.arc_poses(nf, step_deg=float(rng.uniform(5, 40)), radius=float(rng.uniform(0.6, 1.5)),
                             height=float(rng.uniform(-0.3, 0.5)), start_deg=float(rng.uniform(0, 360)))
    trunc = float(rng.uniform(1.5, 5.0)) * max(vs)
    maxw = float(rng.choice([1.5, 2.0, 4.0, 128.0]))
    prims = [(scenes.PRIM_SPHERE, tuple(float(v) for v in rng.uniform(-0.1, 0.1, 3)) +
              (float(rng.uniform(0.1, 0.3)),)),
             (scenes.PRIM_PLANE_Z, (float(rng.uniform(-0.3, -0.1)),))]
    w = scenes.Workload(name=f"random_{seed}", grid=grid, voxel_size=vs[0], offset=offset,
                        truncation=trunc, max_weight=maxw, intrinsics=intr, image=(cols, rows),
                        prims=prims, poses=poses, noise=(0.002, 0.0, 0.05, 100 + seed))
    cfg = recon.make_config(grid, voxel_size=vs, offset=offset, max_weight=maxw, intrinsics=intr,
                            image=(cols, rows), truncation=trunc)
    return w, cfg


@pytest.mark.parametrize("seed", range(16))
def test_random_configuration(cuda, seed):
    w, cfg = _case(seed)
    depth = w.render()
    print(f"random config {seed}: grid {w.grid} image {w.image} frames {w.num_frames}", flush=True)
    o = oracle_run(cfg, depth, w.poses, f64=True)
    err = np.abs(o["D"].astype(np.float64) - o["D64"]).max() if o["D"].size else 0.0
    assert err <= 1e-6, err
    # per-frame and fused integrate
    for fused in (False, True):
        g = gpu_run(cfg, depth, w.poses, fused=fused, count=True)
        assert_bitexact(g["W"], o["W"], f"W fused={fused}")
        assert_bitexact(g["D"], o["D"], f"D fused={fused}")
        total = int(sum(o["counts"]))
        if fused:
            assert int(g["counts"][0]) == total
        else:
            assert g["counts"].tolist() == [int(c) for c in o["counts"]]
        g["vol"].close()
    # colour integrate
    rng = np.random.default_rng(seed)
    cols, rows = w.image
    rgbs = rng.integers(0, 256, (w.num_frames, rows, cols, 3), dtype=np.uint8)
    ov = O.volume_from_config(cfg)
    Dc, Wc = O.init(ov)
    Cc = np.zeros(Dc.shape + (4,), np.float32)
    for f in range(w.num_frames):
        O.integrate_color_f32(ov, Dc, Wc, Cc, depth[f], rgbs[f], w.poses[f])
    vol = recon.TSDFVolume(cfg)
    vol.enable_color(True)
    dev = cuda.from_numpy(np.ascontiguousarray(depth)).cuda()
    drgb = cuda.from_numpy(rgbs).cuda()
    for f in range(w.num_frames):
        vol.integrate(dev[f], w.poses[f], rgb=drgb[f])
    cuda.cuda.synchronize()
    D, W = vol.download()
    assert_bitexact(D, Dc, "colour D")
    assert_bitexact(W, Wc, "colour W")
    assert_bitexact(vol.download_colors(), Cc, "colour")
    vol.close()
    # meshes of the integrated field
    vsz = list(ov.voxel_size)
    off = list(ov.offset)
    field = cuda.from_numpy(np.ascontiguousarray(o["D"])).cuda()
    verts, tris = recon.mc_extract(field, w.grid, vsz, off)
    ovv, ott, _, _ = O.mc_extract(o["D"], w.grid, vsz, off)
    assert verts.shape[0] == ovv.shape[0]
    assert_bitexact(verts.cpu().numpy().reshape(-1, 9), ovv.reshape(-1, 9), "soup vertices")
    assert np.array_equal(tris.cpu().numpy(), ott)
    v, n, t = recon.mc_extract_welded(field, w.grid, vsz, off)
    wv, wn, wt = O.mc_weld(o["D"], w.grid, vsz, off)
    assert v.shape[0] == wv.shape[0] and t.shape[0] == wt.shape[0]
    assert_bitexact(v.cpu().numpy(), wv, "welded vertices")
    assert_bitexact(n.cpu().numpy(), wn, "welded normals")
    assert np.array_equal(t.cpu().numpy(), wt)


@pytest.mark.parametrize("seed", range(8))
def test_random_configuration_variance_weights(cuda, seed):
    # WEIGHT_BY_VARIANCE on the same sweep: per-frame, fused and colour
    # integrate bit-exact with the oracle; a random noise model, or the
    # library's default (all zeros) on odd seeds
    w, cfg = _case(100 + seed)
    rng = np.random.default_rng(9000 + seed)
    cfg.weight_type = 1  # RECON_WEIGHT_BY_VARIANCE
    if seed % 2 == 0:
        nm = (float(rng.uniform(0.0005, 0.003)), float(rng.uniform(0.0005, 0.004)),
              float(rng.uniform(0.2, 0.6)))
        for i in range(3):
            cfg.noise_model[i] = nm[i]
    depth = w.render()
    print(f"variance weights {seed}: grid {w.grid} image {w.image} frames {w.num_frames}", flush=True)
    o = oracle_run(cfg, depth, w.poses)
    for fused in (False, True):
        g = gpu_run(cfg, depth, w.poses, fused=fused)
        assert_bitexact(g["W"], o["W"], f"W fused={fused}")
        assert_bitexact(g["D"], o["D"], f"D fused={fused}")
        g["vol"].close()
    cols, rows = w.image
    rgbs = rng.integers(0, 256, (w.num_frames, rows, cols, 3), dtype=np.uint8)
    ov = O.volume_from_config(cfg)
    Dc, Wc = O.init(ov)
    Cc = np.zeros(Dc.shape + (4,), np.float32)
    for f in range(w.num_frames):
        O.integrate_color_f32(ov, Dc, Wc, Cc, depth[f], rgbs[f], w.poses[f])
    vol = recon.TSDFVolume(cfg)
    vol.enable_color(True)
    dev = cuda.from_numpy(np.ascontiguousarray(depth)).cuda()
    drgb = cuda.from_numpy(rgbs).cuda()
    for f in range(w.num_frames):
        vol.integrate(dev[f], w.poses[f], rgb=drgb[f])
    cuda.cuda.synchronize()
    D, W = vol.download()
    assert_bitexact(D, Dc, "colour D")
    assert_bitexact(W, Wc, "colour W")
    assert_bitexact(vol.download_colors(), Cc, "colour")
    vol.close()


@pytest.mark.parametrize("seed", range(8))
def test_random_configuration_sparse(cuda, seed):
    # sparse brick volume on the same sweep (unbounded, so bricks reach
    # negative coordinates and far outside the dense grid): the same brick set
    # and bit-identical brick contents as the oracle after every frame
    w, cfg = _case(200 + seed)
    depth = w.render()
    print(f"sparse {seed}: grid {w.grid} image {w.image} frames {w.num_frames}", flush=True)
    sv = recon.SparseVolume(cfg, 1 << 16)
    so = O.Sparse(O.volume_from_config(cfg))
    dev = cuda.from_numpy(np.ascontiguousarray(depth)).cuda()
    for f in range(w.num_frames):
        sv.integrate(dev[f], w.poses[f])
        so.integrate(depth[f], w.poses[f])
        cuda.cuda.synchronize()
        g, o = sv.bricks(), so.bricks()
        assert set(g) == set(o), f"frame {f}: {len(g)} vs {len(o)} bricks"
        for k, (od, ow) in o.items():
            gd, gw = g[k]
            assert np.array_equal(gd.view(np.uint32), od.view(np.uint32)), f"D of brick {k}"
            assert np.array_equal(gw.view(np.uint32), ow.view(np.uint32)), f"W of brick {k}"
    assert sv.stats()[1] == 0
    sv.close()


@pytest.mark.parametrize("seed", range(8))
def test_random_configuration_differentiable_forward(cuda, seed):
    # the differentiable (bilinear) integrate on the same sweep: D and W
    # bit-exact with the float32 oracle of the forward
    w, cfg = _case(300 + seed)
    depth = w.render()
    print(f"differentiable {seed}: grid {w.grid} image {w.image} frames {w.num_frames}", flush=True)
    rt32 = recon.pose_to_matrix(w.poses)
    vol = recon.TSDFVolume(cfg)
    vol.diff_integrate(cuda.from_numpy(np.ascontiguousarray(depth)).cuda(), cuda.from_numpy(rt32).cuda())
    cuda.cuda.synchronize()
    D, W = vol.download()
    D32, W32 = O.diff_forward_f32(O.volume_from_config(cfg), depth, rt32)
    assert_bitexact(W, W32, "W")
    assert_bitexact(D, D32, "D")
    vol.close()


def _camera_to_world(pose):
    """The inverse (camera->world) of a world->camera pose (qw, qx, qy, qz, t)."""
    R, t = scenes.pose_matrix(pose)
    return np.r_[pose[0], -np.asarray(pose[1:4]), -(R.T @ t)]


@pytest.mark.parametrize("seed", range(8))
def test_random_configuration_points_and_render(cuda, seed):
    # depth -> points of every frame (with colours) and the z-buffer render of
    # the whole cloud at another pose of the arc: bit-exact with the oracle
    w, _ = _case(400 + seed)
    depth = w.render()
    cols, rows = w.image
    rng = np.random.default_rng(seed)
    rgb = rng.integers(0, 256, (rows, cols, 3), dtype=np.uint8)
    print(f"points and render {seed}: image {w.image} frames {w.num_frames}", flush=True)
    xyz, valid, colors = [], [], []
    for f in range(w.num_frames):
        T = _camera_to_world(w.poses[f])
        p, v, c = recon.depth_to_points(cuda.from_numpy(depth[f]).cuda(), w.intrinsics, T,
                                        rgb=cuda.from_numpy(rgb).cuda(), with_colors=True)
        op, ovl, oc = O.depth_to_points(depth[f], w.intrinsics, T, rgb=rgb)
        assert np.array_equal(v.cpu().numpy(), ovl)
        assert_bitexact(p.cpu().numpy(), op, "points")
        assert np.array_equal(c.cpu().numpy(), oc)
        xyz.append(op)
        valid.append(ovl)
        colors.append(oc)
    xyz, valid, colors = np.concatenate(xyz), np.concatenate(valid), np.concatenate(colors)
    view = scenes.arc_poses(1, radius=1.1, height=0.2, start_deg=float(rng.uniform(0, 360)))[0]
    d, c = recon.render_points(cuda.from_numpy(xyz).cuda(), view, w.intrinsics, w.image,
                               valid=cuda.from_numpy(valid).cuda(), colors=cuda.from_numpy(colors).cuda(),
                               with_colors=True)
    od, oc = O.render_points(xyz, view, w.intrinsics, w.image, valid=valid, colors=colors)
    assert_bitexact(d.cpu().numpy(), od, "rendered depth")
    assert np.array_equal(c.cpu().numpy(), oc)

"""Pins the oracle's colour integrate and colour sampler (no reference
counterpart on the dense path; the octree's add_observation,
octree.cu:118-137, is the running average they follow) with closed forms:
colour leaves D and W untouched; a constant image gives exactly that colour;
two frames of constant colours a and b give a, b or (a + b) / 2 exactly, as
the colour weight says; only band updates (sdf <= trunc) take colour.  The
sampler returns a voxel's own colour at its centre, a constant field's
colour everywhere, the linear interpolation of a linear field, and ignores
neighbours of colour weight 0."""
import numpy as np

import oracle_ffi as O
from recon import scenes


def _setup(frames=2):
    w = scenes.plain_small(num_frames=frames, grid=64, voxel_size=0.016)
    vol = O.volume_from_config(w.config())
    return w, vol, w.render()


def _color_run(vol, depth, poses, rgbs):
    D, W = O.init(vol)
    Cv = np.zeros(D.shape + (4,), np.float32)
    for d, p, c in zip(depth, poses, rgbs):
        O.integrate_color_f32(vol, D, W, Cv, d, c, p)
    return D, W, Cv


def test_color_leaves_tsdf_and_takes_constant_colours():
    w, vol, depth = _setup(frames=2)
    rows, cols = depth.shape[1:]
    a, b = np.array([10, 200, 77], np.uint8), np.array([255, 0, 31], np.uint8)
    rgbs = [np.broadcast_to(a, (rows, cols, 3)), np.broadcast_to(b, (rows, cols, 3))]
    D, W, Cv = _color_run(vol, depth, w.poses, rgbs)
    D0, W0 = O.init(vol)
    for d, p in zip(depth, w.poses):
        O.integrate_f32(vol, D0, W0, d, p)
    assert np.array_equal(D.view(np.uint32), D0.view(np.uint32))
    assert np.array_equal(W.view(np.uint32), W0.view(np.uint32))
    wc = Cv[..., 3]
    assert set(np.unique(wc)) <= {0.0, 1.0, 2.0} and (wc == 2).sum() > 1000
    assert (wc <= W).all()
    rgb = Cv[..., :3]
    both = wc == 2
    assert np.array_equal(rgb[both], np.broadcast_to((a.astype(np.float32) + b) / 2, rgb[both].shape))
    one = wc == 1
    is_a = (rgb[one] == a).all(axis=1)
    is_b = (rgb[one] == b).all(axis=1)
    assert (is_a | is_b).all() and is_a.any() and is_b.any()
    assert (rgb[wc == 0] == 0).all()


def test_only_band_updates_take_colour():
    w, vol, depth = _setup(frames=1)
    rows, cols = depth.shape[1:]
    rgb = np.full((rows, cols, 3), 9, np.uint8)
    D, W, Cv = _color_run(vol, depth, w.poses, [rgb])
    # after one frame from the fill, D = tsdf: band voxels (sdf <= trunc)
    # have D < trunc, except sdf == trunc exactly; free space has D = trunc
    updated = W > 0
    coloured = Cv[..., 3] > 0
    band = updated & (D < vol.truncation)
    assert coloured[band].all()
    assert not coloured[~updated].any()
    extra = coloured & ~band
    assert extra.sum() <= 0.001 * coloured.sum()
    assert (updated & ~coloured).sum() > coloured.sum()   # free space is most of it


def test_sampler_closed_forms():
    vol = O.volume((10, 9, 8), np.array([0.1, 0.2, 0.15], np.float32), [-0.5, 0.1, 0.3], 0.3, 64.0,
                   (500.0, 500.0, 320.0, 240.0), 24.0)
    g = (8, 9, 10)                      # numpy order (gz, gy, gx)
    rng = np.random.default_rng(0)
    vs, off = np.array([0.1, 0.2, 0.15], np.float32), np.array([-0.5, 0.1, 0.3], np.float32)
    # voxel centres sample their own colour
    Cv = np.zeros(g + (4,), np.float32)
    Cv[..., :3] = rng.integers(0, 256, g + (3,))
    Cv[..., 3] = 1.0
    idx = np.stack(np.meshgrid(np.arange(1, 9), np.arange(1, 8), np.arange(1, 7), indexing="ij"), -1).reshape(-1, 3)
    centres = ((idx.astype(np.float32) + np.float32(0.5)) * vs) + off
    s = O.sample_colors_f32(vol, Cv, centres)
    assert np.array_equal(s, Cv[idx[:, 2], idx[:, 1], idx[:, 0], :3].astype(np.uint8))
    # constant field everywhere inside; no support outside
    Cv[..., :3] = np.array([17, 128, 250], np.float32)
    lo, hi = off + vs * 0.6, off + vs * np.array([9.3, 8.3, 7.3], np.float32)
    pts = rng.uniform(lo, hi, (500, 3)).astype(np.float32)
    assert (O.sample_colors_f32(vol, Cv, pts) == [17, 128, 250]).all()
    far = np.array([[10.0, 10.0, 10.0], [np.nan, 0.5, 0.5]], np.float32)
    assert (O.sample_colors_f32(vol, Cv, far) == 0).all()
    # linear in x: r = 3 * ix + 10 interpolates linearly
    Cv[..., 0] = 3.0 * np.arange(10, dtype=np.float32)[None, None, :] + 10.0
    s = O.sample_colors_f32(vol, Cv, pts).astype(np.float64)
    ix = (pts[:, 0].astype(np.float64) - off[0]) / vs[0] - 0.5
    assert np.abs(s[:, 0] - (3.0 * ix + 10.0)).max() <= 0.5 + 1e-3
    # neighbours without colour weight do not count
    Cv[..., :3] = 0.0
    Cv[..., 3] = 1.0
    mask = rng.random(g) < 0.5
    Cv[mask, :3] = 255.0
    Cv[mask, 3] = 0.0
    s = O.sample_colors_f32(vol, Cv, pts)
    assert (s == 0).all()

"""Parity of the colour integrate and colour sampler (recon_tsdf_enable_color,
recon_tsdf_integrate with an rgb image, recon_tsdf_sample_colors) with the
oracle: D, W and the colour volume bit-identical after several frames of
random colour images (one of them row-pitched), vertex colours of the welded
mesh identical, and the switches behave (reset clears, frames without an
image leave colour alone, a volume without colour ignores the image)."""
import numpy as np
import pytest

import oracle_ffi as O
import recon
from helpers import assert_bitexact
from recon import scenes

pytestmark = pytest.mark.gpu


def _frames():
    w = scenes.plain_small(num_frames=4)
    depth = w.render()
    rows, cols = depth.shape[1:]
    rng = np.random.default_rng(21)
    rgbs = [rng.integers(0, 256, (rows, cols, 3)).astype(np.uint8) for _ in range(4)]
    return w, depth, rgbs


def test_color_integrate_and_vertex_colours(cuda):
    w, depth, rgbs = _frames()
    rows, cols = depth.shape[1:]
    vol = recon.TSDFVolume(w.config())
    vol.enable_color()
    print("colour integrate 4 frames", flush=True)
    for f in range(4):
        if f == 1:   # a row-pitched colour image
            buf = cuda.zeros((rows, cols + 7, 3), dtype=cuda.uint8, device="cuda")
            buf[:, :cols] = cuda.from_numpy(rgbs[f]).cuda()
            rgb = buf[:, :cols]
        else:
            rgb = cuda.from_numpy(rgbs[f]).cuda()
        vol.integrate(cuda.from_numpy(depth[f]).cuda(), w.poses[f], rgb=rgb)
    cuda.cuda.synchronize()
    D, W = vol.download()
    Cg = vol.download_colors()
    ov = O.volume_from_config(w.config())
    Do, Wo = O.init(ov)
    Co = np.zeros(Do.shape + (4,), np.float32)
    for f in range(4):
        O.integrate_color_f32(ov, Do, Wo, Co, depth[f], rgbs[f], w.poses[f])
    assert_bitexact(D, Do, "D")
    assert_bitexact(W, Wo, "W")
    assert_bitexact(Cg, Co, "colour")
    assert (Co[..., 3] > 0).sum() > 10000
    # vertex colours of the welded mesh
    v, _, _ = vol.extract_welded_mesh(with_normals=False)
    col = vol.sample_colors(v)
    cuda.cuda.synchronize()
    oc = O.sample_colors_f32(ov, Co, v.cpu().numpy())
    assert np.array_equal(col.cpu().numpy(), oc)
    assert (oc.sum(axis=1) > 0).mean() > 0.95
    # a frame without an image leaves colour alone
    vol.integrate(cuda.from_numpy(depth[0]).cuda(), w.poses[0])
    cuda.cuda.synchronize()
    assert np.array_equal(vol.download_colors(), Cg)
    # reset clears colour; disabling frees it
    vol.reset()
    cuda.cuda.synchronize()
    assert (vol.download_colors() == 0).all()
    vol.enable_color(False)
    with pytest.raises(RuntimeError):
        vol.download_colors()
    with pytest.raises(RuntimeError):
        vol.sample_colors(v)
    vol.close()


def test_image_ignored_without_colour(cuda):
    w, depth, rgbs = _frames()
    a = recon.TSDFVolume(w.config())
    b = recon.TSDFVolume(w.config())
    for f in range(2):
        d = cuda.from_numpy(depth[f]).cuda()
        a.integrate(d, w.poses[f], rgb=cuda.from_numpy(rgbs[f]).cuda())
        b.integrate(d, w.poses[f])
    cuda.cuda.synchronize()
    Da, Wa = a.download()
    Db, Wb = b.download()
    assert_bitexact(Da, Db, "D")
    assert_bitexact(Wa, Wb, "W")
    a.close()
    b.close()

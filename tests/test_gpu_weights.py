"""WEIGHT_BY_VARIANCE on the device (tsdf_volume.cuh:12-15): per-frame,
fused, counting and colour integrates bit-identical to the oracle's; the
paths that weigh every observation 1 (batches, the differentiable integrate,
sparse volumes) refuse the mode."""
import numpy as np
import pytest

import oracle_ffi as O
import recon
from helpers import assert_bitexact
from recon import scenes

pytestmark = pytest.mark.gpu


def _workload():
    w = scenes.plain_small(num_frames=5)
    cfg = w.config()
    cfg.weight_type = 1
    return w, cfg, w.render()


def _oracle(cfg, depth, poses, rgbs=None):
    vol = O.volume_from_config(cfg)
    D, W = O.init(vol)
    Cv = np.zeros(D.shape + (4,), np.float32)
    for f in range(depth.shape[0]):
        if rgbs is None:
            O.integrate_f32(vol, D, W, depth[f], poses[f])
        else:
            O.integrate_color_f32(vol, D, W, Cv, depth[f], rgbs[f], poses[f])
    return D, W, Cv


def test_variance_weights_all_paths(cuda):
    w, cfg, depth = _workload()
    Do, Wo, _ = _oracle(cfg, depth, w.poses)
    dev = cuda.from_numpy(depth).cuda()
    print("variance weights: per-frame", flush=True)
    v = recon.TSDFVolume(cfg)
    for f in range(depth.shape[0]):
        v.integrate(dev[f], w.poses[f])
    cuda.cuda.synchronize()
    D, W = v.download()
    assert_bitexact(D, Do, "per-frame D")
    assert_bitexact(W, Wo, "per-frame W")
    assert len(np.unique(W[W > 0])) > 100        # real weights, not counts
    print("variance weights: fused", flush=True)
    v.reset()
    v.integrate_frames(dev, w.poses)
    cuda.cuda.synchronize()
    D, W = v.download()
    assert_bitexact(D, Do, "fused D")
    assert_bitexact(W, Wo, "fused W")
    print("variance weights: counting", flush=True)
    v.reset()
    ctr = cuda.zeros(1, dtype=cuda.int64, device="cuda")
    v.set_update_counter(ctr)
    for f in range(depth.shape[0]):
        v.integrate(dev[f], w.poses[f])
    cuda.cuda.synchronize()
    D, W = v.download()
    assert_bitexact(D, Do, "counting D")
    assert ctr.item() > 0
    v.set_update_counter(None)
    v.close()
    print("variance weights: colour", flush=True)
    rng = np.random.default_rng(4)
    rgbs = [rng.integers(0, 256, depth.shape[1:] + (3,)).astype(np.uint8) for _ in range(depth.shape[0])]
    Dc, Wc, Cc = _oracle(cfg, depth, w.poses, rgbs)
    v = recon.TSDFVolume(cfg)
    v.enable_color()
    for f in range(depth.shape[0]):
        v.integrate(dev[f], w.poses[f], rgb=cuda.from_numpy(rgbs[f]).cuda())
    cuda.cuda.synchronize()
    D, W = v.download()
    assert_bitexact(D, Dc, "colour D")
    assert_bitexact(W, Wc, "colour W")
    assert_bitexact(v.download_colors(), Cc, "colour C")
    v.close()


def test_unit_weight_paths_refuse(cuda):
    w, cfg, depth = _workload()
    dev = cuda.from_numpy(depth).cuda()
    with pytest.raises(recon.ReconError):
        recon.Batch([(cfg, dev, w.poses)])
    with pytest.raises(recon.ReconError):
        recon.SparseVolume(cfg, 1024)
    v = recon.TSDFVolume(cfg)
    rt = cuda.from_numpy(recon.pose_to_matrix(w.poses)).cuda()
    with pytest.raises(recon.ReconError):
        v.diff_integrate(dev, rt)
    v.close()

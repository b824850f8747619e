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

F64_TOL = 1e-6  # absolute, metres


def _check(w, depth, fused=False, pitch_pad=0, f64=True):
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
        assert err <= F64_TOL, err
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

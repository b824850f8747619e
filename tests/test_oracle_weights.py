"""Pins WEIGHT_BY_VARIANCE in the oracle (tsdf_volume.cuh:12-15 declares the
WEIGHT_TYPE enum and never uses it): observations weigh
sigma(1 m)^2 / sigma(d)^2 with sigma(d) = a + b (d - c)^2.  Closed forms: a
frame whose every depth is 1 m weighs exactly 1, so the result equals the
constant-weight update bit for bit; two frames at depths d1 and d2 give the
sigma-weighted mean of the two single-frame values and W = w1 + w2; nearer
depths weigh more."""
import numpy as np
import pytest

import oracle_ffi as O
import recon
from recon import ffi


def _cfg(weight_type, noise=(0.0, 0.0, 0.0)):
    return recon.make_config((40, 36, 44), voxel_size=0.01, offset=(-0.2, -0.18, 0.5), max_weight=64.0,
                             intrinsics=(120.0, 120.0, 31.5, 23.5), image=(64, 48), truncation=0.03,
                             weight_type=weight_type, noise_model=noise)


def _sigma(d, a=0.0012, b=0.0019, c=0.4):
    return a + b * (d - c) ** 2


def _run(cfg, depths):
    vol = O.volume_from_config(cfg)
    D, W = O.init(vol)
    for d in depths:
        O.integrate_f32(vol, D, W, np.full((48, 64), d, np.float32), [1, 0, 0, 0, 0, 0, 0])
    return D, W


def test_one_metre_equals_constant_weight():
    for frames in ([1.0], [1.0, 1.0, 1.0]):
        D1, W1 = _run(_cfg(1), frames)
        D0, W0 = _run(_cfg(0), frames)
        assert np.array_equal(D1.view(np.uint32), D0.view(np.uint32))
        assert np.array_equal(W1.view(np.uint32), W0.view(np.uint32))
        assert (W0 > 0).sum() > 1000


@pytest.mark.parametrize("d1,d2", [(0.62, 0.71), (0.8, 0.55)])
def test_weighted_mean_of_two_depths(d1, d2):
    D, W = _run(_cfg(1), [d1, d2])
    t1, w_1 = _run(_cfg(0), [d1])
    t2, w_2 = _run(_cfg(0), [d2])
    both = (w_1 > 0) & (w_2 > 0)
    assert both.sum() > 1000
    a1 = (_sigma(1.0) / _sigma(np.float64(np.float32(d1)))) ** 2
    a2 = (_sigma(1.0) / _sigma(np.float64(np.float32(d2)))) ** 2
    expect = (a1 * t1[both].astype(np.float64) + a2 * t2[both]) / (a1 + a2)
    np.testing.assert_allclose(D[both], expect, atol=2e-8, rtol=1e-5)
    np.testing.assert_allclose(W[both], a1 + a2, rtol=1e-6)
    # nearer depths weigh more
    only1 = (w_1 > 0) & (w_2 == 0)
    if only1.any():
        np.testing.assert_allclose(W[only1], a1, rtol=1e-6)
    assert (a1 > a2) == (d1 < d2)


def test_weight_type_validation():
    info = ffi.TsdfInfo()
    lib = ffi.load()
    import ctypes as C
    for wt, noise in [(2, (0, 0, 0)), (-1, (0, 0, 0)), (1, (0.0, 0.001, 0.4)), (1, (0.001, -1.0, 0.4)),
                      (1, (float("nan"), 0.001, 0.4))]:
        assert lib.recon_tsdf_describe(C.byref(_cfg(wt, noise)), C.byref(info)) == ffi.ERR_CONFIG, (wt, noise)
    assert lib.recon_tsdf_describe(C.byref(_cfg(1)), C.byref(info)) == 0
    assert info.weight_type == 1
    np.testing.assert_allclose(list(info.noise_model), [0.0012, 0.0019, 0.4], rtol=1e-6)

"""The Python front refuses device tensors of the wrong element type or
size with ValueError before any kernel runs (a kernel would otherwise read
outside the buffer)."""
import numpy as np
import pytest

import recon
from recon import scenes

pytestmark = pytest.mark.gpu


def test_wrong_types_and_sizes_are_refused(cuda):
    w = scenes.tiny()
    vol = recon.TSDFVolume(w.config())
    cols, rows = w.image
    depth = cuda.zeros((rows, cols), dtype=cuda.float32, device="cuda")
    with pytest.raises(ValueError):
        vol.integrate(depth.double(), w.poses[0])
    with pytest.raises(ValueError):    # colour image one row short
        vol.integrate(depth, w.poses[0], rgb=cuda.zeros((rows - 1, cols, 3), dtype=cuda.uint8,
                                                         device="cuda"))
    with pytest.raises(ValueError):    # colour as float
        vol.integrate(depth, w.poses[0], rgb=cuda.zeros((rows, cols, 3), device="cuda"))
    with pytest.raises(ValueError):    # planar colour (channel stride != 1)
        vol.integrate(depth, w.poses[0],
                      rgb=cuda.zeros((3, rows, cols), dtype=cuda.uint8, device="cuda").permute(1, 2, 0))
    with pytest.raises(ValueError):
        vol.set_update_counter(cuda.zeros(1, dtype=cuda.int32, device="cuda"))
    with pytest.raises(ValueError):
        recon.depth_to_points(depth, w.intrinsics, w.poses[0],
                              rgb=cuda.zeros((rows, cols - 2, 3), dtype=cuda.uint8, device="cuda"))
    small = cuda.zeros(7 * 8 * 8, dtype=cuda.float32, device="cuda")
    for call in (lambda: recon.mc_count(small, (8, 8, 8), (0.1,) * 3, (0.0,) * 3),
                 lambda: recon.mc_extract(small, (8, 8, 8), (0.1,) * 3, (0.0,) * 3),
                 lambda: recon.mc_class_histogram(small, (8, 8, 8)),
                 lambda: recon.mc_extract_welded(small, (8, 8, 8), (0.1,) * 3, (0.0,) * 3)):
        with pytest.raises(ValueError):
            call()
    with pytest.raises(ValueError):
        recon.mc_count(cuda.zeros(512, dtype=cuda.float64, device="cuda"), (8, 8, 8), (0.1,) * 3,
                       (0.0,) * 3)
    # the volume still works after the refusals
    vol.integrate(cuda.from_numpy(w.render()[0]).cuda(), w.poses[0])
    cuda.cuda.synchronize()
    D, W = vol.download()
    assert W.max() == 1.0 and np.isfinite(D).all()
    vol.close()

"""The per-voxel transformation records of TSDFVolume::transformation_voxels_data
(tsdf_volume.cuh:33-38,84-92; filled by init_voxel_transofrmation,
tsdf_volume.cu:106-127): translation = the voxel centre, quaternion (0, 0, 0, 1).
Bar: bit-exact against the oracle's float32 voxel centre on every voxel of a
ragged grid with unequal voxel sizes and a non-zero offset."""
import ctypes as C

import numpy as np
import pytest

import oracle_ffi as O
import recon
from recon import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("grid", [(13, 7, 5), (64, 33, 9)])
def test_records_are_centres_and_identity(cuda, grid):
    w = scenes.tiny(grid=grid, voxel_size=0.6 / max(grid), image=(40, 30), num_frames=1)
    cfg = w.config()
    vol = recon.TSDFVolume(cfg)
    t = vol.transformation_voxels().cpu().numpy()
    ov = O.volume_from_config(cfg)
    gx, gy, gz = vol.info.grid[:]
    assert t.shape == (gx * gy * gz, 7)
    want = np.zeros_like(t)
    want[:, 6] = 1.0
    c = (C.c_float * 3)()
    fn = O.lib().oracle_voxel_centre_f32
    for i in range(gx * gy * gz):
        fn(C.byref(ov), i % gx, (i // gx) % gy, i // (gx * gy), c)
        want[i, :3] = c[:]
    assert np.array_equal(t.view(np.uint32), want.view(np.uint32))
    vol.close()

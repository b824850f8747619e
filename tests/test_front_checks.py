"""The Python front refuses arguments the kernels cannot read safely -- host
tensors, non-tensors, wrong element types, fields smaller than their grid,
colour images smaller than their depth -- with ValueError, before any call
into librecon.so.  Host-only: nothing here touches a GPU (the calls fail
before a device pointer would be formed)."""
import numpy as np
import pytest
import torch

import recon
from recon import ffi

POSE = [1.0, 0, 0, 0, 0, 0, 0]


def _volume_without_handle():
    # the argument checks run before the handle is used
    v = object.__new__(recon.TSDFVolume)
    v._lib = ffi.load()
    v._h = None
    return v


def test_depth_view_refuses_host_and_non_tensors():
    with pytest.raises(ValueError, match="GPU memory"):
        recon.depth_view(torch.zeros(4, 5))
    with pytest.raises(ValueError, match="torch tensor"):
        recon.depth_view(np.zeros((4, 5), np.float32))


def test_volume_methods_refuse_host_tensors():
    v = _volume_without_handle()
    with pytest.raises(ValueError):
        v.integrate(torch.zeros(4, 5), POSE)
    with pytest.raises(ValueError):
        v.integrate_frames(torch.zeros(2, 4, 5), [POSE, POSE])
    with pytest.raises(ValueError):
        v.set_update_counter(torch.zeros(1, dtype=torch.int64))
    with pytest.raises(ValueError):
        v.surface_loss(torch.zeros(3, 3))
    with pytest.raises(ValueError):
        v.sample_colors(torch.zeros(3, 3))
    with pytest.raises(ValueError):
        v.extract_mesh(verts=torch.zeros(2, 3, 3), tris=torch.zeros(2, 3, dtype=torch.int32))
    with pytest.raises(ValueError):
        v.diff_integrate(torch.zeros(2, 4, 5), torch.zeros(2, 12))


def test_field_calls_refuse_host_fields():
    f = torch.zeros(8, 8, 8)
    with pytest.raises(ValueError):
        recon.mc_count(f, (8, 8, 8), (0.1,) * 3, (0.0,) * 3)
    with pytest.raises(ValueError):
        recon.mc_extract(f, (8, 8, 8), (0.1,) * 3, (0.0,) * 3)
    with pytest.raises(ValueError):
        recon.mc_class_histogram(f, (8, 8, 8))
    with pytest.raises(ValueError):
        recon.mc_extract_welded(f, (8, 8, 8), (0.1,) * 3, (0.0,) * 3)


def test_points_calls_refuse_host_tensors():
    with pytest.raises(ValueError):
        recon.depth_to_points(torch.zeros(4, 5), (1.0, 1.0, 2.0, 1.5), POSE)
    with pytest.raises(ValueError):
        recon.render_points(torch.zeros(10, 3), POSE, (1.0, 1.0, 2.0, 1.5), (5, 4))


def test_batches_refuse_host_depth():
    cfg = recon.make_config((8, 8, 8), voxel_size=0.1)
    with pytest.raises(ValueError):
        recon.Batch([(cfg, torch.zeros(1, 4, 5), [POSE])])
    with pytest.raises(ValueError):
        recon.DiffBatch([(cfg, torch.zeros(1, 4, 5), torch.zeros(3, 3), None)])

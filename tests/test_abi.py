"""The C-ABI library without a GPU: it loads, exports every function that
include/*.h declares, and rejects every invalid argument with its own status
before any device call (tsdf_volume.cu:260-268 only logs and returns false)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import oracle_ffi as O
import recon
from recon import ffi

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = []
    for h in sorted((ROOT / "include").rglob("*.h")):
        names += re.findall(r"^RECON_API\s+[\w\s\*]+?\b(recon_\w+)\s*\(", h.read_text(), re.M)
    return names


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 20
    lib = ffi.load()
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(ffi.SIGNATURES), set(names) ^ set(ffi.SIGNATURES)


def test_abi_version():
    hdr = (ROOT / "include" / "recon" / "recon.h").read_text()
    v = int(re.search(r"#define RECON_ABI_VERSION (\d+)", hdr).group(1))
    assert ffi.load().recon_abi_version() == v


def _cfg(**kw):
    args = dict(grid=(64, 48, 32), voxel_size=0.01, offset=(-0.3, -0.2, 0.5))
    args.update(kw)
    return recon.make_config(**args)


def test_describe_matches_oracle_derivation():
    for trunc in (0.0, 0.05):
        cfg = _cfg(truncation=trunc, physical_size=(1.0, 0.7, 0.45))
        info = recon.describe(cfg)
        ov = O.volume_from_config(cfg)
        assert list(info.voxel_size) == list(ov.voxel_size)
        assert info.truncation == ov.truncation
        assert info.max_depth == 24.0
        assert info.num_voxels == 64 * 48 * 32
    # reference default: 1.1 * |voxel_size| (tsdf_volume.cu:199)
    info = recon.describe(_cfg())
    vs = np.float32(np.float32(0.64) / np.float32(64))
    assert abs(info.truncation - 1.1 * np.sqrt(3) * vs) < 1e-7


def _status(cfg):
    info = ffi.TsdfInfo()
    return ffi.load().recon_tsdf_describe(C.byref(cfg), C.byref(info))


@pytest.mark.parametrize("field,value,status", [
    ("voxel_grid_size", (0, 4, 4), ffi.ERR_CONFIG),
    ("voxel_grid_size", (4, -1, 4), ffi.ERR_CONFIG),
    ("physical_size", (1.0, 0.0, 1.0), ffi.ERR_CONFIG),
    ("physical_size", (1.0, float("nan"), 1.0), ffi.ERR_CONFIG),
    ("global_offset", (float("inf"), 0.0, 0.0), ffi.ERR_CONFIG),
    ("max_weight", 0.5, ffi.ERR_CONFIG),
    ("focal_x", 0.0, ffi.ERR_CONFIG),
    ("cy", float("nan"), ffi.ERR_CONFIG),
    ("truncation", float("inf"), ffi.ERR_CONFIG),
    ("voxel_grid_size", (2048, 1024, 1024), ffi.ERR_RANGE),
])
def test_invalid_configs(field, value, status):
    cfg = _cfg()
    if isinstance(value, tuple):
        arr = getattr(cfg, field)
        for i, v in enumerate(value):
            arr[i] = v
    else:
        setattr(cfg, field, value)
    assert _status(cfg) == status
    assert ffi.load().recon_last_error()
    # create validates before touching a device
    h = C.c_void_p()
    assert ffi.load().recon_tsdf_create(C.byref(cfg), C.byref(h)) == status
    assert not h.value


def test_null_arguments():
    lib = ffi.load()
    info = ffi.TsdfInfo()
    assert lib.recon_tsdf_describe(None, C.byref(info)) == ffi.ERR_NULL
    assert lib.recon_tsdf_create(C.byref(_cfg()), None) == ffi.ERR_NULL
    dv = ffi.DepthView(1024, 640 * 4, 480, 640)
    pose = ffi.pose_array([1, 0, 0, 0, 0, 0, 0])
    n = C.c_int64()
    assert lib.recon_tsdf_integrate(None, C.byref(dv), None, 0, pose, None) == ffi.ERR_NULL
    assert lib.recon_tsdf_integrate_frames(None, C.byref(dv), 0, 1, pose, None) == ffi.ERR_NULL
    assert lib.recon_tsdf_reset(None, None) == ffi.ERR_NULL
    assert lib.recon_tsdf_count_triangles(None, C.byref(n), None) == ffi.ERR_NULL
    assert lib.recon_tsdf_extract_mesh(None, None, None, 0, C.byref(n), None) == ffi.ERR_NULL
    assert lib.recon_tsdf_download(None, None, None, None) == ffi.ERR_NULL
    assert lib.recon_tsdf_upload(None, None, None, None) == ffi.ERR_NULL
    assert lib.recon_tsdf_set_update_counter(None, None) == ffi.ERR_NULL
    assert lib.recon_tsdf_get_info(None, C.byref(info)) == ffi.ERR_NULL
    assert lib.recon_tsdf_distances(None) is None and lib.recon_tsdf_weights(None) is None
    lib.recon_tsdf_destroy(None)  # no-op
    g, vs, off = ffi.i32x3((4, 4, 4)), ffi.f32x3((1, 1, 1)), ffi.f32x3((0, 0, 0))
    assert lib.recon_mc_count(None, g, vs, off, C.byref(n), None) == ffi.ERR_NULL
    assert lib.recon_mc_count(1024, g, vs, off, None, None) == ffi.ERR_NULL
    assert lib.recon_mc_extract(None, g, vs, off, None, None, 0, C.byref(n), None) == ffi.ERR_NULL
    assert lib.recon_mc_class_histogram(None, g, None, None) == ffi.ERR_NULL
    xyz = C.c_void_p(4096)
    assert lib.recon_depth_to_points(None, 1, 1, 0, 0, pose, None, 0, xyz, xyz, None, None) == ffi.ERR_NULL
    assert lib.recon_depth_to_points(C.byref(dv), 1, 1, 0, 0, None, None, 0, xyz, xyz, None, None) == ffi.ERR_NULL
    assert lib.recon_depth_to_points(C.byref(dv), 1, 1, 0, 0, pose, None, 0, None, xyz, None, None) == ffi.ERR_NULL


def test_bad_sizes_are_rejected_before_any_launch():
    lib = ffi.load()
    pose = ffi.pose_array([1, 0, 0, 0, 0, 0, 0])
    xyz = C.c_void_p(4096)
    for rows, cols, step in [(0, 640, 2560), (480, 0, 2560), (480, 640, 2556), (480, 640, 2562)]:
        dv = ffi.DepthView(1024, step, rows, cols)
        st = lib.recon_depth_to_points(C.byref(dv), 500.0, 500.0, 320.0, 240.0, pose, None, 0, xyz,
                                       xyz, None, None)
        assert st == ffi.ERR_SIZE, (rows, cols, step)
    dv = ffi.DepthView(1024, 2560, 480, 640)
    for fx in (0.0, -1.0, float("nan")):
        st = lib.recon_depth_to_points(C.byref(dv), fx, 500.0, 320.0, 240.0, pose, None, 0, xyz,
                                       xyz, None, None)
        assert st == ffi.ERR_CONFIG
    bad_pose = ffi.pose_array([1, 0, 0, float("nan"), 0, 0, 0])
    st = lib.recon_depth_to_points(C.byref(dv), 500.0, 500.0, 320.0, 240.0, bad_pose, None, 0, xyz,
                                   xyz, None, None)
    assert st == ffi.ERR_CONFIG
    n = C.c_int64()
    st = lib.recon_mc_count(C.c_void_p(1024), ffi.i32x3((0, 4, 4)), ffi.f32x3((1, 1, 1)),
                            ffi.f32x3((0, 0, 0)), C.byref(n), None)
    assert st == ffi.ERR_SIZE
    st = lib.recon_mc_extract(C.c_void_p(1024), ffi.i32x3((4, 4, 4)), ffi.f32x3((1, 1, 1)),
                              ffi.f32x3((0, 0, 0)), None, None, 10, C.byref(n), None)
    assert st == ffi.ERR_NULL
    st = lib.recon_mc_extract(C.c_void_p(1024), ffi.i32x3((4, 4, 4)), ffi.f32x3((1, 1, 1)),
                              ffi.f32x3((0, 0, 0)), C.c_void_p(4096), None, -1, C.byref(n), None)
    assert st == ffi.ERR_SIZE


def test_differentiable_entry_points_reject_null_and_bad_sizes():
    lib = ffi.load()
    dv = ffi.DepthView(1024, 640 * 4, 480, 640)
    p = C.c_void_p(4096)
    d = C.c_double()
    assert lib.recon_tsdf_diff_integrate(None, C.byref(dv), 0, 1, p, None) == ffi.ERR_NULL
    assert lib.recon_tsdf_surface_loss(None, p, None, 1, None, C.byref(d), None, None,
                                       None) == ffi.ERR_NULL
    assert lib.recon_tsdf_diff_adjoint(None, C.byref(dv), 0, 1, p, p, p, 1, p, None,
                                       None) == ffi.ERR_NULL
    assert lib.recon_pose_to_matrix(None, 1, p) == ffi.ERR_NULL
    assert lib.recon_pose_to_matrix(None, 0, None) == ffi.OK


def test_pose_to_matrix_matches_oracle_narrowing():
    """recon_pose_to_matrix narrows like the plain integrate (the oracle's
    oracle_pose_matrix_f32), bit for bit."""
    import numpy as np
    import oracle_ffi as O
    import recon
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(5)
    q = Rotation.random(50, random_state=6).as_quat()  # x, y, z, w
    poses = np.c_[q[:, 3], q[:, :3], rng.normal(size=(50, 3))]
    rt = recon.pose_to_matrix(poses)
    for k in range(50):
        R, t = O.pose_matrix_f32(poses[k])
        assert (rt[k, :9] == R.ravel()).all() and (rt[k, 9:] == t).all()
    bad = poses[:2].copy()
    bad[1, 2] = np.nan
    with pytest.raises(recon.ReconError):
        recon.pose_to_matrix(bad)


def test_render_arguments_rejected_before_any_launch():
    lib = ffi.load()
    pose = ffi.pose_array([1, 0, 0, 0, 0, 0, 0])
    p = C.c_void_p(4096)

    def call(xyz=p, n=10, pose=pose, fx=500.0, cx=320.0, depth=p, step=2560, rows=480, cols=640,
             color=None, cstep=0, zb=p):
        return lib.recon_points_render(xyz, None, None, n, pose, fx, 500.0, cx, 240.0, depth, step,
                                       rows, cols, color, cstep, zb, None)
    assert call(xyz=None) == ffi.ERR_NULL
    assert call(pose=None) == ffi.ERR_NULL
    assert call(depth=None) == ffi.ERR_NULL
    assert call(zb=None) == ffi.ERR_NULL
    assert call(n=-1) == ffi.ERR_SIZE
    assert call(n=1 << 32) == ffi.ERR_RANGE
    for rows, cols, step in [(0, 640, 2560), (480, 0, 2560), (480, 640, 2556), (480, 640, 2562)]:
        assert call(rows=rows, cols=cols, step=step) == ffi.ERR_SIZE, (rows, cols, step)
    assert call(rows=1 << 18, cols=4, step=16) == ffi.ERR_RANGE
    assert call(color=p, cstep=1919) == ffi.ERR_SIZE
    for fx in (0.0, float("nan"), float("inf"), 1e39):
        assert call(fx=fx) == ffi.ERR_CONFIG, fx
    assert call(cx=float("nan")) == ffi.ERR_CONFIG
    assert call(pose=ffi.pose_array([1, 0, 0, float("inf"), 0, 0, 0])) == ffi.ERR_CONFIG


def test_welded_arguments_rejected_before_any_launch():
    lib = ffi.load()
    nv, nt = C.c_int64(7), C.c_int64(7)
    g, vs, off = ffi.i32x3((4, 4, 4)), ffi.f32x3((1, 1, 1)), ffi.f32x3((0, 0, 0))
    f = C.c_void_p(4096)
    assert lib.recon_mc_extract_welded(f, g, vs, off, None, None, 0, None, 0, None, C.byref(nt), None) == ffi.ERR_NULL
    assert lib.recon_mc_extract_welded(None, g, vs, off, None, None, 0, None, 0, C.byref(nv), C.byref(nt), None) == ffi.ERR_NULL
    assert nv.value == 0 and nt.value == 0
    assert lib.recon_mc_extract_welded(f, ffi.i32x3((0, 4, 4)), vs, off, None, None, 0, None, 0,
                                       C.byref(nv), C.byref(nt), None) == ffi.ERR_SIZE
    assert lib.recon_mc_extract_welded(f, ffi.i32x3((2048, 1024, 1024)), vs, off, None, None, 0, None,
                                       0, C.byref(nv), C.byref(nt), None) == ffi.ERR_RANGE
    assert lib.recon_mc_extract_welded(f, g, vs, off, None, None, -1, None, 0, C.byref(nv),
                                       C.byref(nt), None) == ffi.ERR_SIZE
    assert lib.recon_tsdf_extract_welded_mesh(None, None, None, 0, None, 0, C.byref(nv), C.byref(nt),
                                              None) == ffi.ERR_NULL


def test_colour_arguments_rejected_before_any_launch():
    lib = ffi.load()
    assert lib.recon_tsdf_enable_color(None, 1, None) == ffi.ERR_NULL
    assert lib.recon_tsdf_colors(None) is None
    assert lib.recon_tsdf_download_colors(None, C.c_void_p(4096), None) == ffi.ERR_NULL
    assert lib.recon_tsdf_sample_colors(None, C.c_void_p(4096), 1, C.c_void_p(4096), None) == ffi.ERR_NULL


def test_transformation_voxels_rejects_null_before_any_launch():
    lib = ffi.load()
    assert lib.recon_tsdf_transformation_voxels(None, C.c_void_p(4096), None) == ffi.ERR_NULL


def _gpu_present() -> bool:
    import torch
    return torch.cuda.device_count() > 0


@pytest.mark.skipif(_gpu_present(), reason="host-only check: it passes a fake field address, which "
                                           "must never reach a kernel")
def test_noalloc_mc_refuses_grids_above_its_static_scratch():
    """recon_mc_extract_noalloc checks its 48 MiB static scratch before any
    device call: a 65537 x 160 x 160 grid needs 2 B for each of the 1024
    64-cube chunks of each of its 159^2 cube rows (51.8 MB) and is refused
    with RECON_ERR_RANGE; the message names the limit."""
    lib = ffi.load()
    n = C.c_int64(-1)
    st = lib.recon_mc_extract_noalloc(C.c_void_p(1024), ffi.i32x3((65537, 160, 160)),
                                      ffi.f32x3((1, 1, 1)), ffi.f32x3((0, 0, 0)), None, None, 0,
                                      C.byref(n), None)
    assert st == ffi.ERR_RANGE
    assert b"static maximum" in lib.recon_last_error()
    assert n.value == 0
    assert lib.recon_mc_extract_noalloc(None, ffi.i32x3((4, 4, 4)), ffi.f32x3((1, 1, 1)),
                                        ffi.f32x3((0, 0, 0)), None, None, 0, C.byref(n),
                                        None) == ffi.ERR_NULL

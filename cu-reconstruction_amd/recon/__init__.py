"""Python front of the MI355X TSDF fusion path (librecon.so).

`TSDFVolume` mirrors the reference's TSDFVolume / generate_triangular_surface
API (cu-Reconstruction modules/include/cuda/tsdf_volume.cuh:31-105,
modules/include/cuda/marching_cubes.cuh:13-15) over the C-ABI.  Device
buffers are torch tensors on a ROCm device; torch is plumbing only, every
computation runs in librecon.so's HIP kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import ffi
from .ffi import ReconError, TsdfConfig, TsdfInfo, DepthView, check

__all__ = ["TSDFVolume", "Batch", "make_config", "depth_view", "depth_to_points", "mc_extract", "mc_count",
           "mc_class_histogram", "ReconError", "TsdfConfig", "TsdfInfo", "ffi"]


def _torch():
    import torch  # noqa: F401  (imported before librecon.so so one HIP runtime is shared)
    return torch


def _stream(stream) -> int:
    if stream is None:
        return _torch().cuda.current_stream().cuda_stream
    return int(getattr(stream, "cuda_stream", stream))


def make_config(grid, voxel_size=None, physical_size=None, offset=(0.0, 0.0, 0.0),
                max_weight=128.0, intrinsics=(525.0, 525.0, 319.5, 239.5), image=(640, 480),
                truncation=0.0, max_depth=0.0, weight_type=0, noise_model=(0.0, 0.0, 0.0)) -> TsdfConfig:
    """TSDFVolumeConfig (tsdf_volume.cuh:17-29); give voxel_size or physical_size."""
    cfg = TsdfConfig()
    grid = [int(g) for g in grid]
    if physical_size is None:
        vs = voxel_size if np.ndim(voxel_size) else [voxel_size] * 3
        physical_size = [float(np.float32(v) * np.float32(g)) for v, g in zip(vs, grid)]
    cfg.voxel_grid_size = ffi.i32x3(grid)
    cfg.physical_size = ffi.f32x3(physical_size)
    cfg.global_offset = ffi.f32x3(offset)
    cfg.max_weight = float(max_weight)
    cfg.focal_x, cfg.focal_y, cfg.cx, cfg.cy = [float(v) for v in intrinsics]
    cfg.camera_width, cfg.camera_height = int(image[0]), int(image[1])
    cfg.max_cell_size = 0.0
    cfg.truncation = float(truncation)
    cfg.max_depth = float(max_depth)
    cfg.weight_type = int(weight_type)
    cfg.noise_model = ffi.f32x3(noise_model)
    return cfg


def describe(cfg: TsdfConfig) -> TsdfInfo:
    info = TsdfInfo()
    check(ffi.load().recon_tsdf_describe(C.byref(cfg), C.byref(info)))
    return info


def _device_tensor(t, name, dtype):
    """Raises ValueError unless `t` is a torch tensor of `dtype` in GPU
    memory: the kernels read raw device addresses, so a host tensor or a
    wrong element type must be refused before any call."""
    torch = _torch()
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor")
    if not t.is_cuda:
        raise ValueError(f"{name} must be in GPU memory")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, not {t.dtype}")
    return t


def depth_view(t) -> DepthView:
    """View of a 2-D float32 device tensor (row stride may exceed cols)."""
    torch = _torch()
    _device_tensor(t, "depth", torch.float32)
    if t.dim() != 2 or t.stride(1) != 1 or t.stride(0) < t.shape[1]:
        raise ValueError("depth must be a 2-D float32 tensor with unit column stride")
    return DepthView(t.data_ptr(), t.stride(0) * 4, t.shape[0], t.shape[1])


def _frames_view(depths, name="depths"):
    """View of frame 0 of a (F, H, W) float32 device stack."""
    torch = _torch()
    _device_tensor(depths, name, torch.float32)
    if depths.dim() != 3 or depths.stride(2) != 1 or depths.stride(1) < depths.shape[2] or \
            (depths.shape[0] > 1 and depths.stride(0) < depths.shape[1] * depths.stride(1)):
        raise ValueError(f"{name} must be (F, H, W) with unit column stride and non-overlapping frames")
    return DepthView(depths.data_ptr(), depths.stride(1) * 4, depths.shape[1], depths.shape[2])


def _rgb_arg(rgb, rows, cols):
    """(pointer, row step) of an optional (rows, cols, 3) uint8 device image
    with packed pixels (the row stride may exceed 3*cols)."""
    if rgb is None:
        return None, 0
    torch = _torch()
    _device_tensor(rgb, "rgb", torch.uint8)
    if rgb.dim() != 3 or rgb.shape[2] != 3 or rgb.stride(2) != 1 or rgb.stride(1) != 3 or \
            rgb.stride(0) < 3 * rgb.shape[1]:
        raise ValueError("rgb must be a (rows, cols, 3) uint8 tensor with packed pixels")
    if rgb.shape[0] < rows or rgb.shape[1] < cols:
        raise ValueError(f"rgb ({rgb.shape[0]}x{rgb.shape[1]}) is smaller than the depth "
                         f"({rows}x{cols})")
    return rgb.data_ptr(), rgb.stride(0)


def _field_arg(field, grid):
    """Checks a dense float32[gz][gy][gx] device field covers `grid`."""
    torch = _torch()
    _device_tensor(field, "field", torch.float32)
    g = [int(v) for v in grid]
    if len(g) != 3 or min(g) < 0:
        raise ValueError("grid must be three non-negative sizes")
    if not field.is_contiguous() or field.numel() < g[0] * g[1] * g[2]:
        raise ValueError(f"field must be contiguous with at least {g[0] * g[1] * g[2]} elements")
    return field.data_ptr()


class TSDFVolume:
    """Dense float32 TSDF + weight volume on the current device."""

    def __init__(self, cfg: TsdfConfig):
        _torch()
        self._lib = ffi.load()
        h = C.c_void_p()
        check(self._lib.recon_tsdf_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self.cfg = cfg
        self.info = TsdfInfo()
        check(self._lib.recon_tsdf_get_info(self._h, C.byref(self.info)))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.recon_tsdf_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def grid(self):
        return tuple(self.info.grid)

    @property
    def shape(self):  # numpy order
        g = self.info.grid
        return (g[2], g[1], g[0])

    @property
    def num_voxels(self) -> int:
        return int(self.info.num_voxels)

    @property
    def truncation(self) -> float:
        return float(self.info.truncation)

    def distances_ptr(self) -> int:
        return self._lib.recon_tsdf_distances(self._h)

    def weights_ptr(self) -> int:
        return self._lib.recon_tsdf_weights(self._h)

    def reset(self, stream=None):
        check(self._lib.recon_tsdf_reset(self._h, _stream(stream)))

    def integrate(self, depth, pose, stream=None, rgb=None):
        """TSDFVolume::integrate (tsdf_volume.cu:257-290); async on the stream."""
        dv = depth_view(depth)
        rgb_ptr, rgb_step = _rgb_arg(rgb, dv.rows, dv.cols)
        check(self._lib.recon_tsdf_integrate(self._h, C.byref(dv), rgb_ptr, rgb_step,
                                             ffi.pose_array(pose), _stream(stream)))

    def integrate_frames(self, depths, poses, stream=None):
        """All frames of a (F, H, W) float32 tensor in one sweep per 16 frames."""
        dv = _frames_view(depths)
        poses = np.asarray(poses, dtype=np.float64).reshape(-1, 7)
        if poses.shape[0] != depths.shape[0]:
            raise ValueError("one pose per frame")
        check(self._lib.recon_tsdf_integrate_frames(
            self._h, C.byref(dv), depths.stride(0) * 4, int(depths.shape[0]),
            ffi.pose_array(poses.ravel()), _stream(stream)))

    # ---- differentiable integrate -------------------------------------------
    @staticmethod
    def _frames_view(depths):
        return _frames_view(depths)

    @staticmethod
    def _check_rt(rt, nframes):
        torch = _torch()
        if rt.dtype != torch.float32 or not rt.is_cuda or not rt.is_contiguous() or \
                tuple(rt.shape) != (nframes, 12):
            raise ValueError("poses_rt must be a contiguous (F, 12) float32 device tensor")

    def diff_integrate(self, depths, poses_rt, stream=None):
        """Bilinear-sampling integrate of (F, H, W) depths seen from device
        world->camera matrices poses_rt (F, 12): row-major R, then t."""
        self._check_rt(poses_rt, int(depths.shape[0]))
        dv = self._frames_view(depths)
        check(self._lib.recon_tsdf_diff_integrate(self._h, C.byref(dv), depths.stride(0) * 4,
                                                  int(depths.shape[0]), poses_rt.data_ptr(),
                                                  _stream(stream)))

    def surface_loss(self, points, weights=None, want_lambda=True, stream=None):
        """(loss (1,) float64, residuals (K,), lambda ids (8K,) int64, lambda
        (8K,) float32) for (K, 3) float32 device points."""
        torch = _torch()
        _device_tensor(points, "points", torch.float32)
        if not points.is_contiguous() or points.dim() != 2 or points.shape[1] != 3:
            raise ValueError("points must be a contiguous (K, 3) float32 device tensor")
        K = int(points.shape[0])
        dev = points.device
        loss = torch.empty(1, dtype=torch.float64, device=dev)
        res = torch.empty(K, dtype=torch.float32, device=dev)
        ids = torch.empty(8 * K, dtype=torch.int64, device=dev) if want_lambda else None
        lam = torch.empty(8 * K, dtype=torch.float32, device=dev) if want_lambda else None
        wptr = None
        if weights is not None:
            _device_tensor(weights, "weights", torch.float32)
            if not weights.is_contiguous() or weights.numel() != K:
                raise ValueError("weights must be a contiguous (K,) float32 device tensor")
            wptr = weights.data_ptr()
        check(self._lib.recon_tsdf_surface_loss(
            self._h, points.data_ptr(), wptr, K, res.data_ptr(), loss.data_ptr(),
            ids.data_ptr() if ids is not None else None, lam.data_ptr() if lam is not None else None,
            _stream(stream)))
        return loss, res, ids, lam

    def diff_adjoint(self, depths, poses_rt, lambda_ids, lam, want_depth=True, stream=None):
        """(grad_pose (F, 6) float64, grad_depth like depths or None): the
        adjoint of diff_integrate for dL/dD given as (voxel index, value)
        pairs; xi = (rho, omega), left perturbation."""
        torch = _torch()
        F = int(depths.shape[0])
        self._check_rt(poses_rt, F)
        _device_tensor(lambda_ids, "lambda_ids", torch.int64)
        _device_tensor(lam, "lam", torch.float32)
        if lambda_ids.dtype != torch.int64 or lam.dtype != torch.float32 or \
                lambda_ids.numel() != lam.numel() or not lambda_ids.is_contiguous() or \
                not lam.is_contiguous():
            raise ValueError("lambda_ids (int64) and lam (float32) must be contiguous and equal-sized")
        dv = self._frames_view(depths)
        gp = torch.empty((F, 6), dtype=torch.float64, device=depths.device)
        gd = torch.empty_strided(depths.size(), depths.stride(), dtype=torch.float32,
                                 device=depths.device) if want_depth else None
        check(self._lib.recon_tsdf_diff_adjoint(
            self._h, C.byref(dv), depths.stride(0) * 4, F, poses_rt.data_ptr(),
            lambda_ids.data_ptr(), lam.data_ptr(), int(lam.numel()), gp.data_ptr(),
            gd.data_ptr() if gd is not None else None, _stream(stream)))
        return gp, gd

    def set_update_counter(self, counter_tensor):
        """Later integrate calls add their update counts to element 0 of
        `counter_tensor` (int64 on the GPU); None switches counting off."""
        ptr = None
        if counter_tensor is not None:
            _device_tensor(counter_tensor, "counter", _torch().int64)
            if counter_tensor.numel() < 1:
                raise ValueError("counter must hold at least one element")
            ptr = counter_tensor.data_ptr()
        check(self._lib.recon_tsdf_set_update_counter(self._h, ptr))

    def set_touch_counter(self, counter_tensor):
        """Later integrate calls add to element 0 of `counter_tensor` (int64
        on the GPU) the number of distinct voxels each sweep updates (one
        sweep per call, or per 16 frames of integrate_frames); None switches
        it off."""
        ptr = None
        if counter_tensor is not None:
            _device_tensor(counter_tensor, "counter", _torch().int64)
            if counter_tensor.numel() < 1:
                raise ValueError("counter must hold at least one element")
            ptr = counter_tensor.data_ptr()
        check(self._lib.recon_tsdf_set_touch_counter(self._h, ptr))

    def download(self, stream=None):
        D = np.empty(self.shape, np.float32)
        W = np.empty(self.shape, np.float32)
        check(self._lib.recon_tsdf_download(self._h, D.ctypes.data, W.ctypes.data, _stream(stream)))
        return D, W

    def transformation_voxels(self, stream=None):
        """(N, 7) float32 device tensor: voxel centre and quaternion (x, y, z, w)
        per voxel, the reference's packed TransformationVoxel records."""
        torch = _torch()
        out = torch.empty((self.num_voxels, 7), dtype=torch.float32, device="cuda")
        check(self._lib.recon_tsdf_transformation_voxels(self._h, out.data_ptr(), _stream(stream)))
        return out

    def enable_color(self, enable=True, stream=None):
        """Allocates (or frees) the colour volume (recon_tsdf_enable_color)."""
        check(self._lib.recon_tsdf_enable_color(self._h, 1 if enable else 0, _stream(stream)))

    def download_colors(self, stream=None):
        """(gz, gy, gx, 4) float32: r, g, b, colour weight."""
        out = np.empty(self.shape + (4,), np.float32)
        check(self._lib.recon_tsdf_download_colors(self._h, out.ctypes.data, _stream(stream)))
        return out

    def sample_colors(self, points, stream=None):
        """Colours (N, 3) uint8 at (N, 3) float32 device points."""
        torch = _torch()
        _device_tensor(points, "points", torch.float32)
        if points.numel() % 3:
            raise ValueError("points must hold 3 floats per point")
        pts = points.reshape(-1, 3).contiguous()
        out = torch.empty((pts.shape[0], 3), dtype=torch.uint8, device=pts.device)
        check(self._lib.recon_tsdf_sample_colors(self._h, pts.data_ptr(), pts.shape[0],
                                                 out.data_ptr(), _stream(stream)))
        return out

    def upload(self, D=None, W=None, stream=None):
        Dp = np.ascontiguousarray(D, np.float32) if D is not None else None
        Wp = np.ascontiguousarray(W, np.float32) if W is not None else None
        for a in (Dp, Wp):
            if a is not None and a.size != self.num_voxels:
                raise ValueError("array size does not match the volume")
        check(self._lib.recon_tsdf_upload(self._h, Dp.ctypes.data if Dp is not None else None,
                                          Wp.ctypes.data if Wp is not None else None,
                                          _stream(stream)))

    def count_triangles(self, stream=None) -> int:
        n = C.c_int64()
        check(self._lib.recon_tsdf_count_triangles(self._h, C.byref(n), _stream(stream)))
        return int(n.value)

    def extract_mesh(self, verts=None, tris=None, stream=None):
        """generate_triangular_surface (marching_cubes.cu:191-236) as a
        triangle soup: verts (T, 3, 3) float32, tris (T, 3) int32."""
        torch = _torch()
        if verts is None:
            n = self.count_triangles(stream)
            verts = torch.empty((max(n, 1), 3, 3), dtype=torch.float32, device="cuda")
            tris = torch.empty((max(n, 1), 3), dtype=torch.int32, device="cuda")
        _check_soup_buffers(verts, tris)
        cap = verts.shape[0]
        n = C.c_int64()
        check(self._lib.recon_tsdf_extract_mesh(self._h, verts.data_ptr(),
                                                tris.data_ptr() if tris is not None else None,
                                                cap, C.byref(n), _stream(stream)))
        t = int(n.value)
        return verts[:t], (tris[:t] if tris is not None else None)

    def count_welded(self, stream=None):
        """(vertices, triangles) of the welded mesh."""
        nv, nt = C.c_int64(), C.c_int64()
        check(self._lib.recon_tsdf_extract_welded_mesh(self._h, None, None, 0, None, 0, C.byref(nv),
                                                       C.byref(nt), _stream(stream)))
        return int(nv.value), int(nt.value)

    def extract_welded_mesh(self, with_normals=True, out=None, stream=None):
        """Welded mesh of the zero level set (recon_tsdf_extract_welded_mesh):
        vertices (V, 3), normals (V, 3) or None, triangles (T, 3) int32.
        `out` = (vertices, normals or None, triangles) buffers large enough
        for the mesh make it one call (no counting call first)."""
        return _welded(lambda *a: self._lib.recon_tsdf_extract_welded_mesh(self._h, *a),
                       with_normals, stream, out)


class SparseVolume:
    """Sparse brick volume over the lattice of a dense config
    (recon_sparse_*): unbounded, 8x8x8-voxel bricks allocated along each
    frame's truncation band, the dense update rule voxel for voxel."""

    def __init__(self, cfg, brick_capacity):
        _torch()
        self._lib = ffi.load()
        h = C.c_void_p()
        check(self._lib.recon_sparse_create(C.byref(cfg), int(brick_capacity), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.recon_sparse_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def reset(self, stream=None):
        check(self._lib.recon_sparse_reset(self._h, _stream(stream)))

    def integrate(self, depth, pose, stream=None):
        dv = depth_view(depth)
        check(self._lib.recon_sparse_integrate(self._h, C.byref(dv), ffi.pose_array(pose),
                                               _stream(stream)))

    def set_bounds(self, lo=None, hi=None):
        """Limit allocation to bricks meeting the voxel box [lo, hi) (None: unbounded)."""
        check(self._lib.recon_sparse_set_bounds(self._h, ffi.i32x3(lo) if lo is not None else None,
                                                ffi.i32x3(hi) if hi is not None else None))

    def allocate_box(self, lo, hi, stream=None):
        check(self._lib.recon_sparse_allocate_box(self._h, ffi.i32x3(lo), ffi.i32x3(hi),
                                                  _stream(stream)))

    def stats(self, stream=None):
        """(allocated bricks, overflow flags)."""
        n, f = C.c_int64(), C.c_uint32()
        check(self._lib.recon_sparse_stats(self._h, C.byref(n), C.byref(f), _stream(stream)))
        return int(n.value), int(f.value)

    def export(self, origin, size, stream=None):
        """Dense (sz, sy, sx) D and W device tensors of the voxel box."""
        torch = _torch()
        D = torch.empty(tuple(size[::-1]), dtype=torch.float32, device="cuda")
        W = torch.empty(tuple(size[::-1]), dtype=torch.float32, device="cuda")
        check(self._lib.recon_sparse_export(self._h, ffi.i32x3(origin), ffi.i32x3(size),
                                            D.data_ptr(), W.data_ptr(), _stream(stream)))
        return D, W

    def bricks(self, stream=None):
        """{(bx, by, bz): (D, W)} host arrays (8, 8, 8) in [z][y][x] order."""
        n, _ = self.stats(stream)
        c = np.empty((max(n, 1), 3), np.int32)
        D = np.empty((max(n, 1), 8, 8, 8), np.float32)
        W = np.empty((max(n, 1), 8, 8, 8), np.float32)
        got = C.c_int64()
        check(self._lib.recon_sparse_download_bricks(self._h, c.ctypes.data, D.ctypes.data,
                                                     W.ctypes.data, n, C.byref(got),
                                                     _stream(stream)))
        return {tuple(int(q) for q in c[i]): (D[i], W[i]) for i in range(int(got.value))}


def _check_soup_buffers(verts, tris):
    torch = _torch()
    _device_tensor(verts, "verts", torch.float32)
    if verts.dim() != 3 or tuple(verts.shape[1:]) != (3, 3) or not verts.is_contiguous():
        raise ValueError("verts must be a contiguous (T, 3, 3) float32 device tensor")
    if tris is not None:
        _device_tensor(tris, "tris", torch.int32)
        if tris.dim() != 2 or tris.shape[1] != 3 or not tris.is_contiguous() or \
                tris.shape[0] < verts.shape[0]:
            raise ValueError("tris must be a contiguous (T, 3) int32 device tensor as long as verts")


def _welded(call, with_normals, stream, out=None):
    torch = _torch()
    nv, nt = C.c_int64(), C.c_int64()
    if out is None:
        check(call(None, None, 0, None, 0, C.byref(nv), C.byref(nt), _stream(stream)))
        V, T = int(nv.value), int(nt.value)
        verts = torch.empty((max(V, 1), 3), dtype=torch.float32, device="cuda")
        normals = (torch.empty((max(V, 1), 3), dtype=torch.float32, device="cuda")
                   if with_normals else None)
        tris = torch.empty((max(T, 1), 3), dtype=torch.int32, device="cuda")
    else:
        verts, normals, tris = out
        for name, t, dt in (("vertices", verts, torch.float32), ("normals", normals, torch.float32),
                            ("triangles", tris, torch.int32)):
            if t is None and name != "normals":
                raise ValueError(f"{name} buffer missing")
            if t is not None:
                _device_tensor(t, name, dt)
                if t.dim() != 2 or t.shape[1] != 3 or not t.is_contiguous():
                    raise ValueError(f"{name} must be a contiguous (n, 3) device tensor")
        if normals is not None and normals.shape[0] < verts.shape[0]:
            raise ValueError("normals must be as long as vertices")
    check(call(verts.data_ptr(), normals.data_ptr() if normals is not None else None,
               verts.shape[0], tris.data_ptr(), tris.shape[0], C.byref(nv), C.byref(nt),
               _stream(stream)))
    V, T = int(nv.value), int(nt.value)
    return verts[:V], (normals[:V] if normals is not None else None), tris[:T]


def mc_extract_welded(field, grid, voxel_size, offset, with_normals=True, out=None, stream=None):
    """Welded marching cubes of a dense field (recon_mc_extract_welded)."""
    fp = _field_arg(field, grid)
    g, vs, off = _grid_args(grid, voxel_size, offset)
    return _welded(lambda *a: ffi.load().recon_mc_extract_welded(fp, g, vs, off, *a),
                   with_normals, stream, out)


def _grid_args(grid, voxel_size, offset):
    return ffi.i32x3(grid), ffi.f32x3(voxel_size), ffi.f32x3(offset)


def mc_count(field, grid, voxel_size, offset, stream=None) -> int:
    fp = _field_arg(field, grid)
    g, vs, off = _grid_args(grid, voxel_size, offset)
    n = C.c_int64()
    check(ffi.load().recon_mc_count(fp, g, vs, off, C.byref(n), _stream(stream)))
    return int(n.value)


def mc_extract(field, grid, voxel_size, offset, capacity=None, stream=None, with_tris=True):
    torch = _torch()
    fp = _field_arg(field, grid)
    if capacity is None:
        capacity = mc_count(field, grid, voxel_size, offset, stream)
    verts = torch.empty((max(capacity, 1), 3, 3), dtype=torch.float32, device=field.device)
    tris = torch.empty((max(capacity, 1), 3), dtype=torch.int32, device=field.device) if with_tris else None
    g, vs, off = _grid_args(grid, voxel_size, offset)
    n = C.c_int64()
    st = ffi.load().recon_mc_extract(fp, g, vs, off, verts.data_ptr(),
                                     tris.data_ptr() if tris is not None else None, capacity,
                                     C.byref(n), _stream(stream))
    check(st)
    t = int(n.value)
    return verts[:t], (tris[:t] if tris is not None else None)


def mc_class_histogram(field, grid, stream=None):
    torch = _torch()
    fp = _field_arg(field, grid)
    hist = torch.zeros(256, dtype=torch.int64, device=field.device)
    check(ffi.load().recon_mc_class_histogram(fp, ffi.i32x3(grid), hist.data_ptr(),
                                              _stream(stream)))
    return hist


def pose_to_matrix(poses) -> np.ndarray:
    """(F, 7) {qw,qx,qy,qz,tx,ty,tz} poses -> (F, 12) float32 [R row-major | t],
    narrowed as the plain integrate narrows them."""
    poses = np.ascontiguousarray(np.asarray(poses, np.float64).reshape(-1, 7))
    out = np.empty((poses.shape[0], 12), np.float32)
    check(ffi.load().recon_pose_to_matrix(poses.ctypes.data, poses.shape[0], out.ctypes.data))
    return out


def depth_to_points(depth, intrinsics, transform, rgb=None, with_colors=False, stream=None):
    """cudaPointCloud::extract_points (cuda_point_cloud.cu:179-226) for one
    frame: returns xyz (H*W, 3) float32, valid (H*W,) uint8 and optionally
    colours (H*W, 3) uint8."""
    torch = _torch()
    dv = depth_view(depth)
    rgb_ptr, rgb_step = _rgb_arg(rgb, dv.rows, dv.cols)
    n = depth.shape[0] * depth.shape[1]
    xyz = torch.empty((n, 3), dtype=torch.float32, device=depth.device)
    valid = torch.empty((n,), dtype=torch.uint8, device=depth.device)
    colors = torch.empty((n, 3), dtype=torch.uint8, device=depth.device) if with_colors else None
    fx, fy, cx, cy = [float(v) for v in intrinsics]
    check(ffi.load().recon_depth_to_points(
        C.byref(dv), fx, fy, cx, cy, ffi.pose_array(transform),
        rgb_ptr, rgb_step,
        xyz.data_ptr(), valid.data_ptr(), colors.data_ptr() if colors is not None else None,
        _stream(stream)))
    return (xyz, valid, colors) if with_colors else (xyz, valid)


def render_points(xyz, pose, intrinsics, image, valid=None, colors=None, with_colors=False,
                  depth_out=None, color_out=None, zbuffer=None, stream=None):
    """cudaPointCloud::project_to_depth / project_to_color_depth
    (cuda_point_cloud.cu:229-353): z-buffer render of world points xyz (N, 3)
    float32 at the world->camera `pose` into a (rows, cols) depth map and,
    with with_colors, a (rows, cols, 3) uint8 image.  `image` is (cols, rows).
    depth_out / color_out may be row-pitched views (unit inner stride)."""
    torch = _torch()
    cols, rows = int(image[0]), int(image[1])
    dev = xyz.device
    _device_tensor(xyz, "xyz", torch.float32)
    if xyz.dim() != 2 or xyz.shape[1] != 3 or not xyz.is_contiguous():
        raise ValueError("xyz must be a contiguous (N, 3) float32 tensor")
    n = xyz.shape[0]
    for name, t, shape in (("valid", valid, (n,)), ("colors", colors, (n, 3))):
        if t is not None:
            _device_tensor(t, name, torch.uint8)
            if tuple(t.shape) != shape or not t.is_contiguous():
                raise ValueError(f"{name} must be a contiguous uint8 tensor of shape {shape}")
    for name, t, dt in (("depth_out", depth_out, torch.float32), ("color_out", color_out, torch.uint8),
                        ("zbuffer", zbuffer, torch.int64)):
        if t is not None:
            _device_tensor(t, name, dt)
    if depth_out is None:
        depth_out = torch.empty((rows, cols), dtype=torch.float32, device=dev)
    if with_colors and color_out is None:
        color_out = torch.empty((rows, cols, 3), dtype=torch.uint8, device=dev)
    if depth_out.shape != (rows, cols) or depth_out.stride(1) != 1 or depth_out.dtype != torch.float32:
        raise ValueError("depth_out must be (rows, cols) float32 with unit column stride")
    if color_out is not None and (color_out.shape != (rows, cols, 3) or color_out.stride(2) != 1
                                  or color_out.stride(1) != 3 or color_out.dtype != torch.uint8):
        raise ValueError("color_out must be (rows, cols, 3) uint8 with packed pixels")
    if zbuffer is None:
        zbuffer = torch.empty((rows * cols,), dtype=torch.int64, device=dev)
    elif zbuffer.dtype != torch.int64 or zbuffer.numel() < rows * cols:
        raise ValueError("zbuffer must be an int64 tensor of at least rows*cols elements")
    fx, fy, cx, cy = [float(v) for v in intrinsics]
    check(ffi.load().recon_points_render(
        xyz.data_ptr() if n else None, valid.data_ptr() if valid is not None else None,
        colors.data_ptr() if colors is not None else None, n, ffi.pose_array(pose), fx, fy, cx, cy,
        depth_out.data_ptr(), depth_out.stride(0) * 4, rows, cols,
        color_out.data_ptr() if color_out is not None else None,
        color_out.stride(0) if color_out is not None else 0, zbuffer.data_ptr(), _stream(stream)))
    return (depth_out, color_out) if with_colors or color_out is not None else depth_out


class Batch:
    """Many volumes integrated and meshed by one graph launch
    (recon_batch_*).  `volumes` is a list of (config, depth tensor (F, H, W)
    on the device, poses (F, 7)); the depth tensors are kept alive here
    because the graph reads them on every run."""

    def __init__(self, volumes, mesh_capacity=0, tune_cache=None):
        _torch()
        self._lib = ffi.load()
        self._keep = []
        arr = (ffi.BatchVolume * len(volumes))()
        for i, (cfg, depth, poses) in enumerate(volumes):
            poses = np.ascontiguousarray(np.asarray(poses, np.float64).reshape(-1, 7))
            self._keep.append((depth, poses))
            v = arr[i]
            v.config = cfg
            v.num_frames = int(poses.shape[0])
            if v.num_frames:
                _frames_view(depth, "depth")
                if depth.shape[0] != v.num_frames:
                    raise ValueError("depth must be (F, H, W) with one frame per pose")
                v.rows, v.cols = int(depth.shape[1]), int(depth.shape[2])
                v.depth_step = depth.stride(1) * 4
                v.frame_stride = depth.stride(0) * 4
                v.depth = depth.data_ptr()
                v.poses = poses.ctypes.data
        self._arr = arr
        self.n = len(volumes)
        h = C.c_void_p()
        check(self._lib.recon_batch_create(arr, self.n, int(mesh_capacity),
                                           tune_cache.encode() if tune_cache else None, C.byref(h)))
        self._h = h

    @staticmethod
    def autotune(volumes, cache_path, iterations=3, stream=None):
        b = (ffi.BatchVolume * len(volumes))()
        keep = []
        for i, (cfg, depth, poses) in enumerate(volumes):
            poses = np.ascontiguousarray(np.asarray(poses, np.float64).reshape(-1, 7))
            if poses.shape[0]:
                _frames_view(depth, "depth")
                if depth.shape[0] != poses.shape[0]:
                    raise ValueError("depth must be (F, H, W) with one frame per pose")
            keep.append(poses)
            b[i].config = cfg
            b[i].num_frames = int(poses.shape[0])
            b[i].rows, b[i].cols = int(depth.shape[1]), int(depth.shape[2])
            b[i].depth_step = depth.stride(1) * 4
            b[i].frame_stride = depth.stride(0) * 4
            b[i].depth = depth.data_ptr()
            b[i].poses = poses.ctypes.data
        check(ffi.load().recon_batch_autotune(b, len(volumes), str(cache_path).encode(), iterations,
                                              _stream(stream)))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.recon_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def run(self, stream=None):
        check(self._lib.recon_batch_run(self._h, _stream(stream)))

    def info(self, i) -> TsdfInfo:
        info = TsdfInfo()
        check(self._lib.recon_batch_volume_data(self._h, i, None, None, C.byref(info)))
        return info

    def download(self, i, stream=None):
        info = self.info(i)
        shape = (info.grid[2], info.grid[1], info.grid[0])
        D = np.empty(shape, np.float32)
        W = np.empty(shape, np.float32)
        check(self._lib.recon_batch_download(self._h, i, D.ctypes.data, W.ctypes.data,
                                             _stream(stream)))
        return D, W

    def mesh_ranges(self, stream=None):
        r = np.zeros(2 * self.n, np.int64)
        p = C.c_void_p()
        check(self._lib.recon_batch_meshes(self._h, r.ctypes.data_as(C.POINTER(C.c_int64)),
                                           C.byref(p), _stream(stream)))
        return r.reshape(-1, 2)

    def download_mesh(self, i, stream=None):
        n = int(self.mesh_ranges(stream)[i, 1])
        out = np.empty((max(n, 1), 3, 3), np.float32)
        got = C.c_int64()
        check(self._lib.recon_batch_download_mesh(self._h, i, out.ctypes.data, n, C.byref(got),
                                                  _stream(stream)))
        return out[:got.value]


class DiffBatch:
    """Pose refinement over many fragments (recon_diff_batch_*).  `volumes` is
    a list of (config, depth tensor (F, H, W) on the device, points (K, 3)
    float32 device tensor, weights (K,) float32 device tensor or None), each
    optionally followed by model-to-frame pairs ((P, 2) int32 device tensor
    of (frame, point), grouped by frame), their weight in the loss, the
    residual's truncation (metres; 0: none) and the points' unit normals
    ((K, 3) float32 device tensor or None, read by refine_gn); the tensors are
    kept alive here because every step reads them."""

    def __init__(self, volumes):
        torch = _torch()
        self._lib = ffi.load()
        self._keep = []
        arr = (ffi.DiffVolume * len(volumes))()
        frames = 0
        for i, spec in enumerate(volumes):
            cfg, depth, points, weights = spec[:4]
            pairs, pair_weight = (spec[4], float(spec[5])) if len(spec) > 4 else (None, 0.0)
            pair_trunc = float(spec[6]) if len(spec) > 6 else 0.0
            normals = spec[7] if len(spec) > 7 else None
            if pairs is not None:
                _device_tensor(pairs, "frame pairs", torch.int32)
                if pairs.dim() != 2 or pairs.shape[1] != 2 or not pairs.is_contiguous():
                    raise ValueError("frame pairs must be a contiguous (P, 2) int32 tensor")
                self._keep.append(pairs)
                arr[i].frame_pairs = pairs.data_ptr()
                arr[i].num_frame_pairs = int(pairs.shape[0])
                arr[i].frame_weight = pair_weight
                arr[i].frame_trunc = pair_trunc
            _frames_view(depth, "depth")
            _device_tensor(points, "points", torch.float32)
            if not points.is_contiguous() or points.dim() != 2 or points.shape[1] != 3:
                raise ValueError("points must be a contiguous (K, 3) float32 tensor")
            if weights is not None:
                _device_tensor(weights, "weights", torch.float32)
                if not weights.is_contiguous() or weights.numel() != points.shape[0]:
                    raise ValueError("weights must be a contiguous (K,) float32 tensor")
            if normals is not None:
                _device_tensor(normals, "normals", torch.float32)
                if not normals.is_contiguous() or tuple(normals.shape) != tuple(points.shape):
                    raise ValueError("normals must be a contiguous (K, 3) float32 tensor")
                arr[i].normals = normals.data_ptr()
            self._keep.append((depth, points, weights, normals))
            v = arr[i]
            v.config = cfg
            v.num_frames = int(depth.shape[0])
            v.rows, v.cols = int(depth.shape[1]), int(depth.shape[2])
            v.depth_step = depth.stride(1) * 4
            v.frame_stride = depth.stride(0) * 4
            v.depth = depth.data_ptr()
            v.points = points.data_ptr()
            v.weights = weights.data_ptr() if weights is not None else None
            v.num_points = int(points.shape[0])
            frames += v.num_frames
        self.n = len(volumes)
        self.total_frames = frames
        h = C.c_void_p()
        check(self._lib.recon_diff_batch_create(arr, self.n, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.recon_diff_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check_rt(self, rt):
        torch = _torch()
        _device_tensor(rt, "poses_rt", torch.float32)
        if rt.dtype != torch.float32 or not rt.is_contiguous() or \
                tuple(rt.shape) != (self.total_frames, 12):
            raise ValueError("poses_rt must be a contiguous (total frames, 12) float32 tensor")

    def forward(self, poses_rt, stream=None):
        """(n,) float64 device tensor of per-volume losses."""
        torch = _torch()
        self._check_rt(poses_rt)
        loss = torch.empty(self.n, dtype=torch.float64, device=poses_rt.device)
        check(self._lib.recon_diff_batch_forward(self._h, poses_rt.data_ptr(), loss.data_ptr(),
                                                 _stream(stream)))
        return loss

    def backward(self, stream=None):
        """(total frames, 6) float64 gradients for the last forward."""
        torch = _torch()
        g = torch.empty((self.total_frames, 6), dtype=torch.float64, device="cuda")
        check(self._lib.recon_diff_batch_backward(self._h, g.data_ptr(), _stream(stream)))
        return g

    def refine(self, poses_rt, steps, lr_translation=1e-3, lr_rotation=2e-3, beta1=0.9,
               beta2=0.999, eps=1e-12, decay=0.95, stream=None):
        """Refines poses_rt in place; returns the (steps, n) loss history."""
        torch = _torch()
        self._check_rt(poses_rt)
        hist = torch.empty((max(steps, 1), self.n), dtype=torch.float64, device=poses_rt.device)
        opt = ffi.AdamParams(lr_translation, lr_rotation, beta1, beta2, eps, decay)
        check(self._lib.recon_diff_batch_refine(self._h, poses_rt.data_ptr(), int(steps),
                                                C.byref(opt), hist.data_ptr(), _stream(stream)))
        return hist[:steps]

    def refine_gn(self, poses_rt, steps, step=0.7, damping=1e-3, trunc_ratio=1.0, stream=None):
        """Damped Gauss-Newton refinement of poses_rt in place (every volume
        needs frame pairs); the model-to-frame truncation shrinks by
        trunc_ratio over the first half of the steps.  Returns the (steps, n)
        loss history."""
        torch = _torch()
        self._check_rt(poses_rt)
        hist = torch.empty((max(steps, 1), self.n), dtype=torch.float64, device=poses_rt.device)
        opt = ffi.GnParams(step, damping, trunc_ratio)
        check(self._lib.recon_diff_batch_refine_gn(self._h, poses_rt.data_ptr(), int(steps),
                                                   C.byref(opt), hist.data_ptr(), _stream(stream)))
        return hist[:steps]

    def volume(self, i):
        """(D, W) device pointers and the info of volume i."""
        d, w = C.c_void_p(), C.c_void_p()
        info = TsdfInfo()
        check(self._lib.recon_diff_batch_volume_data(self._h, i, C.byref(d), C.byref(w),
                                                     C.byref(info)))
        return d.value, w.value, info

    def download(self, i, stream=None):
        """(D, W) of volume i as numpy arrays in (z, y, x) order."""
        _, _, info = self.volume(i)
        shape = (int(info.grid[2]), int(info.grid[1]), int(info.grid[0]))
        D = np.empty(shape, np.float32)
        W = np.empty(shape, np.float32)
        check(self._lib.recon_diff_batch_download(self._h, i, D.ctypes.data, W.ctypes.data,
                                                  _stream(stream)))
        return D, W

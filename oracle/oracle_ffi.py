"""ctypes binding of oracle/lib/liboracle.so (the CPU restatement).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline,
never as the product path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "lib" / "liboracle.so"


class OracleVolume(C.Structure):
    _fields_ = [("grid", C.c_int32 * 3), ("voxel_size", C.c_float * 3), ("offset", C.c_float * 3),
                ("truncation", C.c_float), ("max_weight", C.c_float), ("max_depth", C.c_float),
                ("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float),
                ("weight_type", C.c_int32), ("noise", C.c_float * 3)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: run make -C oracle")
        L = C.CDLL(str(LIB))
        vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        sig = {
            "oracle_pose_matrix_f32": (None, [vp, vp, vp]),
            "oracle_voxel_centre_f32": (None, [vp, i32, i32, i32, vp]),
            "oracle_tsdf_init": (None, [vp, vp, vp]),
            "oracle_tsdf_integrate_f32": (i64, [vp, vp, vp, vp, i32, i32, i64, vp, vp, C.c_int]),
            "oracle_tsdf_integrate_f64": (i64, [vp, vp, vp, vp, i32, i32, i64, vp, vp, C.c_int]),
            "oracle_mc_count": (i64, [vp, vp, vp, C.c_int]),
            "oracle_mc_extract": (i64, [vp, vp, vp, vp, vp, vp, vp, vp, i64, C.c_int]),
            "oracle_depth_to_points": (None, [vp, i32, i32, i64, C.c_double, C.c_double, C.c_double,
                                              C.c_double, vp, vp, i64, vp, vp, vp]),
            "oracle_point_cloud_from_depth_f64": (i64, [vp, i32, i32, i64, C.c_double, C.c_double,
                                                        C.c_double, C.c_double, vp, C.c_double,
                                                        C.c_double, vp, vp, i64]),
        }
        sig.update({
            "oracle_diff_forward_f64": (None, [vp, vp, vp, vp, i32, i32, vp, i32, C.c_int]),
            "oracle_trilinear_support": (C.c_int, [vp, vp, vp, vp]),
            "oracle_diff_forward_f32": (None, [vp, vp, vp, vp, i32, i32, vp, i32, C.c_int]),
            "oracle_surface_loss_f64": (C.c_double, [vp, vp, vp, vp, i64, vp]),
            "oracle_diff_adjoint_f64": (None, [vp, vp, i32, i32, vp, i32, vp, vp, vp]),
            "oracle_frame_loss_f64": (C.c_double, [vp, vp, i32, i32, vp, i32, vp, vp, vp, i64,
                                                   C.c_double, vp, vp]),
            "oracle_frame_normal_matrix_f64": (None, [vp, vp, i32, i32, vp, i32, vp, vp, vp, vp,
                                                      i64, C.c_double, vp, vp]),
            "oracle_stable_points": (None, [vp, vp, i32, i32, vp, i32, vp, i64, C.c_double,
                                            C.c_double, vp]),
            "oracle_render_points": (None, [vp, vp, vp, i64, vp, C.c_float, C.c_float, C.c_float,
                                            C.c_float, i32, i32, i64, vp, vp, i64]),
            "oracle_mc_weld": (i64, [vp, vp, vp, vp, vp, vp, i64, vp, i64, vp]),
            "oracle_tsdf_integrate_color_f32": (i64, [vp, vp, vp, vp, vp, i32, i32, i64, vp, i64,
                                                      vp, C.c_int]),
            "oracle_sample_colors_f32": (None, [vp, vp, vp, i64, vp]),
            "oracle_sparse_create": (vp, [vp]),
            "oracle_sparse_destroy": (None, [vp]),
            "oracle_sparse_allocate_box": (None, [vp, vp, vp]),
            "oracle_sparse_set_bounds": (None, [vp, vp, vp]),
            "oracle_sparse_integrate": (i64, [vp, vp, i32, i32, i64, vp]),
            "oracle_sparse_num_bricks": (i64, [vp]),
            "oracle_sparse_export": (None, [vp, vp, vp, vp, vp]),
            "oracle_sparse_bricks": (i64, [vp, vp, vp, vp, i64]),
        })
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray | None):
    # a ctypes pointer that keeps `a` alive until the call has read it (a bare
    # address would let a temporary array be freed, and its memory reused,
    # before the C function runs)
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def volume(grid, voxel_size, offset, truncation, max_weight, intrinsics, max_depth=24.0):
    """Oracle volume; voxel size in float32 as physical/grid is computed
    (tsdf_volume.cu:197) when `voxel_size` is given as the physical size
    divided out already."""
    v = OracleVolume()
    v.grid = (C.c_int32 * 3)(*[int(g) for g in grid])
    v.voxel_size = (C.c_float * 3)(*[float(np.float32(s)) for s in voxel_size])
    v.offset = (C.c_float * 3)(*[float(np.float32(o)) for o in offset])
    v.truncation = float(np.float32(truncation))
    v.max_weight = float(np.float32(max_weight))
    v.max_depth = float(np.float32(max_depth))
    v.fx, v.fy, v.cx, v.cy = [float(np.float32(k)) for k in intrinsics]
    return v


def volume_from_config(cfg) -> OracleVolume:
    """From a recon_tsdf_config: voxel_size = physical / grid in float32,
    truncation default 1.1 * |voxel_size| (tsdf_volume.cu:197-199)."""
    grid = [int(g) for g in cfg.voxel_grid_size]
    phys = np.array(list(cfg.physical_size), np.float32)
    vs = phys / np.array(grid, np.float32)
    if cfg.truncation > 0:
        trunc = np.float32(cfg.truncation)
    else:
        dot = (vs[0] * vs[0] + vs[1] * vs[1]) + vs[2] * vs[2]
        trunc = np.float32(1.1) * np.sqrt(np.float32(dot))
    maxd = np.float32(cfg.max_depth) if cfg.max_depth > 0 else np.float32(24.0)
    v = volume(grid, vs, list(cfg.global_offset), trunc, cfg.max_weight,
               (cfg.focal_x, cfg.focal_y, cfg.cx, cfg.cy), maxd)
    v.weight_type = int(getattr(cfg, "weight_type", 0))
    nm = list(getattr(cfg, "noise_model", (0.0, 0.0, 0.0)))
    if not any(nm):
        nm = [0.0012, 0.0019, 0.4]       # the library's default noise model
    v.noise = (C.c_float * 3)(*[float(np.float32(q)) for q in nm])
    return v


def shape(vol: OracleVolume):
    g = vol.grid
    return (g[2], g[1], g[0])


def init(vol: OracleVolume):
    D = np.empty(shape(vol), np.float32)
    W = np.empty(shape(vol), np.float32)
    lib().oracle_tsdf_init(C.byref(vol), _p(D), _p(W))
    return D, W


def pose_arr(pose):
    return np.ascontiguousarray(np.asarray(pose, np.float64).reshape(7))


def integrate_f32(vol, D, W, depth, pose, mask=False, nthreads=0):
    depth = np.ascontiguousarray(depth, np.float32)
    rows, cols = depth.shape
    m = np.empty(shape(vol), np.uint8) if mask else None
    n = lib().oracle_tsdf_integrate_f32(C.byref(vol), _p(D), _p(W), _p(depth), rows, cols, cols,
                                        _p(pose_arr(pose)), _p(m), nthreads)
    return (n, m) if mask else n


def integrate_f64(vol, D64, W64, depth, pose, nthreads=0):
    depth = np.ascontiguousarray(depth, np.float32)
    rows, cols = depth.shape
    flips = C.c_int64()
    n = lib().oracle_tsdf_integrate_f64(C.byref(vol), _p(D64), _p(W64), _p(depth), rows, cols, cols,
                                        _p(pose_arr(pose)), C.addressof(flips), nthreads)
    return n, flips.value


def pose_matrix_f32(pose):
    R = np.empty(9, np.float32)
    t = np.empty(3, np.float32)
    lib().oracle_pose_matrix_f32(_p(pose_arr(pose)), _p(R), _p(t))
    return R.reshape(3, 3), t


def mc_count(field, grid, hist=False, nthreads=0):
    field = np.ascontiguousarray(field, np.float32)
    g = np.array(grid, np.int32)
    h = np.zeros(256, np.int64) if hist else None
    n = lib().oracle_mc_count(_p(field), _p(g), _p(h), nthreads)
    return (n, h) if hist else n


def mc_extract(field, grid, voxel_size, offset, nthreads=0):
    """Returns verts (T,3,3) float32, tris (T,3) int32, edge ids (T,3) int32,
    cube ids (T,) int64."""
    field = np.ascontiguousarray(field, np.float32)
    g = np.array(grid, np.int32)
    vs = np.array(voxel_size, np.float32)
    off = np.array(offset, np.float32)
    n = mc_count(field, grid, nthreads=nthreads)
    verts = np.empty((max(n, 1), 3, 3), np.float32)
    tris = np.empty((max(n, 1), 3), np.int32)
    edges = np.empty((max(n, 1), 3), np.int32)
    cubes = np.empty((max(n, 1),), np.int64)
    t = lib().oracle_mc_extract(_p(field), _p(g), _p(vs), _p(off), _p(verts), _p(tris), _p(edges),
                                _p(cubes), n, nthreads)
    assert t == n
    return verts[:n], tris[:n], edges[:n], cubes[:n]


def depth_to_points(depth, intrinsics, transform, rgb=None):
    depth = np.ascontiguousarray(depth, np.float32)
    rows, cols = depth.shape
    xyz = np.empty((rows * cols, 3), np.float32)
    valid = np.empty(rows * cols, np.uint8)
    colors = np.empty((rows * cols, 3), np.uint8)
    fx, fy, cx, cy = [float(v) for v in intrinsics]
    rgbp = np.ascontiguousarray(rgb, np.uint8) if rgb is not None else None
    lib().oracle_depth_to_points(_p(depth), rows, cols, cols, fx, fy, cx, cy,
                                 _p(pose_arr(transform)), _p(rgbp), cols * 3, _p(xyz), _p(valid),
                                 _p(colors))
    return xyz, valid, colors


def integrate_color_f32(vol, D, W, Cv, depth, rgb, pose, nthreads=0):
    """Colour integrate: D, W (gz, gy, gx) float32 and Cv (gz, gy, gx, 4)
    float32 updated in place; rgb (rows, cols, 3) uint8."""
    depth = np.ascontiguousarray(depth, np.float32)
    rgb = np.ascontiguousarray(rgb, np.uint8)
    rows, cols = depth.shape
    return lib().oracle_tsdf_integrate_color_f32(C.byref(vol), _p(D), _p(W), _p(Cv), _p(depth),
                                                 rows, cols, cols, _p(rgb), 3 * cols,
                                                 _p(pose_arr(pose)), nthreads)


def sample_colors_f32(vol, Cv, points):
    points = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    out = np.empty((points.shape[0], 3), np.uint8)
    lib().oracle_sample_colors_f32(C.byref(vol), _p(np.ascontiguousarray(Cv, np.float32)),
                                   _p(points), points.shape[0], _p(out))
    return out


class Sparse:
    """The sparse (brick) volume oracle over `vol`'s lattice."""

    def __init__(self, vol):
        self.vol = vol
        self._h = lib().oracle_sparse_create(C.byref(vol))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_sparse_destroy(self._h)
            self._h = None

    def allocate_box(self, lo, hi):
        lib().oracle_sparse_allocate_box(self._h, _p(np.array(lo, np.int32)),
                                         _p(np.array(hi, np.int32)))

    def set_bounds(self, lo, hi):
        lib().oracle_sparse_set_bounds(self._h, _p(np.array(lo, np.int32)),
                                       _p(np.array(hi, np.int32)))

    def integrate(self, depth, pose):
        depth = np.ascontiguousarray(depth, np.float32)
        rows, cols = depth.shape
        return lib().oracle_sparse_integrate(self._h, _p(depth), rows, cols, cols,
                                             _p(pose_arr(pose)))

    def num_bricks(self):
        return lib().oracle_sparse_num_bricks(self._h)

    def export(self, origin, size):
        """Dense (sz, sy, sx) D and W of the box [origin, origin + size)."""
        D = np.empty(tuple(size[::-1]), np.float32)
        W = np.empty(tuple(size[::-1]), np.float32)
        lib().oracle_sparse_export(self._h, _p(np.array(origin, np.int32)),
                                   _p(np.array(size, np.int32)), _p(D), _p(W))
        return D, W

    def bricks(self):
        """{(bx, by, bz): (D, W)} with D, W (8, 8, 8) in [z][y][x] order."""
        n = self.num_bricks()
        c = np.empty((max(n, 1), 3), np.int32)
        D = np.empty((max(n, 1), 8, 8, 8), np.float32)
        W = np.empty((max(n, 1), 8, 8, 8), np.float32)
        lib().oracle_sparse_bricks(self._h, _p(c), _p(D), _p(W), n)
        return {tuple(int(q) for q in c[i]): (D[i], W[i]) for i in range(n)}


def mc_weld(field, grid, voxel_size, offset):
    """Welded mesh: vertices (V, 3), normals (V, 3) float32, triangles (T, 3)
    int32."""
    field = np.ascontiguousarray(field, np.float32)
    g = np.array(grid, np.int32)
    vs = np.array(voxel_size, np.float32)
    off = np.array(offset, np.float32)
    nv = C.c_int64()
    nt = lib().oracle_mc_weld(_p(field), _p(g), _p(vs), _p(off), None, None, 0, None, 0, C.byref(nv))
    verts = np.empty((max(nv.value, 1), 3), np.float32)
    normals = np.empty((max(nv.value, 1), 3), np.float32)
    tris = np.empty((max(nt, 1), 3), np.int32)
    lib().oracle_mc_weld(_p(field), _p(g), _p(vs), _p(off), _p(verts), _p(normals), nv.value,
                         _p(tris), nt, C.byref(nv))
    return verts[:nv.value], normals[:nv.value], tris[:nt]


def render_points(xyz, pose, intrinsics, image, valid=None, colors=None, want_color=True):
    """Z-buffer render of world points (N, 3) float32 into (rows, cols) depth
    and (rows, cols, 3) colour (None when want_color is False)."""
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    n = xyz.shape[0]
    cols, rows = image
    fx, fy, cx, cy = [float(np.float32(v)) for v in intrinsics]
    depth = np.empty((rows, cols), np.float32)
    color = np.empty((rows, cols, 3), np.uint8) if want_color else None
    valid = np.ascontiguousarray(valid, np.uint8) if valid is not None else None
    colors = np.ascontiguousarray(colors, np.uint8) if colors is not None else None
    lib().oracle_render_points(_p(xyz), _p(valid), _p(colors), n, _p(pose_arr(pose)), fx, fy, cx,
                               cy, rows, cols, cols, _p(depth), _p(color), cols * 3)
    return depth, color


def point_cloud_from_depth_f64(depth, intrinsics, transform, dmin=0.0, dmax=1e30):
    depth = np.ascontiguousarray(depth, np.float32)
    rows, cols = depth.shape
    fx, fy, cx, cy = [float(v) for v in intrinsics]
    xyz = np.empty((rows * cols, 3), np.float64)
    idx = np.empty(rows * cols, np.int64)
    n = lib().oracle_point_cloud_from_depth_f64(_p(depth), rows, cols, cols, fx, fy, cx, cy,
                                                _p(pose_arr(transform)), dmin, dmax, _p(xyz),
                                                _p(idx), rows * cols)
    return xyz[:n], idx[:n]


# ---- differentiable integrate (float64) ----------------------------------------

def rt_from_poses(poses) -> np.ndarray:
    """(F, 12) float64 world->camera [R row-major | t] from (F, 7) poses."""
    poses = np.asarray(poses, np.float64).reshape(-1, 7)
    out = np.empty((poses.shape[0], 12))
    for f, p in enumerate(poses):
        w, x, y, z = p[:4] / np.linalg.norm(p[:4])
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        out[f, :9] = R.ravel()
        out[f, 9:] = p[4:]
    return out


def diff_forward(vol, depth, Rt, nthreads=0):
    depth = np.ascontiguousarray(depth, np.float32)
    F, rows, cols = depth.shape
    D = np.full(shape(vol), float(np.float32(vol.truncation)), np.float64)
    W = np.zeros(shape(vol), np.float64)
    Rt = np.ascontiguousarray(Rt, np.float64)
    lib().oracle_diff_forward_f64(C.byref(vol), _p(D), _p(W), _p(depth), rows, cols, _p(Rt), F,
                                  nthreads)
    return D, W


def diff_forward_f32(vol, depth, rt32, nthreads=0):
    """Float32 bilinear forward from init (D = trunc, W = 0); rt32 (F, 12)."""
    depth = np.ascontiguousarray(depth, np.float32)
    F, rows, cols = depth.shape
    D, W = init(vol)
    lib().oracle_diff_forward_f32(C.byref(vol), _p(D), _p(W), _p(depth), rows, cols,
                                  _p(np.ascontiguousarray(rt32, np.float32)), F, nthreads)
    return D, W


def trilinear_support(vol, point):
    """(ids (8,) int64, coef (8,) float64) of a world point, or None when the
    support leaves the grid."""
    pw = np.ascontiguousarray(point, np.float64)
    ids = np.empty(8, np.int64)
    coef = np.empty(8, np.float64)
    if not lib().oracle_trilinear_support(C.byref(vol), _p(pw), _p(ids), _p(coef)):
        return None
    return ids, coef


def surface_loss(vol, D, points, weights=None, want_lambda=True):
    pts = np.ascontiguousarray(points, np.float64)
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    lam = np.empty(shape(vol), np.float64) if want_lambda else None
    L = lib().oracle_surface_loss_f64(C.byref(vol), _p(np.ascontiguousarray(D, np.float64)),
                                      _p(pts), _p(w), pts.shape[0], _p(lam))
    return (L, lam) if want_lambda else L


def diff_adjoint(vol, depth, Rt, lam, want_depth=True):
    depth = np.ascontiguousarray(depth, np.float32)
    F, rows, cols = depth.shape
    gp = np.empty((F, 6), np.float64)
    gd = np.empty((F, rows, cols), np.float64) if want_depth else None
    lib().oracle_diff_adjoint_f64(C.byref(vol), _p(depth), rows, cols,
                                  _p(np.ascontiguousarray(Rt, np.float64)), F,
                                  _p(np.ascontiguousarray(lam, np.float64)), _p(gp), _p(gd))
    return gp, gd


def frame_loss(vol, depth, Rt, points, pairs, weights=None, want_grad=True, tau=0.0):
    """Model-to-frame loss (float64): (L, dL/dxi (F, 6) or None, residuals
    (K,)) for pairs (K, 2) int32 of (frame, point); residuals of |r| >= tau
    (tau > 0) count tau^2 without gradient."""
    depth = np.ascontiguousarray(depth, np.float32)
    F, rows, cols = depth.shape
    pts = np.ascontiguousarray(points, np.float64)
    pr = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    gp = np.zeros((F, 6), np.float64) if want_grad else None
    res = np.empty(pr.shape[0], np.float64)
    L = lib().oracle_frame_loss_f64(C.byref(vol), _p(depth), rows, cols,
                                    _p(np.ascontiguousarray(Rt, np.float64)), F, _p(pts), _p(pr),
                                    _p(w), pr.shape[0], float(tau), _p(gp), _p(res))
    return L, gp, res


def frame_normal_matrix(vol, depth, Rt, points, pairs, normals=None, weights=None, tau=0.0,
                        want_grad=False):
    """(F, 6, 6) Gauss-Newton normal matrix of the model-to-frame term (and,
    with want_grad, the (F, 6) gradient of the same linearisation): the
    point-to-plane linearisation with world `normals`, or the bilinear depth
    slopes without."""
    depth = np.ascontiguousarray(depth, np.float32)
    F, rows, cols = depth.shape
    pts = np.ascontiguousarray(points, np.float64)
    nrm = None if normals is None else np.ascontiguousarray(normals, np.float64)
    pr = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    H = np.zeros((F, 6, 6), np.float64)
    g = np.zeros((F, 6), np.float64) if want_grad else None
    lib().oracle_frame_normal_matrix_f64(C.byref(vol), _p(depth), rows, cols,
                                         _p(np.ascontiguousarray(Rt, np.float64)), F, _p(pts),
                                         _p(nrm), _p(pr), _p(w), pr.shape[0], float(tau), _p(H),
                                         _p(g))
    return (H, g) if want_grad else H


def gn_update(Rt, grad, H, step=1.0, damping=0.0):
    """float32 (F, 12) poses after one damped Gauss-Newton step per frame:
    M = H + damping diag(H) + 1e-12 max(diag H) I factored M = L L^T column
    by column, a column whose pivot is at most 1e-9 max(diag M) skipped (its
    coordinate of x is 0), x = M^-1 grad over the rest, T <- Exp(-step x) T in
    float64, R's rows re-orthonormalised (Gram-Schmidt, third row from the
    cross product).  Frames with H = 0 keep their pose.  H is (F, 6, 6)."""
    Rt = np.asarray(Rt, np.float32)
    out = Rt.copy()
    for f in range(Rt.shape[0]):
        M = np.array(H[f], np.float64)
        dmax = np.max(np.diag(M))
        if not dmax > 0.0:
            continue
        M[np.diag_indices(6)] = np.diag(M) * (1.0 + damping) + 1e-12 * dmax
        mmax = np.max(np.diag(M))
        Lc = np.zeros((6, 6))
        live = np.zeros(6, bool)
        for c in range(6):
            d = M[c, c] - np.dot(Lc[c, :c], Lc[c, :c])
            live[c] = d > 1e-9 * mmax
            if live[c]:
                Lc[c, c] = np.sqrt(d)
                for r in range(c + 1, 6):
                    Lc[r, c] = (M[r, c] - np.dot(Lc[r, :c], Lc[c, :c])) / Lc[c, c]
        g = np.asarray(grad[f], np.float64)
        y = np.zeros(6)
        for r in range(6):
            if live[r]:
                y[r] = (g[r] - np.dot(Lc[r, :r], y[:r])) / Lc[r, r]
        x = np.zeros(6)
        for r in range(5, -1, -1):
            if live[r]:
                x[r] = (y[r] - np.dot(Lc[r + 1:, r], x[r + 1:])) / Lc[r, r]
        T = se3_exp_left(Rt[f].astype(np.float64), -step * x)
        R = T[:9].reshape(3, 3).copy()
        R[0] /= np.linalg.norm(R[0])
        R[1] -= np.dot(R[0], R[1]) * R[0]
        R[1] /= np.linalg.norm(R[1])
        R[2] = np.cross(R[0], R[1])
        out[f, :9] = R.ravel().astype(np.float32)
        out[f, 9:] = T[9:].astype(np.float32)
    return out


def stable_points(vol, depth, Rt, points, h=1e-4, kink=1e-5):
    depth = np.ascontiguousarray(depth, np.float32)
    F, rows, cols = depth.shape
    pts = np.ascontiguousarray(points, np.float64)
    w = np.empty(pts.shape[0], np.float64)
    lib().oracle_stable_points(C.byref(vol), _p(depth), rows, cols,
                               _p(np.ascontiguousarray(Rt, np.float64)), F, _p(pts), pts.shape[0],
                               float(h), float(kink), _p(w))
    return w


def se3_exp_left(Rt, xi):
    """Exp(xi) T for xi = (rho, omega) (float64, Rodrigues)."""
    R = Rt[:9].reshape(3, 3)
    t = Rt[9:]
    rho, om = np.asarray(xi[:3], np.float64), np.asarray(xi[3:], np.float64)
    th = np.linalg.norm(om)
    K = np.array([[0, -om[2], om[1]], [om[2], 0, -om[0]], [-om[1], om[0], 0]])
    if th < 1e-12:
        Re, V = np.eye(3) + K, np.eye(3) + 0.5 * K
    else:
        Re = np.eye(3) + np.sin(th) / th * K + (1 - np.cos(th)) / th ** 2 * K @ K
        V = np.eye(3) + (1 - np.cos(th)) / th ** 2 * K + (th - np.sin(th)) / th ** 3 * K @ K
    out = np.empty(12)
    out[:9] = (Re @ R).ravel()
    out[9:] = Re @ t + V @ rho
    return out

"""Synthetic workloads of the TSDF fusion path.

Each workload fixes a volume, a camera path, a scene of analytic primitives
and a depth-noise model; depth frames are ray-cast by librecon_scene.so.
All randomness comes from a project-owned splitmix64 stream, so every frame
is bit-identical on every machine.

  plain_small      128^3, 8 mm, trunc 4 voxels, cap 128; sphere r=0.3 m at the
                   origin + plane z=-0.35; 10 cameras on a 1 m arc 10 deg apart
                   looking at the origin; 640x480, f=525, c=(319.5, 239.5);
                   noise 1 mm, 2 % dropout.
  plain_large      512^3, 4 mm, offset (-1.024,-1.024,-0.1), trunc 5 voxels,
                   cap 128; 40 seeded boxes (sides 0.05-0.6 m) on the floor
                   z=0; 300 frames on a seeded smooth orbit (radius
                   1.2-1.8 m); noise 0.5 mm + 0.0015 z^2, 3 % dropout; mesh
                   after every 50 frames.
  batched_uneven   64 scenes of 1-5 spheres/boxes, grids {64..256}^3,
                   1-40 frames, 320x240 / 640x480 / 1280x720, f = 0.82 W.
  differentiable   256^3, 6 mm, the plain_small scene, 20 frames.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import make_config

PKG = Path(__file__).resolve().parents[1]

PRIM_SPHERE, PRIM_BOX, PRIM_PLANE_Z = 0, 1, 2
MASK64 = (1 << 64) - 1


class Prim(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad_", C.c_int32), ("p", C.c_double * 6)]


class Noise(C.Structure):
    _fields_ = [("sigma0", C.c_double), ("sigma_z2", C.c_double), ("dropout", C.c_double),
                ("seed", C.c_uint64)]


def splitmix64(x: int) -> int:
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & MASK64
    return x ^ (x >> 31)


class Rng:
    """Sequential splitmix64 stream."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next(self) -> int:
        self.state = (self.state + 1) & MASK64
        return splitmix64(self.state)

    def uniform(self, lo=0.0, hi=1.0) -> float:
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / 9007199254740992.0))

    def integer(self, lo: int, hi: int) -> int:  # inclusive
        return lo + int(self.uniform() * (hi - lo + 1)) if hi > lo else lo


def quat_from_matrix(R: np.ndarray) -> np.ndarray:
    """Unit quaternion (w, x, y, z), w >= 0, of a rotation matrix."""
    m = R
    tr = m[0, 0] + m[1, 1] + m[2, 2]
    if tr > 0:
        s = math.sqrt(tr + 1.0) * 2
        q = [0.25 * s, (m[2, 1] - m[1, 2]) / s, (m[0, 2] - m[2, 0]) / s, (m[1, 0] - m[0, 1]) / s]
    elif m[0, 0] > m[1, 1] and m[0, 0] > m[2, 2]:
        s = math.sqrt(1.0 + m[0, 0] - m[1, 1] - m[2, 2]) * 2
        q = [(m[2, 1] - m[1, 2]) / s, 0.25 * s, (m[0, 1] + m[1, 0]) / s, (m[0, 2] + m[2, 0]) / s]
    elif m[1, 1] > m[2, 2]:
        s = math.sqrt(1.0 + m[1, 1] - m[0, 0] - m[2, 2]) * 2
        q = [(m[0, 2] - m[2, 0]) / s, (m[0, 1] + m[1, 0]) / s, 0.25 * s, (m[1, 2] + m[2, 1]) / s]
    else:
        s = math.sqrt(1.0 + m[2, 2] - m[0, 0] - m[1, 1]) * 2
        q = [(m[1, 0] - m[0, 1]) / s, (m[0, 2] + m[2, 0]) / s, (m[1, 2] + m[2, 1]) / s, 0.25 * s]
    q = np.array(q, dtype=np.float64)
    q /= np.linalg.norm(q)
    return -q if q[0] < 0 else q


def matrix_from_quat(q) -> np.ndarray:
    w, x, y, z = np.asarray(q, np.float64) / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def look_at_pose(centre, target, up=(0.0, 0.0, 1.0)) -> np.ndarray:
    """World->camera pose {qw,qx,qy,qz,tx,ty,tz} of a camera at `centre`
    looking at `target`, OpenCV axes (x right, y down, z forward)."""
    c = np.asarray(centre, np.float64)
    z = np.asarray(target, np.float64) - c
    z /= np.linalg.norm(z)
    x = np.cross(z, np.asarray(up, np.float64))
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    R_cw = np.stack([x, y, z])  # rows: camera axes in world coordinates
    t = -R_cw @ c
    return np.concatenate([quat_from_matrix(R_cw), t])


def pose_matrix(pose) -> tuple[np.ndarray, np.ndarray]:
    pose = np.asarray(pose, np.float64)
    return matrix_from_quat(pose[:4]), pose[4:7].copy()


@dataclass
class Workload:
    name: str
    grid: tuple
    voxel_size: float
    offset: tuple
    truncation: float
    max_weight: float
    intrinsics: tuple          # fx, fy, cx, cy
    image: tuple               # cols, rows
    prims: list                # (kind, params)
    poses: np.ndarray          # (F, 7)
    noise: tuple               # sigma0, sigma_z2, dropout, seed
    mesh_every: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def num_frames(self) -> int:
        return int(self.poses.shape[0])

    @property
    def num_voxels(self) -> int:
        return int(np.prod(self.grid))

    def config(self):
        return make_config(self.grid, voxel_size=self.voxel_size, offset=self.offset,
                           max_weight=self.max_weight, intrinsics=self.intrinsics,
                           image=self.image, truncation=self.truncation)

    def prim_array(self):
        arr = (Prim * max(len(self.prims), 1))()
        for i, (kind, p) in enumerate(self.prims):
            arr[i].kind = kind
            for k, v in enumerate(p):
                arr[i].p[k] = float(v)
        return arr

    def render(self, frames=None, noise=True, nthreads=0) -> np.ndarray:
        """Depth frames (F, rows, cols) float32 in metres."""
        frames = np.arange(self.num_frames) if frames is None else np.asarray(frames)
        cols, rows = self.image
        out = np.empty((len(frames), rows, cols), np.float32)
        poses = np.ascontiguousarray(self.poses[frames], np.float64)
        ids = np.ascontiguousarray(frames, np.uint64)
        nz = None
        if noise:
            nz = Noise(*self.noise[:3], self.noise[3])
        fx, fy, cx, cy = self.intrinsics
        rc = scene_lib().recon_scene_render(
            self.prim_array(), len(self.prims), poses.ctypes.data, ids.ctypes.data, len(frames),
            fx, fy, cx, cy, cols, rows, C.byref(nz) if nz is not None else None, out.ctypes.data,
            nthreads)
        if rc != 0:
            raise RuntimeError(f"recon_scene_render failed ({rc})")
        return out

    def sdf(self, points: np.ndarray) -> np.ndarray:
        lib = scene_lib()
        prims = self.prim_array()
        pts = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
        return np.array([lib.recon_scene_sdf(prims, len(self.prims), p.ctypes.data) for p in pts])


_scene = None


def scene_lib() -> C.CDLL:
    global _scene
    if _scene is None:
        p = PKG / "lib" / "librecon_scene.so"
        if not p.exists():
            raise RuntimeError(f"{p} is missing: build first")
        _scene = C.CDLL(str(p))
        _scene.recon_scene_render.restype = C.c_int
        _scene.recon_scene_render.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                              C.c_int32, C.c_double, C.c_double, C.c_double,
                                              C.c_double, C.c_int32, C.c_int32, C.c_void_p,
                                              C.c_void_p, C.c_int32]
        _scene.recon_scene_sdf.restype = C.c_double
        _scene.recon_scene_sdf.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
    return _scene


# ---- workloads ---------------------------------------------------------------

def arc_poses(n: int, step_deg: float = 10.0, radius: float = 1.0, height: float = 0.0,
              start_deg: float = 0.0, target=(0.0, 0.0, 0.0)) -> np.ndarray:
    out = []
    for k in range(n):
        th = math.radians(start_deg + step_deg * k)
        out.append(look_at_pose((radius * math.cos(th), radius * math.sin(th), height), target))
    return np.array(out)


def plain_small(num_frames: int = 10, grid: int = 128, voxel_size: float = 0.008) -> Workload:
    half = grid * voxel_size / 2
    return Workload(
        name=f"plain_small_{grid}^3x{num_frames}",
        grid=(grid, grid, grid), voxel_size=voxel_size, offset=(-half, -half, -half),
        truncation=4 * voxel_size, max_weight=128.0,
        intrinsics=(525.0, 525.0, 319.5, 239.5), image=(640, 480),
        prims=[(PRIM_SPHERE, (0.0, 0.0, 0.0, 0.3)), (PRIM_PLANE_Z, (-0.35,))],
        poses=arc_poses(num_frames), noise=(0.001, 0.0, 0.02, 0xC0FFEE))


def differentiable(num_frames: int = 20) -> Workload:
    w = plain_small(num_frames=num_frames, grid=256, voxel_size=0.006)
    w.name = f"differentiable_256^3x{num_frames}"
    return w


def catmull_rom(p0, p1, p2, p3, t):
    t2, t3 = t * t, t * t * t
    return 0.5 * ((2 * p1) + (-p0 + p2) * t + (2 * p0 - 5 * p1 + 4 * p2 - p3) * t2 +
                  (-p0 + 3 * p1 - 3 * p2 + p3) * t3)


def plain_large(num_frames: int = 300, seed: int = 0x5EED0001) -> Workload:
    rng = Rng(seed)
    prims = [(PRIM_PLANE_Z, (0.0,))]
    for _ in range(40):
        sx, sy, sz = (rng.uniform(0.05, 0.6) for _ in range(3))
        cx, cy = rng.uniform(-0.75, 0.75), rng.uniform(-0.75, 0.75)
        prims.append((PRIM_BOX, (cx - sx / 2, cy - sy / 2, 0.0, cx + sx / 2, cy + sy / 2, sz)))
    # periodic Catmull-Rom spline over 8 control points: radius 1.2-1.8 m,
    # height 0.4-1.2 m, one revolution over the frames
    nctl = 8
    rad = [rng.uniform(1.2, 1.8) for _ in range(nctl)]
    hgt = [rng.uniform(0.4, 1.2) for _ in range(nctl)]
    phase = rng.uniform(0.0, 2 * math.pi)
    poses = []
    for f in range(num_frames):
        s = nctl * f / num_frames
        i = int(s)
        t = s - i
        idx = [(i - 1) % nctl, i % nctl, (i + 1) % nctl, (i + 2) % nctl]
        r = min(max(catmull_rom(*[rad[j] for j in idx], t), 1.2), 1.8)
        h = min(max(catmull_rom(*[hgt[j] for j in idx], t), 0.4), 1.2)
        th = phase + 2 * math.pi * f / num_frames
        poses.append(look_at_pose((r * math.cos(th), r * math.sin(th), h), (0.0, 0.0, 0.15)))
    return Workload(
        name=f"plain_large_512^3x{num_frames}",
        grid=(512, 512, 512), voxel_size=0.004, offset=(-1.024, -1.024, -0.1),
        truncation=5 * 0.004, max_weight=128.0,
        intrinsics=(525.0, 525.0, 319.5, 239.5), image=(640, 480),
        prims=prims, poses=np.array(poses), noise=(0.0005, 0.0015, 0.03, seed ^ 0xABCDEF),
        mesh_every=50)


RESOLUTIONS = [(320, 240), (640, 480), (1280, 720)]
GRID_SIZES = [64, 96, 128, 192, 256]


def batched_uneven(num_volumes: int = 64, seed: int = 0x5EED0002) -> list[Workload]:
    rng = Rng(seed)
    out = []
    extent = 1.0
    for v in range(num_volumes):
        s = GRID_SIZES[rng.integer(0, 4)]
        frames = rng.integer(1, 40)
        cols, rows = RESOLUTIONS[rng.integer(0, 2)]
        vs = extent / s
        prims = []
        for _ in range(rng.integer(1, 5)):
            if rng.uniform() < 0.5:
                prims.append((PRIM_SPHERE, (rng.uniform(-0.25, 0.25), rng.uniform(-0.25, 0.25),
                                            rng.uniform(-0.25, 0.25), rng.uniform(0.05, 0.2))))
            else:
                c = [rng.uniform(-0.25, 0.25) for _ in range(3)]
                h = [rng.uniform(0.03, 0.15) for _ in range(3)]
                prims.append((PRIM_BOX, tuple(c[i] - h[i] for i in range(3)) +
                              tuple(c[i] + h[i] for i in range(3))))
        start = rng.uniform(0.0, 360.0)
        elev = rng.uniform(-0.3, 0.5)
        poses = arc_poses(frames, step_deg=9.0, radius=1.2, height=elev, start_deg=start)
        fx = 0.82 * cols
        out.append(Workload(
            name=f"batched_{v}_{s}^3x{frames}@{cols}x{rows}",
            grid=(s, s, s), voxel_size=vs, offset=(-extent / 2,) * 3, truncation=4 * vs,
            max_weight=128.0, intrinsics=(fx, fx, cols / 2 - 0.5, rows / 2 - 0.5),
            image=(cols, rows), prims=prims, poses=poses,
            noise=(0.001, 0.0, 0.02, seed ^ (v * 0x9E37)), meta={"index": v}))
    return out


def tiny(grid=(16, 12, 10), voxel_size=0.05, image=(32, 24), num_frames=3, seed=7) -> Workload:
    """A small scene for exhaustive checks: sphere + floor seen by a short arc."""
    gx, gy, gz = grid
    off = (-gx * voxel_size / 2, -gy * voxel_size / 2, -gz * voxel_size / 2)
    cols, rows = image
    f = 0.9 * cols
    return Workload(
        name=f"tiny_{gx}x{gy}x{gz}", grid=tuple(grid), voxel_size=voxel_size, offset=off,
        truncation=3 * voxel_size, max_weight=2.0, intrinsics=(f, f, cols / 2 - 0.5, rows / 2 - 0.5),
        image=image, prims=[(PRIM_SPHERE, (0.0, 0.0, 0.0, 0.18)), (PRIM_PLANE_Z, (-0.2,))],
        poses=arc_poses(num_frames, step_deg=25.0, radius=1.0, height=0.2, start_deg=10.0),
        noise=(0.002, 0.0, 0.05, seed))


def _union_sdf(prims, p: np.ndarray) -> np.ndarray:
    """Signed distance to the union of primitives, vectorised over (n, 3)
    points: the same conventions as recon_scene_sdf (solid below a z plane)."""
    best = np.full(p.shape[0], np.inf)
    for kind, q in prims:
        if kind == PRIM_SPHERE:
            d = np.linalg.norm(p - np.asarray(q[:3]), axis=1) - q[3]
        elif kind == PRIM_BOX:
            lo, hi = np.asarray(q[:3]), np.asarray(q[3:6])
            r = np.abs(p - 0.5 * (lo + hi)) - 0.5 * (hi - lo)
            d = np.linalg.norm(np.maximum(r, 0.0), axis=1) + np.minimum(r.max(axis=1), 0.0)
        else:
            d = p[:, 2] - q[0]
        best = np.minimum(best, d)
    return best


def _sample_prim_surface(kind, q, m: int, lo: np.ndarray, hi: np.ndarray,
                         rng: np.random.Generator) -> tuple[np.ndarray, np.ndarray, float]:
    """m points uniform on one primitive's surface (a z plane: its part over
    the box [lo, hi]), their outward normals and that surface's area."""
    if kind == PRIM_SPHERE:
        v = rng.normal(size=(m, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        return np.asarray(q[:3]) + q[3] * v, v, 4.0 * math.pi * q[3] ** 2
    if kind == PRIM_BOX:
        b0, b1 = np.asarray(q[:3], np.float64), np.asarray(q[3:6], np.float64)
        s = b1 - b0
        areas = np.array([s[1] * s[2], s[1] * s[2], s[0] * s[2], s[0] * s[2], s[0] * s[1], s[0] * s[1]])
        face = rng.choice(6, size=m, p=areas / areas.sum())
        pts = b0 + rng.uniform(size=(m, 3)) * s
        axis = face // 2
        pts[np.arange(m), axis] = np.where(face % 2 == 0, b0[axis], b1[axis])
        nrm = np.zeros((m, 3))
        nrm[np.arange(m), axis] = np.where(face % 2 == 0, -1.0, 1.0)
        return pts, nrm, float(areas.sum())
    pts = np.c_[rng.uniform(lo[0], hi[0], m), rng.uniform(lo[1], hi[1], m), np.full(m, q[0])]
    return pts, np.tile([0.0, 0.0, 1.0], (m, 1)), float((hi[0] - lo[0]) * (hi[1] - lo[1]))


def surface_points(w: Workload, n: int, seed: int = 0x5EED0003, margin_voxels: float = 1.5,
                   with_normals: bool = False):
    """n points (float64, (n, 3)) uniform over the visible surface of the
    scene inside the volume shrunk by `margin_voxels`: each primitive's
    surface, less the parts inside another primitive, drawn in proportion to
    its clipped area.  With `with_normals`, also their outward normals."""
    rng = np.random.default_rng(seed)
    vs = w.voxel_size
    lo = np.asarray(w.offset, np.float64) + margin_voxels * vs
    hi = np.asarray(w.offset, np.float64) + np.asarray(w.grid, np.float64) * vs - margin_voxels * vs
    pools, areas = [], []
    m = max(4 * n, 4096)
    for kind, q in w.prims:
        pts, nrm, area = _sample_prim_surface(kind, q, m, lo, hi, rng)
        keep = np.all((pts >= lo) & (pts <= hi), axis=1) & (_union_sdf(w.prims, pts) >= -1e-9)
        pools.append((pts[keep], nrm[keep]))
        areas.append(area * keep.mean())
    areas = np.asarray(areas)
    if areas.sum() <= 0:
        return (np.zeros((0, 3)), np.zeros((0, 3))) if with_normals else np.zeros((0, 3))
    counts = rng.multinomial(n, areas / areas.sum())
    picks = [(pool, nrm, rng.choice(len(pool), size=c, replace=c > len(pool)))
             for (pool, nrm), c in zip(pools, counts) if c > 0]
    pts = np.concatenate([pool[i] for pool, _, i in picks], axis=0)
    if not with_normals:
        return pts
    return pts, np.concatenate([nrm[i] for _, nrm, i in picks], axis=0)


def perturb_pose(pose, rot_rad: float, trans_m: float, rng: Rng) -> np.ndarray:
    """Left perturbation of a world->camera pose {qw,qx,qy,qz,tx,ty,tz}: a
    rotation by up to `rot_rad` about a random axis through the camera centre
    and a shift of the centre by up to `trans_m` (uniform magnitudes)."""
    R, t = pose_matrix(pose)
    axis = np.array([rng.uniform(-1, 1) for _ in range(3)])
    axis /= max(np.linalg.norm(axis), 1e-12)
    ang = rng.uniform(0.0, rot_rad)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    dR = np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K
    d = np.array([rng.uniform(-1, 1) for _ in range(3)])
    d *= rng.uniform(0.0, trans_m) / max(np.linalg.norm(d), 1e-12)
    R2 = dR @ R
    t2 = dR @ t + d
    return np.r_[quat_from_matrix(R2), t2]


def pose_refinement(num_fragments: int = 32, frames: int = 10, seed: int = 0x5EED0004,
                    grid: int = 128, voxel_size: float = 0.008) -> list[Workload]:
    """Fragments for batched pose refinement: each a 128^3 volume (8 mm) of a
    distinct seeded scene on a floor (z = -0.3 m): one box, two spheres and
    up to two more boxes or spheres, seen by `frames` 640x480 cameras on an
    arc 10 degrees apart.  Every frame then sees surfaces whose normals fix
    all six pose coordinates (a box or the floor alone leaves in-plane shifts
    free, a sphere alone rotations about its centre).  `poses` are the
    ground truth; meta["init"] are the poses perturbed by up to 1 degree and
    1 cm."""
    rng = Rng(seed)
    out = []
    half = grid * voxel_size / 2
    for v in range(num_fragments):
        prims = []
        nprim = rng.integer(3, 5)
        for k in range(nprim):
            if k == 0 or (k >= 3 and rng.uniform() < 0.5):
                c = [rng.uniform(-0.2, 0.2) for _ in range(3)]
                h = [rng.uniform(0.05, 0.18) for _ in range(3)]
                prims.append((PRIM_BOX, tuple(c[i] - h[i] for i in range(3)) +
                              tuple(c[i] + h[i] for i in range(3))))
            else:
                prims.append((PRIM_SPHERE, (rng.uniform(-0.2, 0.2), rng.uniform(-0.2, 0.2),
                                            rng.uniform(-0.2, 0.2), rng.uniform(0.06, 0.2))))
        prims.append((PRIM_PLANE_Z, (-0.3,)))
        poses = arc_poses(frames, step_deg=10.0, radius=1.1, height=rng.uniform(0.1, 0.6),
                          start_deg=rng.uniform(0.0, 360.0))
        init = np.stack([perturb_pose(p, math.radians(1.0), 0.01, rng) for p in poses])
        out.append(Workload(
            name=f"pose_refinement_{v}_{grid}^3x{frames}", grid=(grid, grid, grid),
            voxel_size=voxel_size, offset=(-half, -half, -half), truncation=4 * voxel_size,
            max_weight=128.0, intrinsics=(525.0, 525.0, 319.5, 239.5), image=(640, 480),
            prims=prims, poses=poses, noise=(0.001, 0.0, 0.02, seed ^ (v * 0x9E37)),
            meta={"index": v, "init": init}))
    return out


def pose_errors(est, gt) -> tuple[np.ndarray, np.ndarray]:
    """(rotation error in degrees, camera-centre error in metres) per frame of
    world->camera poses given as (F, 12) [R | t] or (F, 7) quaternion poses."""
    def mats(p):
        p = np.asarray(p, np.float64)
        if p.shape[-1] == 12:
            return p[:, :9].reshape(-1, 3, 3), p[:, 9:]
        Rs, ts = zip(*[pose_matrix(q) for q in p])
        return np.stack(Rs), np.stack(ts)
    Re, te = mats(est)
    Rg, tg = mats(gt)
    dR = Re @ np.transpose(Rg, (0, 2, 1))
    # angle from both the symmetric and the skew part (exact near zero)
    skew = np.stack([dR[:, 2, 1] - dR[:, 1, 2], dR[:, 0, 2] - dR[:, 2, 0], dR[:, 1, 0] - dR[:, 0, 1]], 1)
    ang = np.degrees(np.arctan2(0.5 * np.linalg.norm(skew, axis=1),
                                0.5 * (np.trace(dR, axis1=1, axis2=2) - 1)))
    ce = -np.einsum("fji,fj->fi", Re, te)
    cg = -np.einsum("fji,fj->fi", Rg, tg)
    return ang, np.linalg.norm(ce - cg, axis=1)


def observed_points(w: Workload, W: np.ndarray, points: np.ndarray) -> np.ndarray:
    """float32 weights: 1 for points whose eight trilinear support voxels
    (lower corner floor(q/vs - 0.5), q = p - offset) all have weight > 0 in
    the (z, y, x) weight volume W, else 0."""
    g = np.asarray(w.grid)
    vs = np.float32(w.voxel_size)
    q = (np.asarray(points, np.float64) - np.asarray(w.offset, np.float64)) / np.float64(vs) - 0.5
    lo = np.floor(q).astype(np.int64)
    inside = np.all((lo >= 0) & (lo + 1 <= g - 1), axis=1)
    lo = np.clip(lo, 0, g - 2)
    ok = inside.copy()
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                ok &= W[lo[:, 2] + dz, lo[:, 1] + dy, lo[:, 0] + dx] > 0
    return ok.astype(np.float32)



def visible_points(w: Workload, depth: np.ndarray, poses, points: np.ndarray, tol: float,
                   min_frames: int = 1, normals: np.ndarray | None = None,
                   min_cos: float = 0.0, window: int = 0) -> np.ndarray:
    """float32 weights: 1 for points that at least `min_frames` frames see,
    i.e. whose camera z lies within `tol` of the depth at their projected
    pixel (nearest pixel, valid depth, in front of the camera).  With
    `normals`, a frame counts only when |n . view direction| >= min_cos; with
    `window` > 0, only when every depth of the (2 window + 1)^2 pixels around
    the projection is valid and within `tol` of the point (no silhouette)."""
    fx, fy, cx, cy = w.intrinsics
    cols, rows = w.image
    P = np.asarray(points, np.float64)
    seen = np.zeros(len(P), np.int32)
    for f, pose in enumerate(np.asarray(poses)):
        if pose.shape[-1] == 12:
            R, t = pose[:9].reshape(3, 3).astype(np.float64), pose[9:].astype(np.float64)
        else:
            R, t = pose_matrix(pose)
        pc = P @ R.T + t
        z = pc[:, 2]
        with np.errstate(all="ignore"):
            u = np.floor(pc[:, 0] * fx / z + cx + 0.5).astype(np.int64)
            v = np.floor(pc[:, 1] * fy / z + cy + 0.5).astype(np.int64)
        r = window
        ok = (z > 0) & (u >= r) & (u < cols - r) & (v >= r) & (v < rows - r)
        good = ok.copy()
        for dy in range(-r, r + 1):
            for dx in range(-r, r + 1):
                d = np.zeros(len(P))
                d[ok] = depth[f][v[ok] + dy, u[ok] + dx]
                good &= (d > 0) & (np.abs(d - z) < tol)
        if normals is not None:
            view = pc / np.maximum(np.linalg.norm(pc, axis=1, keepdims=True), 1e-12)
            nc = np.asarray(normals, np.float64) @ R.T
            good &= np.abs(np.sum(nc * view, axis=1)) >= min_cos
        seen += good.astype(np.int32)
    return (seen >= min_frames).astype(np.float32)


def refinement_points(w: Workload, depth: np.ndarray, n: int, seed: int,
                      with_normals: bool = False):
    """Surface points for pose refinement and their float32 weights: n points
    uniform on the scene's surfaces, weight 1 where at least 3 frames see the
    point at its initial pose (meta["init"]) away from silhouettes (5x5
    depth window within 3 cm) and not at grazing angles (|n . view| >= 0.4);
    elsewhere the fused TSDF is not locally linear and its residuals would
    pull the poses.  With `with_normals`, also the points' normals."""
    pts, nrm = surface_points(w, n, seed=seed, with_normals=True)
    wts = visible_points(w, depth, w.meta.get("init", w.poses), pts, tol=0.03, min_frames=3,
                         normals=nrm, min_cos=0.4, window=2)
    return (pts, wts, nrm) if with_normals else (pts, wts)


def frame_pairs(w: Workload, depth: np.ndarray, points: np.ndarray, normals: np.ndarray,
                poses=None, tol: float = 0.03, window: int = 2, min_cos: float = 0.4) -> np.ndarray:
    """(P, 2) int32 (frame, point) pairs, grouped by frame, for the
    model-to-frame term of pose refinement: frame f sees point p at pose
    poses[f] (default meta["init"]) away from silhouettes (every depth of the
    (2 window + 1)^2 pixels around the projection valid and within `tol` of
    the point's camera z) and not at a grazing angle (|n . view| >= min_cos)."""
    fx, fy, cx, cy = w.intrinsics
    cols, rows = w.image
    poses = np.asarray(w.meta.get("init", w.poses) if poses is None else poses)
    P = np.asarray(points, np.float64)
    out = []
    for f, pose in enumerate(poses):
        R, t = pose_matrix(pose)
        pc = P @ R.T + t
        z = pc[:, 2]
        with np.errstate(all="ignore"):
            u = np.floor(pc[:, 0] * fx / z + cx + 0.5).astype(np.int64)
            v = np.floor(pc[:, 1] * fy / z + cy + 0.5).astype(np.int64)
        r = window
        ok = (z > 0) & (u >= r) & (u < cols - r) & (v >= r) & (v < rows - r)
        good = ok.copy()
        for dy in range(-r, r + 1):
            for dx in range(-r, r + 1):
                d = np.zeros(len(P))
                d[ok] = depth[f][v[ok] + dy, u[ok] + dx]
                good &= (d > 0) & (np.abs(d - z) < tol)
        view = pc / np.maximum(np.linalg.norm(pc, axis=1, keepdims=True), 1e-12)
        nc = np.asarray(normals, np.float64) @ R.T
        good &= np.abs(np.sum(nc * view, axis=1)) >= min_cos
        idx = np.nonzero(good)[0]
        out.append(np.stack([np.full(len(idx), f), idx], axis=1))
    return np.ascontiguousarray(np.concatenate(out).astype(np.int32))


def mesh_rms(w: Workload, verts: np.ndarray) -> float:
    """RMS distance of mesh vertices (any shape (..., 3)) to the scene's
    analytic surface (its union signed distance)."""
    v = np.asarray(verts, np.float64).reshape(-1, 3)
    if len(v) == 0:
        return float("nan")
    return float(np.sqrt(np.mean(_union_sdf(w.prims, v) ** 2)))


def world_error_twists(est, gt) -> np.ndarray:
    """(F, 6) world-frame error of each world->camera pose: E = T_est^-1 T_gt,
    as (rho = its translation in metres, omega = its small-angle rotation
    vector in radians).  A rigid motion of a whole fragment against the
    model gives every frame the same E."""
    def mats(p):
        p = np.asarray(p, np.float64)
        return p[:, :9].reshape(-1, 3, 3), p[:, 9:]
    Re, te = mats(est)
    Rg, tg = mats(gt)
    RE = np.transpose(Re, (0, 2, 1)) @ Rg
    tE = np.einsum("fji,fj->fi", Re, tg - te)
    om = 0.5 * np.stack([RE[:, 2, 1] - RE[:, 1, 2], RE[:, 0, 2] - RE[:, 2, 0], RE[:, 1, 0] - RE[:, 0, 1]], 1)
    return np.c_[tE, om]


def observed_vertices(w: Workload, W: np.ndarray, verts: np.ndarray) -> np.ndarray:
    """Mask of mesh vertices whose eight surrounding voxel centres all have
    weight > 0 in the (z, y, x) weight volume W (marching cubes on D alone
    also meets the initial value of unobserved voxels behind surfaces)."""
    v = np.asarray(verts, np.float64).reshape(-1, 3)
    g = np.asarray(w.grid)
    q = (v - np.asarray(w.offset, np.float64)) / np.float64(np.float32(w.voxel_size)) - 0.5
    lo = np.floor(q).astype(np.int64)
    ok = np.all((lo >= 0) & (lo + 1 <= g - 1), axis=1)
    lo = np.clip(lo, 0, g - 2)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                ok &= W[lo[:, 2] + dz, lo[:, 1] + dy, lo[:, 0] + dx] > 0
    return ok


def morton_order(w: Workload, points: np.ndarray) -> np.ndarray:
    """Permutation putting points in 3-D Morton (Z-curve) order of their voxel
    cell: consecutive points then share support voxels, which keeps the
    loss's (voxel, value) pairs, their merge and the adjoint's accesses
    local."""
    q = np.floor((np.asarray(points, np.float64) - np.asarray(w.offset)) / w.voxel_size)
    c = np.clip(q, 0, 1023).astype(np.uint64)

    def spread(x):
        x = (x | (x << np.uint64(16))) & np.uint64(0x030000FF)
        x = (x | (x << np.uint64(8))) & np.uint64(0x0300F00F)
        x = (x | (x << np.uint64(4))) & np.uint64(0x030C30C3)
        x = (x | (x << np.uint64(2))) & np.uint64(0x09249249)
        return x
    key = spread(c[:, 0]) | (spread(c[:, 1]) << np.uint64(1)) | (spread(c[:, 2]) << np.uint64(2))
    return np.argsort(key, kind="stable")

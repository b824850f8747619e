"""The driver pieces of the reference's TSDF example
(examples/tsdf_surface_reconstruction/tsdf_surface.cpp): the flat YAML
parameter file, depth PNG decoding and the parameter file's camera pose.
No GPU calls."""
import ctypes as C
import struct
import zlib
from pathlib import Path

import numpy as np
import pytest
from PIL import Image
from scipy.spatial.transform import Rotation

from recon import ffi

LIB = Path(__file__).resolve().parents[1] / "cu-reconstruction_amd" / "lib" / "librecon_host.so"


@pytest.fixture(scope="module")
def host():
    ffi.load()  # librecon first (RTLD_GLOBAL), then the host library
    lib = C.CDLL(str(LIB))
    lib.recon_png_read_gray.restype = C.c_int
    lib.recon_png_read_gray.argtypes = [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                        C.POINTER(C.c_int32), C.c_void_p, C.c_int64]
    lib.recon_png_read_rgb.restype = C.c_int
    lib.recon_png_read_rgb.argtypes = [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                       C.c_void_p, C.c_int64]
    lib.recon_config_lookup.restype = C.c_int
    lib.recon_config_lookup.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int64]
    lib.recon_euler_pose.restype = None
    lib.recon_euler_pose.argtypes = [C.c_double] * 6 + [C.POINTER(C.c_double)]
    return lib


def read_png(lib, path):
    w, h, b = C.c_int32(), C.c_int32(), C.c_int32()
    st = lib.recon_png_read_gray(str(path).encode(), C.byref(w), C.byref(h), C.byref(b), None, 0)
    if st != 0:
        return st, None
    out = np.empty((h.value, w.value), np.uint16)
    st = lib.recon_png_read_gray(str(path).encode(), C.byref(w), C.byref(h), C.byref(b),
                                 out.ctypes.data, out.size)
    return st, out


def write_png(path, img, filt):
    """A grey PNG written here, every row with filter type `filt` (0-4) or
    cycling through them (filt = -1): independent of the decoder."""
    img = np.asarray(img)
    bits = 16 if img.dtype == np.uint16 else 8
    h, w = img.shape
    bpp = bits // 8
    raw = img.astype(">u2").tobytes() if bits == 16 else img.astype(np.uint8).tobytes()
    stride = w * bpp
    rows = [np.frombuffer(raw[y * stride:(y + 1) * stride], np.uint8).astype(np.int32) for y in range(h)]
    out = bytearray()
    prev = np.zeros(stride, np.int32)
    for y, cur in enumerate(rows):
        f = filt if filt >= 0 else y % 5
        left = np.r_[np.zeros(bpp, np.int32), cur[:-bpp]]
        upleft = np.r_[np.zeros(bpp, np.int32), prev[:-bpp]]
        if f == 0:
            enc = cur
        elif f == 1:
            enc = cur - left
        elif f == 2:
            enc = cur - prev
        elif f == 3:
            enc = cur - ((left + prev) >> 1)
        else:
            p = left + prev - upleft
            pa, pb, pc = np.abs(p - left), np.abs(p - prev), np.abs(p - upleft)
            pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, prev, upleft))
            enc = cur - pred
        out.append(f)
        out += (enc & 0xFF).astype(np.uint8).tobytes()
        prev = cur

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, bits, 0, 0, 0, 0))
    comp = zlib.compress(bytes(out), 6)
    data += chunk(b"IDAT", comp[: len(comp) // 2]) + chunk(b"IDAT", comp[len(comp) // 2:])  # split
    data += chunk(b"IEND", b"")
    Path(path).write_bytes(data)


@pytest.mark.parametrize("filt", [0, 1, 2, 3, 4, -1])
@pytest.mark.parametrize("bits", [8, 16])
def test_png_filters(host, tmp_path, filt, bits):
    rng = np.random.default_rng(filt + 10 * bits)
    dt = np.uint16 if bits == 16 else np.uint8
    img = rng.integers(0, 2 ** bits, size=(37, 53)).astype(dt)
    # smooth regions make the predictors matter
    img[5:20, 10:40] = (np.arange(30)[None, :] * 97 % (2 ** bits)).astype(dt)
    p = tmp_path / "a.png"
    write_png(p, img, filt)
    st, got = read_png(host, p)
    assert st == 0 and (got == img).all()


def test_png_written_by_pil(host, tmp_path):
    rng = np.random.default_rng(3)
    d16 = (rng.integers(0, 65535, size=(480, 640))).astype(np.uint16)
    d16[100:300, 200:500] = 12345
    Image.fromarray(d16).save(tmp_path / "d.png")
    st, got = read_png(host, tmp_path / "d.png")
    assert st == 0 and (got == d16).all()
    g8 = rng.integers(0, 255, size=(24, 31)).astype(np.uint8)
    Image.fromarray(g8).save(tmp_path / "g.png")
    st, got = read_png(host, tmp_path / "g.png")
    assert st == 0 and (got == g8).all()
    # colour images are not depth maps
    Image.fromarray(np.zeros((8, 8, 3), np.uint8)).save(tmp_path / "c.png")
    st, _ = read_png(host, tmp_path / "c.png")
    assert st == ffi.ERR_CONFIG
    (tmp_path / "x.png").write_bytes(b"not a png")
    assert read_png(host, tmp_path / "x.png")[0] == ffi.ERR_CONFIG


def test_rgb_png_written_by_pil(host, tmp_path):
    rng = np.random.default_rng(8)
    img = rng.integers(0, 256, (29, 41, 3)).astype(np.uint8)
    img[5:20, 5:30] = 77
    Image.fromarray(img, "RGB").save(tmp_path / "c.png")
    w, h = C.c_int32(), C.c_int32()
    assert host.recon_png_read_rgb(str(tmp_path / "c.png").encode(), C.byref(w), C.byref(h), None, 0) == 0
    out = np.empty((h.value, w.value, 3), np.uint8)
    assert host.recon_png_read_rgb(str(tmp_path / "c.png").encode(), C.byref(w), C.byref(h),
                                   out.ctypes.data, out.size) == 0
    assert (out == img).all()
    assert host.recon_png_read_rgb(str(tmp_path / "c.png").encode(), C.byref(w), C.byref(h),
                                   out.ctypes.data, out.size - 1) == ffi.ERR_CAPACITY
    # a grey image is not a colour image
    Image.fromarray(img[..., 0]).save(tmp_path / "g.png")
    assert host.recon_png_read_rgb(str(tmp_path / "g.png").encode(), C.byref(w), C.byref(h), None,
                                   0) == ffi.ERR_CONFIG


PARAMS = """%YAML:1.0

rgbd.dataset.path: "./some dir/scene # not a comment"
dataset.color.mode: 0 # 0 - false, 1 - true
# mesh_store.path: "./old.ply"
mesh_store.path: './mesh.ply'
camera.fx: 510.5
depth_scale: 0.0002
tsdf.volume.voxel_grid_size.x: 96
"""


def lookup(lib, text, key):
    buf = C.create_string_buffer(256)
    st = lib.recon_config_lookup(text.encode(), key.encode(), buf, 256)
    return st, buf.value.decode()


def test_parameter_file(host):
    assert lookup(host, PARAMS, "rgbd.dataset.path") == (0, "./some dir/scene # not a comment")
    assert lookup(host, PARAMS, "dataset.color.mode") == (0, "0")
    assert lookup(host, PARAMS, "mesh_store.path") == (0, "./mesh.ply")
    assert lookup(host, PARAMS, "camera.fx") == (0, "510.5")
    assert lookup(host, PARAMS, "tsdf.volume.voxel_grid_size.x") == (0, "96")
    assert lookup(host, PARAMS, "camera.fy")[0] == ffi.ERR_RANGE
    assert lookup(host, "key without colon\n", "a")[0] == ffi.ERR_CONFIG


def test_euler_pose_matches_scipy(host):
    """R = Rz(roll) Ry(yaw) Rx(pitch) (tsdf_surface_reconstruction.cpp:66-79),
    angles formed in float32 as the reference forms them."""
    rng = np.random.default_rng(4)
    half_c = np.float32(np.pi / 180.0)
    for _ in range(50):
        r, y, p = rng.uniform(-180, 180, 3)
        t = rng.normal(size=3)
        out = (C.c_double * 7)()
        host.recon_euler_pose(r, y, p, *t, out)
        q = np.array(out[:4])
        ang = [np.float64(np.float32(a) * half_c) for a in (r, y, p)]
        R = Rotation.from_euler("ZYX", ang).as_matrix()
        Rq = Rotation.from_quat([q[1], q[2], q[3], q[0]]).as_matrix()
        assert abs(np.linalg.norm(q) - 1) < 1e-14
        np.testing.assert_allclose(Rq, R, atol=1e-14)
        np.testing.assert_array_equal(np.array(out[4:]), np.float32(t).astype(np.float64))


def _run_driver(*args):
    import subprocess
    exe = LIB.parent / "tsdf_surface"
    return subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=60)


def test_driver_error_paths(tmp_path):
    """Failures before any device call: bad arguments, a missing parameter
    file, a dataset directory without an RGB-D pair."""
    assert _run_driver("--bogus").returncode != 0
    r = _run_driver(f"--config_path={tmp_path / 'missing.yaml'}")
    assert r.returncode != 0 and "missing.yaml" in r.stderr
    (tmp_path / "scene").mkdir()
    (tmp_path / "scene" / "notes.txt").write_text("not an image")
    cfg = tmp_path / "p.yaml"
    cfg.write_text(f'%YAML:1.0\nrgbd.dataset.path: "{tmp_path / "scene"}"\ndepth_scale: 0.001\n')
    r = _run_driver("--config_path", str(cfg))
    assert r.returncode != 0 and "traversed" in r.stderr

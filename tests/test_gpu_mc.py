"""Parity of the HIP marching cubes (librecon.so) with the CPU oracle
(marching_cubes.cu:23-236 with the gx*gy layer stride): class histogram,
triangle count, triangle index lists and vertex bits identical."""
import numpy as np
import pytest

from helpers import assert_bitexact, oracle_run

import oracle_ffi as O
import recon
from recon import scenes

pytestmark = pytest.mark.gpu


def _mc_compare(torch, field, grid, vs, off):
    dev = torch.from_numpy(np.ascontiguousarray(field, np.float32)).cuda()
    n_o, hist_o = O.mc_count(field, grid, hist=True)
    hist_g = recon.mc_class_histogram(dev, grid).cpu().numpy()
    assert np.array_equal(hist_g, hist_o)
    assert recon.mc_count(dev, grid, vs, off) == n_o
    verts, tris = recon.mc_extract(dev, grid, vs, off)
    ov, ot, _, _ = O.mc_extract(field, grid, vs, off)
    assert verts.shape[0] == n_o
    assert_bitexact(verts.cpu().numpy(), ov, "vertices")
    assert np.array_equal(tris.cpu().numpy(), ot)
    return n_o


def test_single_cube_cases(cuda):
    # all 256 classes on a 2x2x2 grid, where the reference's layer stride
    # gx+gy equals gx*gy (marching_cubes.cu:211)
    corner = [(0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), (0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)]
    for c in range(256):
        f = np.empty((2, 2, 2), np.float32)
        for i, (dx, dy, dz) in enumerate(corner):
            f[dz, dy, dx] = -0.75 if (c >> i) & 1 else 0.25 + 0.1 * i
        _mc_compare(cuda, f, (2, 2, 2), (0.1, 0.2, 0.3), (-1.0, 0.5, 2.0))


def test_integrated_volume(cuda):
    w = scenes.plain_small()
    cfg = w.config()
    o = oracle_run(cfg, w.render(), w.poses)
    vol = o["vol"]
    n = _mc_compare(cuda, o["D"], w.grid, list(vol.voxel_size), list(vol.offset))
    assert n > 50000


# rows of 3, 70, 65 and 129 voxels take the count pass's 4-byte loads; rows
# of a multiple of 4 its 16-byte loads, 256 cubes a span: one span cut short
# (4, 64, 68), two (260) and three (516)
@pytest.mark.parametrize("grid", [(3, 2, 2), (70, 3, 5), (65, 66, 7), (129, 2, 3), (4, 3, 3),
                                  (64, 5, 9), (68, 7, 6), (260, 9, 4), (516, 6, 5)])
def test_random_fields(cuda, grid):
    rng = np.random.default_rng(sum(grid))
    gx, gy, gz = grid
    f = rng.standard_normal((gz, gy, gx)).astype(np.float32)
    f[rng.random(f.shape) < 0.1] = 0.0           # exact zeros: not < 0
    f[rng.random(f.shape) < 0.05] = -0.0         # negative zero: not < 0
    f[rng.random(f.shape) < 0.05] = 1e-41        # denormals
    _mc_compare(cuda, f, grid, (0.01, 0.02, 0.03), (0.1, -0.2, 0.3))


def test_unaligned_field_takes_4_byte_loads(cuda):
    # a field one float past a 16-byte boundary: rows of 68 voxels, but the
    # 4-byte loads; the same results as the aligned copy
    grid = (68, 7, 6)
    rng = np.random.default_rng(7)
    f = rng.standard_normal(grid[::-1]).astype(np.float32)
    buf = cuda.zeros(f.size + 1, dtype=cuda.float32, device="cuda")
    buf[1:] = cuda.from_numpy(f.ravel()).cuda()
    dev = buf[1:]
    assert dev.data_ptr() % 16 == 4
    vs, off = (0.01, 0.02, 0.03), (0.1, -0.2, 0.3)
    ov, ot, _, _ = O.mc_extract(f, grid, vs, off)
    verts, tris = recon.mc_extract(dev, grid, vs, off)
    assert verts.shape[0] == len(ot) > 0
    assert_bitexact(verts.cpu().numpy(), ov, "vertices")
    assert np.array_equal(tris.cpu().numpy(), ot)


def test_many_tiles(cuda):
    # 1024 x 1024 cube rows of 7 cubes: 65536 tiles of 16 rows, so the tile
    # scan makes eight passes and group records carry tile indices past
    # 2^16; a cylinder along x crosses a quarter of the rows, and noise in
    # the last 64 layers puts triangles in most chunks there
    gx, gy, gz = 8, 1025, 1025
    y = np.arange(gy, dtype=np.float32)[None, :, None]
    z = np.arange(gz, dtype=np.float32)[:, None, None]
    f = np.sqrt((y - 500.0) ** 2 + (z - 480.0) ** 2) - np.float32(300.0) + np.zeros(gx, np.float32)
    f = np.ascontiguousarray(f.astype(np.float32))
    rng = np.random.default_rng(11)
    f[-64:] = rng.standard_normal((64, gy, gx)).astype(np.float32)
    n = _mc_compare(cuda, f, (gx, gy, gz), (0.01, 0.01, 0.01), (0.0, 0.0, 0.0))
    assert n > 100000


def test_sparse_chunks_of_wide_rows(cuda):
    # rows of 2199 cubes: 35 chunks of 64, so chunks 31-34 share the last bit
    # of the emit pass's chunk mask; 5 cube rows per layer, so each layer has
    # one group of 4 rows (one mask for the group) and one row of its own.
    # Surface voxels sit in a few chunks only, at chunk boundaries and at the
    # row ends, in some rows and not in others.
    gx, gy, gz = 2200, 6, 4
    f = np.full((gz, gy, gx), 0.5, np.float32)
    xs = [0, 5, 7 * 64 + 10, 30 * 64 + 63, 31 * 64, 31 * 64 + 3, 33 * 64 + 40, gx - 1]
    for k, x in enumerate(xs):
        z, y = k % gz, (2 * k) % gy
        f[z, y, x] = -0.25 - 0.01 * k
    f[2, 5, 32 * 64 + 1] = -1.0  # the layer's row outside the 4-row group
    n = _mc_compare(cuda, f, (gx, gy, gz), (0.01, 0.02, 0.03), (0.1, -0.2, 0.3))
    assert n > 0


def test_volume_api_and_capacity(cuda):
    w = scenes.tiny(grid=(40, 36, 28), voxel_size=0.025, image=(64, 48))
    cfg = w.config()
    depth = w.render()
    v = recon.TSDFVolume(cfg)
    dev = cuda.from_numpy(depth).cuda()
    for f in range(w.num_frames):
        v.integrate(dev[f], w.poses[f])
    n = v.count_triangles()
    assert n > 0
    verts, tris = v.extract_mesh()
    assert verts.shape[0] == n
    small = cuda.empty((n // 2, 3, 3), dtype=cuda.float32, device="cuda")
    with pytest.raises(recon.ReconError) as e:
        v.extract_mesh(verts=small, tris=None)
    assert e.value.status == recon.ffi.ERR_CAPACITY
    # the first n//2 triangles are still written, in order
    assert_bitexact(small.cpu().numpy(), verts[: n // 2].cpu().numpy(), "partial")
    v.close()


def test_degenerate_grids(cuda):
    for grid in [(1, 5, 5), (5, 1, 5), (5, 5, 1), (1, 1, 1)]:
        f = -np.ones(grid[::-1], np.float32)
        dev = cuda.from_numpy(f).cuda()
        assert recon.mc_count(dev, grid, (1, 1, 1), (0, 0, 0)) == 0

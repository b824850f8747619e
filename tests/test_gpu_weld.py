"""Parity of the HIP welded marching cubes (recon_mc_extract_welded,
recon_tsdf_extract_welded_mesh) with the oracle: vertex and normal bits and
triangle indices identical; every welded triangle corner equals the soup's
vertex; capacities and degenerate grids behave as documented."""
import ctypes as C

import numpy as np
import pytest

import oracle_ffi as O
import recon
from helpers import assert_bitexact, oracle_run
from recon import ffi, scenes

pytestmark = pytest.mark.gpu


def _compare(torch, field, grid, vs, off, what):
    dev = torch.from_numpy(np.ascontiguousarray(field, np.float32)).cuda()
    v, n, t = recon.mc_extract_welded(dev, grid, vs, off)
    ov, on, ot = O.mc_weld(field, grid, vs, off)
    assert v.shape[0] == ov.shape[0] and t.shape[0] == ot.shape[0], what
    assert_bitexact(v.cpu().numpy(), ov, f"{what} vertices")
    assert_bitexact(n.cpu().numpy(), on, f"{what} normals")
    assert np.array_equal(t.cpu().numpy(), ot), f"{what} triangles"
    return dev, v, t


def test_integrated_volume_and_handle(cuda):
    w = scenes.plain_small()
    o = oracle_run(w.config(), w.render(), w.poses)
    vol = o["vol"]
    print("welded: integrated 128^3", flush=True)
    dev, v, t = _compare(cuda, o["D"], w.grid, list(vol.voxel_size), list(vol.offset), "integrated")
    assert t.shape[0] > 50000 and v.shape[0] < 0.6 * t.shape[0]
    # corners equal the soup's vertices, up to the rounding of s + (e - s)
    # where an edge end is exactly 0
    sv, _ = recon.mc_extract(dev, w.grid, list(vol.voxel_size), list(vol.offset))
    d = (v[t.long()] - sv).abs()
    assert d.max().item() <= 2e-7 and (d == 0).float().mean().item() > 0.999
    # the handle entry on a volume integrated by the device
    tv = recon.TSDFVolume(w.config())
    for f, d in enumerate(w.render()):
        tv.integrate(cuda.from_numpy(d).cuda(), w.poses[f])
    hv, hn, ht = tv.extract_welded_mesh()
    cuda.cuda.synchronize()
    assert cuda.equal(hv, v) and cuda.equal(ht, t)
    # one call into preallocated, larger buffers
    nv, nt = tv.count_welded()
    out = (cuda.full((nv + 5, 3), 9.0, device="cuda"), cuda.full((nv + 5, 3), 9.0, device="cuda"),
           cuda.full((nt + 7, 3), -5, dtype=cuda.int32, device="cuda"))
    ov, on, ot = tv.extract_welded_mesh(out=out)
    cuda.cuda.synchronize()
    assert cuda.equal(ov, v) and cuda.equal(on, hn) and cuda.equal(ot, t)
    assert (out[0][nv:] == 9.0).all() and (out[2][nt:] == -5).all()
    tv.close()


@pytest.mark.parametrize("grid", [(3, 2, 2), (70, 3, 5), (65, 66, 7), (129, 2, 3), (64, 64, 4)])
def test_random_fields(cuda, grid):
    rng = np.random.default_rng(sum(grid) + 1)
    gx, gy, gz = grid
    f = rng.standard_normal((gz, gy, gx)).astype(np.float32)
    f[rng.random(f.shape) < 0.1] = 0.0
    f[rng.random(f.shape) < 0.05] = -0.0
    f[rng.random(f.shape) < 0.05] = 1e-41
    print(f"welded random {grid}", flush=True)
    _compare(cuda, f, grid, (0.01, 0.02, 0.03), (0.1, -0.2, 0.3), f"random {grid}")


def test_counts_capacity_and_degenerate(cuda):
    lib = ffi.load()
    f = np.random.default_rng(3).standard_normal((6, 7, 8)).astype(np.float32)
    dev = cuda.from_numpy(f).cuda()
    g, vs, off = ffi.i32x3((8, 7, 6)), ffi.f32x3((1, 1, 1)), ffi.f32x3((0, 0, 0))
    nv, nt = C.c_int64(), C.c_int64()
    assert lib.recon_mc_extract_welded(dev.data_ptr(), g, vs, off, None, None, 0, None, 0,
                                       C.byref(nv), C.byref(nt), None) == 0
    ov, _, ot = O.mc_weld(f, (8, 7, 6), (1, 1, 1), (0, 0, 0))
    assert (nv.value, nt.value) == (ov.shape[0], ot.shape[0])
    buf = cuda.full((nv.value, 3), -7.0, dtype=cuda.float32, device="cuda")
    tri = cuda.full((nt.value, 3), -7, dtype=cuda.int32, device="cuda")
    st = lib.recon_mc_extract_welded(dev.data_ptr(), g, vs, off, buf.data_ptr(), None,
                                     nv.value - 1, None, 0, C.byref(nv), C.byref(nt), None)
    assert st == ffi.ERR_CAPACITY
    st = lib.recon_mc_extract_welded(dev.data_ptr(), g, vs, off, None, None, 0, tri.data_ptr(),
                                     nt.value - 1, C.byref(nv), C.byref(nt), None)
    assert st == ffi.ERR_CAPACITY
    # one output too small: the other, large enough, is not written either
    st = lib.recon_mc_extract_welded(dev.data_ptr(), g, vs, off, buf.data_ptr(), None, nv.value,
                                     tri.data_ptr(), nt.value - 1, C.byref(nv), C.byref(nt), None)
    assert st == ffi.ERR_CAPACITY
    assert (nv.value, nt.value) == (ov.shape[0], ot.shape[0])
    cuda.cuda.synchronize()
    assert (buf.cpu().numpy() == -7.0).all() and (tri.cpu().numpy() == -7).all()
    # triangles only
    assert lib.recon_mc_extract_welded(dev.data_ptr(), g, vs, off, None, None, 0, tri.data_ptr(),
                                       nt.value, C.byref(nv), C.byref(nt), None) == 0
    cuda.cuda.synchronize()
    assert np.array_equal(tri.cpu().numpy(), ot)
    for grid in [(1, 5, 5), (5, 1, 5), (5, 5, 1)]:
        d = cuda.ones(grid[::-1], dtype=cuda.float32, device="cuda")
        d.view(-1)[::2] = -1.0
        v, n, t = recon.mc_extract_welded(d, grid, (1, 1, 1), (0, 0, 0))
        assert v.shape[0] == 0 and t.shape[0] == 0

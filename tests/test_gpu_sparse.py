"""Parity of the HIP sparse (brick) volume (recon_sparse_*) with its oracle
and with the dense path: the same set of bricks and bit-identical brick
contents after several frames (also at negative brick coordinates); with
every brick pre-allocated the dense integrate's result; dense export and
marching cubes of the export identical to the oracle's; a full pool is
flagged and stays safe; reset empties the volume."""
import numpy as np
import pytest

import oracle_ffi as O
import recon
from helpers import assert_bitexact
from recon import scenes

pytestmark = pytest.mark.gpu


def _scene(frames=3, offset=None):
    w = scenes.plain_small(num_frames=frames, grid=64, voxel_size=0.016)
    if offset is not None:
        w.offset = offset
    return w, w.render()


def _compare_bricks(gpu, oracle, what):
    g, o = gpu.bricks(), oracle.bricks()
    assert set(g) == set(o), f"{what}: {len(g)} vs {len(o)} bricks"
    for k, (od, ow) in o.items():
        gd, gw = g[k]
        assert np.array_equal(gd.view(np.uint32), od.view(np.uint32)), f"{what} D of brick {k}"
        assert np.array_equal(gw.view(np.uint32), ow.view(np.uint32)), f"{what} W of brick {k}"
    return len(o)


@pytest.mark.parametrize("offset", [None, (0.0, 0.0, 0.0)])
def test_sparse_equals_oracle(cuda, offset):
    w, depth = _scene(3, offset)
    cfg = w.config()
    sv = recon.SparseVolume(cfg, 1 << 16)
    so = O.Sparse(O.volume_from_config(cfg))
    print(f"sparse 3 frames offset {offset}", flush=True)
    for f in range(3):
        sv.integrate(cuda.from_numpy(depth[f]).cuda(), w.poses[f])
        so.integrate(depth[f], w.poses[f])
    cuda.cuda.synchronize()
    n, flags = sv.stats()
    assert flags == 0
    m = _compare_bricks(sv, so, f"offset {offset}")
    assert m == n > 100
    if offset is not None:   # bricks on both sides of the origin
        assert any(min(k) < 0 for k in sv.bricks())
    sv.reset()
    assert sv.stats() == (0, 0)
    sv.close()


def test_full_allocation_equals_dense_and_export_meshes(cuda):
    w, depth = _scene(3)
    cfg = w.config()
    sv = recon.SparseVolume(cfg, 1 << 16)
    sv.allocate_box((0, 0, 0), w.grid)
    dv = recon.TSDFVolume(cfg)
    for f in range(3):
        d = cuda.from_numpy(depth[f]).cuda()
        sv.integrate(d, w.poses[f])
        dv.integrate(d, w.poses[f])
    cuda.cuda.synchronize()
    D, W = dv.download()
    Ds, Ws = sv.export((0, 0, 0), w.grid)
    cuda.cuda.synchronize()
    assert_bitexact(Ds.cpu().numpy(), D, "D")
    assert_bitexact(Ws.cpu().numpy(), W, "W")
    dv.close()
    # export of a box reaching past the allocated bricks, then marching cubes
    sv2 = recon.SparseVolume(cfg, 1 << 16)
    so = O.Sparse(O.volume_from_config(cfg))
    for f in range(3):
        sv2.integrate(cuda.from_numpy(depth[f]).cuda(), w.poses[f])
        so.integrate(depth[f], w.poses[f])
    origin, size = (-12, -5, -9), (90, 77, 83)
    Dg, Wg = sv2.export(origin, size)
    cuda.cuda.synchronize()
    Do, Wo = so.export(origin, size)
    assert_bitexact(Dg.cpu().numpy(), Do, "export D")
    assert_bitexact(Wg.cpu().numpy(), Wo, "export W")
    vo = O.volume_from_config(cfg)
    off = [float(vo.offset[k] + np.float32(origin[k]) * vo.voxel_size[k]) for k in range(3)]
    verts, _ = recon.mc_extract(Dg, size, list(vo.voxel_size), off)
    ov, _, _, _ = O.mc_extract(Do, size, list(vo.voxel_size), off)
    assert verts.shape[0] > 10000
    assert_bitexact(verts.cpu().numpy(), ov, "mesh")
    sv.close()
    sv2.close()


def test_allocation_bounds(cuda):
    w, depth = _scene(2)
    cfg = w.config()
    lo, hi = (5, -3, 10), (50, 40, 70)
    sv = recon.SparseVolume(cfg, 1 << 16)
    sv.set_bounds(lo, hi)
    so = O.Sparse(O.volume_from_config(cfg))
    so.set_bounds(lo, hi)
    for f in range(2):
        sv.integrate(cuda.from_numpy(depth[f]).cuda(), w.poses[f])
        so.integrate(depth[f], w.poses[f])
    cuda.cuda.synchronize()
    n = _compare_bricks(sv, so, "bounded")
    assert n > 20
    for k in sv.bricks():
        for a in range(3):
            assert lo[a] // 8 <= k[a] <= (hi[a] - 1) // 8
    sv.close()


def test_full_pool_is_flagged(cuda):
    w, depth = _scene(1)
    sv = recon.SparseVolume(w.config(), 16)
    sv.integrate(cuda.from_numpy(depth[0]).cuda(), w.poses[0])
    cuda.cuda.synchronize()
    n, flags = sv.stats()
    assert n == 16 and flags & 1
    D, W = sv.export((0, 0, 0), (32, 32, 32))
    cuda.cuda.synchronize()
    assert np.isfinite(D.cpu().numpy()).all()
    sv.close()

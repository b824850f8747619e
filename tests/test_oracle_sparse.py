"""Pins the sparse (brick) volume oracle (sparse_oracle.c; no reference
counterpart on the dense path) against the pinned dense oracle: with every
brick allocated it equals the dense integrate bit for bit over several
frames; from empty, one frame allocates the brick of every voxel the dense
integrate updates inside the truncation band, and every allocated voxel then
equals the dense one; allocation is sparse."""
import numpy as np

import oracle_ffi as O
from recon import scenes


def _scene(frames):
    w = scenes.plain_small(num_frames=frames, grid=64, voxel_size=0.016)
    return w, O.volume_from_config(w.config()), w.render()


def test_full_allocation_equals_dense():
    w, vol, depth = _scene(3)
    s = O.Sparse(vol)
    s.allocate_box((0, 0, 0), w.grid)
    assert s.num_bricks() == 8 * 8 * 8
    D, W = O.init(vol)
    for f in range(3):
        O.integrate_f32(vol, D, W, depth[f], w.poses[f])
        s.integrate(depth[f], w.poses[f])
    assert s.num_bricks() > 512         # the band also covers surfaces beyond the grid
    Ds, Ws = s.export((0, 0, 0), w.grid)
    assert np.array_equal(Ds.view(np.uint32), D.view(np.uint32))
    assert np.array_equal(Ws.view(np.uint32), W.view(np.uint32))


def test_band_voxels_are_allocated_and_equal_dense():
    w, vol, depth = _scene(1)
    s = O.Sparse(vol)
    s.integrate(depth[0], w.poses[0])
    D, W = O.init(vol)
    O.integrate_f32(vol, D, W, depth[0], w.poses[0])
    bricks = s.bricks()
    gz, gy, gx = D.shape
    alloc = np.zeros((gz // 8, gy // 8, gx // 8), bool)
    for (bx, by, bz) in bricks:
        if 0 <= bx < gx // 8 and 0 <= by < gy // 8 and 0 <= bz < gz // 8:
            alloc[bz, by, bx] = True
    assert 0 < alloc.sum() < 0.5 * alloc.size      # sparse inside the grid
    amask = np.repeat(np.repeat(np.repeat(alloc, 8, 0), 8, 1), 8, 2)
    band = (W > 0) & (D < vol.truncation)
    assert band.sum() > 10000 and amask[band].all()
    Ds, Ws = s.export((0, 0, 0), w.grid)
    assert np.array_equal(Ds[amask].view(np.uint32), D[amask].view(np.uint32))
    assert np.array_equal(Ws[amask].view(np.uint32), W[amask].view(np.uint32))
    # unallocated voxels read as the fill value
    assert (Ds[~amask] == np.float32(vol.truncation)).all() and (Ws[~amask] == 0).all()
    # a brick's own arrays agree with the export
    (bx, by, bz), (bd, bw) = next(iter(bricks.items()))
    ed, ew = s.export((8 * bx, 8 * by, 8 * bz), (8, 8, 8))
    assert np.array_equal(ed, bd) and np.array_equal(ew, bw)

"""Parity of the 512^3 x 300-frame workload at full size: 4 mm voxels, 40
boxes on a floor seen from a seeded spline orbit, marching cubes after every
50 frames.  It is the one workload where the weight cap (128) binds, so the
capped running average (W' = min(W + 1, cap), D' = (D*W + tsdf)/(W + 1); the
reference's update of tsdf_volume.cu:179-185 plus the cap of octree.cu:125-127)
is exercised past the cap.

Three GPU volumes follow the same frames: one through recon_tsdf_integrate as
a caller makes it (no counter), one with a per-frame update counter, one
through recon_tsdf_integrate_frames, 50 frames per call.  At each of the 6
extraction points:
  * D and W of the plain volume equal the float32 oracle's bit for bit, and
    the counted and fused volumes equal the plain one;
  * every frame's update count equals the oracle's;
  * the cube-class histogram, triangle count, vertex bits and index lists of
    marching cubes equal the oracle's (marching_cubes.cu:102-236);
and after the last frame D is within 1e-6 of the float64 twin, W equals it,
and W reaches the cap.

The tests share one run and each brings it to its own frame, so any one of
them also runs alone.
"""
import numpy as np
import pytest

from helpers import assert_bitexact

import oracle_ffi as O
import recon
from recon import scenes

pytestmark = pytest.mark.gpu

CHECKPOINTS = [50, 100, 150, 200, 250, 300]
F64_TOL = 1e-6


class _Run:
    def __init__(self, torch):
        self.torch = torch
        self.w = w = scenes.plain_large()
        assert w.grid == (512, 512, 512) and w.num_frames == 300 and w.mesh_every == 50
        self.cfg = w.config()
        print("plain_large: rendering 300 frames", flush=True)
        self.depth = w.render()
        self.dev = torch.from_numpy(self.depth).cuda()
        self.vol = recon.TSDFVolume(self.cfg)
        self.counted = recon.TSDFVolume(self.cfg)
        self.fused = recon.TSDFVolume(self.cfg)
        self.counters = torch.zeros(w.num_frames, dtype=torch.int64, device="cuda")
        self.ov = O.volume_from_config(self.cfg)
        self.D, self.W = O.init(self.ov)
        self.D64, self.W64 = self.D.astype(np.float64), self.W.astype(np.float64)
        self.counts = []
        self.done = 0

    def advance(self, upto):
        """Integrates frames [done, upto) everywhere; the GPU work is queued
        first so that it runs while the oracle computes on the host."""
        w, lo = self.w, self.done
        if upto <= lo:
            return
        print(f"plain_large: frames {lo}..{upto - 1}", flush=True)
        for f in range(lo, upto):
            self.vol.integrate(self.dev[f], w.poses[f])
            self.counted.set_update_counter(self.counters[f:f + 1])
            self.counted.integrate(self.dev[f], w.poses[f])
        self.counted.set_update_counter(None)
        self.fused.integrate_frames(self.dev[lo:upto], w.poses[lo:upto])
        for f in range(lo, upto):
            n = O.integrate_f32(self.ov, self.D, self.W, self.depth[f], w.poses[f])
            n64, _ = O.integrate_f64(self.ov, self.D64, self.W64, self.depth[f], w.poses[f])
            assert n64 == n, (f, n64, n)
            self.counts.append(int(n))
        self.torch.cuda.synchronize()
        self.done = upto


@pytest.fixture(scope="module")
def run(cuda):
    r = _Run(cuda)
    yield r
    for v in (r.vol, r.counted, r.fused):
        v.close()


@pytest.mark.parametrize("frame", CHECKPOINTS)
def test_plain_large_checkpoint(run, frame):
    torch = run.torch
    run.advance(frame)
    print(f"plain_large: checking volumes after frame {frame}", flush=True)
    D, W = run.vol.download()
    assert_bitexact(W, run.W, f"W after {frame} frames")
    assert_bitexact(D, run.D, f"D after {frame} frames")
    got = run.counters[:frame].cpu().numpy().tolist()
    assert got == run.counts[:frame], f"update counts differ after {frame} frames"
    for name, v in (("counted", run.counted), ("fused", run.fused)):
        Dv, Wv = v.download()
        assert_bitexact(Wv, W, f"{name} W after {frame} frames")
        assert_bitexact(Dv, D, f"{name} D after {frame} frames")
        del Dv, Wv

    print(f"plain_large: marching cubes after frame {frame}", flush=True)
    grid = run.w.grid
    field = torch.from_numpy(D).cuda()
    hist = recon.mc_class_histogram(field, grid).cpu().numpy()
    del field
    n_o, hist_o = O.mc_count(run.D, grid, hist=True)
    assert np.array_equal(hist, hist_o), "cube-class histograms differ"
    verts, tris = run.vol.extract_mesh()
    ov, ot, _, _ = O.mc_extract(run.D, grid, list(run.ov.voxel_size), list(run.ov.offset))
    assert verts.shape[0] == n_o == ov.shape[0] > 100000, (verts.shape, n_o, ov.shape)
    assert_bitexact(verts.cpu().numpy(), ov, f"soup vertices after {frame} frames")
    assert np.array_equal(tris.cpu().numpy(), ot), "triangle index lists differ"
    del verts, tris, ov, ot

    if frame == run.w.num_frames:
        err = np.abs(D.astype(np.float64) - run.D64).max()
        assert err <= F64_TOL, err
        assert np.array_equal(W.astype(np.float64), run.W64)
        # 300 frames against a cap of 128: the cap binds
        assert W.max() == run.w.max_weight == 128.0
        assert int((W == 128.0).sum()) > 1000000

"""The batched path (recon_batch_*): one graph launch integrates and meshes
many uneven volumes; every replay must equal, bit for bit, looping over the
volumes with the single-volume path (and, for small volumes, the oracle)."""
import json

import numpy as np
import pytest

from helpers import assert_bitexact, oracle_run

import oracle_ffi as O
import recon
from recon import scenes

pytestmark = pytest.mark.gpu


def _shrink(w, grid=None, frames=None):
    """A smaller copy of a batched workload (same scene and cameras)."""
    if grid is not None:
        ext = w.voxel_size * w.grid[0]
        w.grid = tuple(grid)
        w.voxel_size = ext / grid[0]
        w.truncation = 4 * w.voxel_size
    if frames is not None:
        w.poses = w.poses[:frames]
    return w


def _looped(torch, w, depth_dev):
    v = recon.TSDFVolume(w.config())
    for f in range(w.num_frames):
        v.integrate(depth_dev[f], w.poses[f])
    torch.cuda.synchronize()
    D, W = v.download()
    verts, _ = v.extract_mesh()
    out = (D, W, verts.cpu().numpy())
    v.close()
    return out


def _batch_set(torch, n=7, seed=0x5EED0002):
    ws = scenes.batched_uneven(num_volumes=n, seed=seed)
    ws = [_shrink(w, grid=None, frames=min(w.num_frames, 12)) for w in ws]
    depths = [torch.from_numpy(w.render()).cuda() for w in ws]
    return ws, depths


def _check_batch(torch, ws, depths, replays=3, tune_cache=None):
    eager = [_looped(torch, w, d) for w, d in zip(ws, depths)]
    cap = sum(e[2].shape[0] for e in eager) + 16
    b = recon.Batch([(w.config(), d, w.poses) for w, d in zip(ws, depths)], mesh_capacity=cap,
                    tune_cache=tune_cache)
    for r in range(replays):
        b.run()
        torch.cuda.synchronize()
        ranges = b.mesh_ranges()
        for i, (D0, W0, V0) in enumerate(eager):
            D, W = b.download(i)
            assert_bitexact(W, W0, f"replay {r} volume {i} W")
            assert_bitexact(D, D0, f"replay {r} volume {i} D")
            assert ranges[i, 1] == V0.shape[0]
            assert_bitexact(b.download_mesh(i), V0, f"replay {r} volume {i} mesh")
        # eager work and a synchronisation between replays
        _looped(torch, ws[0], depths[0])
    b.close()
    return eager


def test_batch_equals_looped_and_oracle(cuda):
    ws, depths = _batch_set(cuda)
    assert len({w.grid for w in ws}) > 1 and len({w.image for w in ws}) > 1
    eager = _check_batch(cuda, ws, depths)
    # the two smallest volumes also against the oracle
    order = np.argsort([w.num_voxels * w.num_frames for w in ws])[:2]
    for i in order:
        o = oracle_run(ws[i].config(), depths[i].cpu().numpy(), ws[i].poses)
        assert_bitexact(eager[i][1], o["W"], f"volume {i} W vs oracle")
        assert_bitexact(eager[i][0], o["D"], f"volume {i} D vs oracle")


def test_batch_ragged_and_empty_volumes(cuda):
    ws = []
    for k, grid in enumerate([(37, 29, 23), (1, 1, 1), (66, 8, 9), (40, 36, 28)]):
        w = scenes.tiny(grid=grid, voxel_size=0.6 / max(grid), image=(40 + 8 * k, 30 + 4 * k),
                        num_frames=3 + k)
        ws.append(w)
    depths = [cuda.from_numpy(w.render()).cuda() for w in ws]
    # a volume without frames: reset only, empty mesh
    empty = scenes.tiny(grid=(16, 16, 16), num_frames=1)
    empty.poses = empty.poses[:0]
    ws.append(empty)
    depths.append(cuda.zeros((0, 24, 32), dtype=cuda.float32, device="cuda"))
    eager = []
    for w, d in zip(ws, depths):
        eager.append(_looped(cuda, w, d))
    cap = sum(e[2].shape[0] for e in eager) + 16
    b = recon.Batch([(w.config(), d, w.poses) for w, d in zip(ws, depths)], mesh_capacity=cap)
    b.run()
    cuda.cuda.synchronize()
    for i, (D0, W0, V0) in enumerate(eager):
        D, W = b.download(i)
        assert_bitexact(W, W0, f"volume {i} W")
        assert_bitexact(D, D0, f"volume {i} D")
        assert_bitexact(b.download_mesh(i), V0, f"volume {i} mesh")
    assert (b.download(len(ws) - 1)[1] == 0).all()
    b.close()


def test_autotune_cache_round_trip(cuda, tmp_path):
    ws, depths = _batch_set(cuda, n=4, seed=77)
    cache = tmp_path / "tune.json"
    recon.Batch.autotune([(w.config(), d, w.poses) for w, d in zip(ws, depths)], cache, iterations=2)
    table = json.loads(cache.read_text())
    keys = {f"g{w.grid[0]}x{w.grid[1]}x{w.grid[2]}_{w.image[0]}x{w.image[1]}" for w in ws}
    assert keys <= set(table)
    for v in table.values():
        assert 0 <= v["variant"] < 8 and len(v["batch_us"]) == 8
    # the tuned batch gives the same results
    _check_batch(cuda, ws, depths, replays=1, tune_cache=str(cache))


def test_mesh_capacity_overflow(cuda):
    ws, depths = _batch_set(cuda, n=3, seed=5)
    b = recon.Batch([(w.config(), d, w.poses) for w, d in zip(ws, depths)], mesh_capacity=10)
    b.run()
    with pytest.raises(recon.ReconError) as e:
        b.mesh_ranges()
    assert e.value.status == recon.ffi.ERR_CAPACITY
    b.close()


def test_batch_full_size_64_volumes(cuda):
    """The full 64-volume workload (grids 64-256^3, 1-40 frames, three image
    sizes; about 2.6 GiB of D and W, so arena offsets pass 2^31 bytes): two
    replays equal the per-volume loop for every volume, and one volume of
    each grid size equals the oracle."""
    ws = scenes.batched_uneven()
    assert len(ws) == 64
    print("batched 64 volumes: rendering", flush=True)
    depths = [cuda.from_numpy(w.render()).cuda() for w in ws]
    total = sum(w.num_voxels for w in ws)
    assert 8 * total > 2 ** 31
    eager = _check_batch(cuda, ws, depths, replays=2)
    print("batched 64 volumes: oracle", flush=True)
    seen = {}
    for i, w in enumerate(ws):
        if w.grid not in seen or w.num_frames < ws[seen[w.grid]].num_frames:
            seen[w.grid] = i
    assert len(seen) >= 4
    for g, i in sorted(seen.items()):
        o = oracle_run(ws[i].config(), depths[i].cpu().numpy(), ws[i].poses)
        assert_bitexact(eager[i][1], o["W"], f"volume {i} ({g}) W vs oracle")
        assert_bitexact(eager[i][0], o["D"], f"volume {i} ({g}) D vs oracle")
        ov, _, _, _ = O.mc_extract(o["D"], ws[i].grid, list(o["vol"].voxel_size),
                                   list(o["vol"].offset))
        assert_bitexact(eager[i][2], ov, f"volume {i} ({g}) mesh vs oracle")
